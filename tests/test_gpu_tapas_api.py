"""The C++ Tapas user surface (tapas-reduce_amd/tapas_gpu/tapas_fmm.h) driven
like the reference application: Partition, Map(FMM_Upward),
Map(FMM_DTT, Product(root, root)), Map(FMM_Downward), CopyBackResult,
main.csv and the 100-target check (tapas_exafmm.cxx:448-712)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tapas-reduce_amd", "tapas_exafmm_gpu")


def run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def test_driver_rejects_mutual(tmp_path):
    """mutual = 1 is refused before any device work, as on the reference's
    CUDA path (tapas_exafmm.cxx:507-515)."""
    r = run(["-n", "1000", "-m", "1"], str(tmp_path))
    assert r.returncode == 1 and "Mutual is not supported" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("dist,prec", [("cube", "fp64"), ("plummer", "fp64"), ("cube", "mixed")])
def test_driver_runs_reference_sequence(tmp_path, dist, prec):
    print("driver", dist, prec, flush=True)
    r = run(["-n", "30000", "-c", "64", "-t", "0.4", "-m", "0", "-d", dist, "-p", prec],
            str(tmp_path))
    assert r.returncode == 0, r.stderr
    errs = [float(v) for v in re.findall(r"Rel\. L2 Error \((?:pot|acc)\)\s*:\s*(\S+)", r.stdout)]
    assert len(errs) == 2 and max(errs) < 1e-3, r.stdout
    lines = open(os.path.join(str(tmp_path), "main.csv")).read().splitlines()
    assert lines[0] == "Rank,total,upward,traverse,downward,tree,numP2P,numM2L"
    assert int(lines[1].split(",")[-1]) > 0
