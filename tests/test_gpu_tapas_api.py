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


def read_rank_csv(path):
    """A one-rank file in the reference's RankCSV format: header and row of
    14-character right-aligned columns."""
    head, row = open(path).read().splitlines()[:2]
    return dict(zip(head.split(), row.split()))


def test_driver_rejects_bad_mutual(tmp_path):
    """mutual must be 0 or 1; refused before any device work."""
    r = run(["-n", "1000", "-m", "2"], str(tmp_path))
    assert r.returncode == 1 and "mutual must be 0 or 1" in r.stderr
    r = run(["-n", "1000", "-m", "1", "-i", "1"], str(tmp_path))  # images with mutual
    assert r.returncode == 1 and "-m 0" in r.stderr
    r = run(["-n", "1000", "-b", "legendre"], str(tmp_path))
    assert r.returncode == 1 and "basis must be" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mutual", [0, 1])
def test_driver_cartesian_basis(tmp_path, mutual):
    """-b cartesian (the reference's BASIS = Cartesian build): the driver's
    check against the direct sum passes the reference's 1e-3 gate, and the
    errors equal the spherical run's to the printed digits."""
    errs = {}
    for basis in ("spherical", "cartesian"):
        print("driver basis", basis, "mutual", mutual, flush=True)
        r = run(["-n", "20000", "-c", "32", "-t", "0.4", "-m", str(mutual), "-d", "plummer",
                 "-b", basis], str(tmp_path))
        assert r.returncode == 0, r.stderr
        e = [float(v) for v in re.findall(r"Rel\. L2 Error \((?:pot|acc)\)\s*:\s*(\S+)", r.stdout)]
        assert len(e) == 2 and max(e) < 1e-3, r.stdout
        errs[basis] = e
    for a, b in zip(errs["spherical"], errs["cartesian"]):
        assert abs(a - b) <= 1e-3 * a, errs  # printed to 6 digits


@pytest.mark.gpu
@pytest.mark.parametrize("images", [1, 2])
def test_driver_periodic_images(tmp_path, images):
    """-i images (period 2 pi, the reference's cycle): the FMM against the
    direct sum over the same copies of the bodies, the reference's 1e-3 gate."""
    print("driver periodic images", images, flush=True)
    r = run(["-n", "4000", "-c", "32", "-t", "0.4", "-m", "0", "-d", "cube", "-i", str(images)],
            str(tmp_path))
    assert r.returncode == 0, r.stderr
    errs = [float(v) for v in re.findall(r"Rel\. L2 Error \((?:pot|acc)\)\s*:\s*(\S+)", r.stdout)]
    assert len(errs) == 2 and max(errs) < 1e-3, r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("dist,prec,mutual", [("cube", "fp64", 0), ("plummer", "fp64", 0),
                                              ("cube", "mixed", 0), ("plummer", "fp64", 1),
                                              ("cube", "mixed", 1)])
def test_driver_runs_reference_sequence(tmp_path, dist, prec, mutual):
    print("driver", dist, prec, "mutual", mutual, flush=True)
    r = run(["-n", "30000", "-c", "64", "-t", "0.4", "-m", str(mutual), "-d", dist, "-p", prec],
            str(tmp_path))
    assert r.returncode == 0, r.stderr
    errs = [float(v) for v in re.findall(r"Rel\. L2 Error \((?:pot|acc)\)\s*:\s*(\S+)", r.stdout)]
    assert len(errs) == 2 and max(errs) < 1e-3, r.stdout
    main = read_rank_csv(os.path.join(str(tmp_path), "main.csv"))
    assert list(main) == ["Rank", "total", "upward", "traverse", "downward", "tree", "numP2P",
                          "numM2L"]
    assert int(main["numM2L"]) > 0 and float(main["total"]) > 0
    # root->Report(): the reference's report files (TP/hot/report.h:39-152)
    lb = read_rank_csv(os.path.join(str(tmp_path), "load_balancing.csv"))
    assert int(lb["NumBodies1"]) == 30000 and int(lb["NumCells"]) > int(lb["NumLeaves"]) > 0
    tc = read_rank_csv(os.path.join(str(tmp_path), "tree_construction.csv"))
    assert list(tc) == ["Rank", "all", "sample", "exchange", "grow-local", "grow-global"]
    assert float(tc["all"]) > 0
    m2 = read_rank_csv(os.path.join(str(tmp_path), "map2.csv"))
    assert list(m2) == ["Rank", "all", "let", "net_traverse", "device_call"]
    assert float(m2["net_traverse"]) > 0
    let = read_rank_csv(os.path.join(str(tmp_path), "let_construction.csv"))
    assert len(let) == 10 and float(let["LET-All"]) == 0
    assert read_rank_csv(os.path.join(str(tmp_path), "hostnames.csv"))["Hostname"]
    misc = open(os.path.join(str(tmp_path), "misc.csv")).read().split("\n")
    assert misc[0].split() == ["SamplingRate", "NB"] and int(misc[1].split()[1]) == 30000


FUNCTOR_SRC = os.path.join(ROOT, "tests", "native", "functor_fmm.cpp")
FUNCTOR_EXE = os.path.join(ROOT, "tests", "native", "functor_fmm")


def test_reference_shaped_functors_compile_against_the_surface():
    """FMM_Upward / FMM_DTT / FMM_Downward written with the reference
    application's operator() signatures, kernel calls and Map calls
    (subcells, bodies, products of cells and of bodies; tapas_exafmm.cxx:
    87-286) compile against tapas_gpu/tapas_exafmm_gpu.h."""
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror",
                        FUNCTOR_SRC], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mutual", [0, 1])
def test_gpu_functor_traversal_matches_gpu_traversal(tmp_path, mutual):
    """The functors run on host cells with their M2L and P2P recorded and
    launched on the GPU: the same interaction lists (CSR, bit for bit) and
    fields as the library's own GPU traversal (checked in the program), and
    within 1e-10 of the CPU oracle, on the 16,384-body uniform cube, without
    and with mutual interaction."""
    import numpy as np
    import datasets
    import oracle
    x, y, z, q = datasets.cube(16384, seed=0)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as f:
        f.write(np.int64(len(x)).tobytes())
        for a in (x, y, z, q):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
    print("functor_fmm n=%d ncrit 64 theta 0.4 mutual %d" % (len(x), mutual), flush=True)
    r = subprocess.run([FUNCTOR_EXE, inp, out, "64", "0.4", str(mutual)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    fields = np.fromfile(out, np.float64).reshape(-1, 4)
    o = oracle.Run(x, y, z, q, 64, 0.4, mutual=mutual)
    c = o.counts()
    assert "m2l %d p2p leaf pairs %d" % (c["m2l"], c["p2p_leafpairs"]) in r.stdout
    ep, ea = oracle.rel_l2(fields, o.fields())
    print("functor run vs oracle: %.2e %.2e" % (ep, ea))
    assert ep < 1e-10 and ea < 1e-10


def test_report_files_use_the_reference_rankcsv_format(tmp_path):
    """RankCSV lines as the reference writes them (cpp/include/tapas/util.h:
    114-317): a Rank column first, every column right-aligned in 14
    characters, doubles in scientific notation, and one space after a value
    that fills its column (except the last)."""
    exe = os.path.join(str(tmp_path), "report_format")
    src = os.path.join(ROOT, "tests", "native", "report_format.cpp")
    r = subprocess.run(["g++", "-std=c++11", "-O1", "-o", exe, src], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60).stdout
    assert out.split("\n") == [
        "          Rank         total        numP2P      Hostname",
        "             0  1.500000e+00           123a-very-long-hostname",
        "          Rank LET-Res-acomm   grow-global",
        "             0  0.000000e+00 -2.500000e-07",
        ""]
