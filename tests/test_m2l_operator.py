"""The batched M2L formulation of the MFMA tile kernel, checked on the CPU.

tests/native/m2l_operator_check.cpp applies the kernel's operand codes
(tapas-reduce_amd/csrc/m2l_codes.h), its table of harmonic multiples, the
level factor split between the harmonics and the rows of the A operands, and
the M~/L~ scalings
in FP64, block by block and lane by lane as the kernel does, for pairs up to
three levels apart either way (and pairs on the z axis), and compares the result
with the oracle's per-pair M2L (LaplaceSphericalCPU_tapas.h:208-273).
"""
import os
import shutil
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "m2l_operator_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    oracle.lib()  # builds liboracle.so if needed
    exe = str(tmp_path_factory.mktemp("m2lchk") / "m2l_operator_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, SRC, oracle.LIB_PATH,
                    "-Wl,-rpath," + os.path.dirname(oracle.LIB_PATH)], check=True)
    return exe


def test_block_sparse_m2l_matches_oracle(checker):
    out = subprocess.run([checker, "210"], check=True, capture_output=True, text=True).stdout
    rel = float(out.split()[1])
    # FP64 throughout: only rounding separates the two formulations
    assert rel < 1e-13, out
