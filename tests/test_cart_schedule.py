"""The Cartesian M2L's two formulations, checked on the CPU.

tests/native/cart_schedule_check.cpp builds the schedule tables of
tapas-reduce_amd/csrc/cart_sched.h and applies them as the kernel
(cartesian.hip, k_cart_m2l) does -- derivative recurrence from the tables'
indices and factorials, lane-by-lane bin accumulators over several sources,
the zero slot, the fixed-order sum of each term's pieces -- and compares
with the oracle's per-pair M2L (LaplaceCartesianCPU.cxx:659-673).  It also
checks that every product (n, k) with |n| + |k| <= P - 1 is scheduled once.
"""
import os
import shutil
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "cart_schedule_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    oracle.lib()  # builds liboracle.so if needed
    exe = str(tmp_path_factory.mktemp("cartchk") / "cart_schedule_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, SRC, oracle.LIB_PATH,
                    "-Wl,-rpath," + os.path.dirname(oracle.LIB_PATH)], check=True)
    return exe


def test_cartesian_schedule_matches_oracle(checker):
    out = subprocess.run([checker, "60"], check=True, capture_output=True, text=True).stdout
    rel = float(out.split()[1])
    assert rel < 1e-13, out


BASIS = os.path.join(HERE, "native", "cart_basis_check.cpp")


@pytest.fixture(scope="module")
def basis_checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    oracle.lib()
    exe = str(tmp_path_factory.mktemp("cartbasis") / "cart_basis_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, BASIS, oracle.LIB_PATH,
                    "-Wl,-rpath," + os.path.dirname(oracle.LIB_PATH)], check=True)
    return exe


def test_basis_change_is_exact(basis_checker):
    """T1 (Cartesian -> spherical multipoles) against the spherical P2M, T2
    (spherical -> Cartesian locals) against the spherical L2P, and the
    Cartesian M2L computed as T2 . spherical M2L . T1 against the oracle's
    Cartesian M2L: the factorisation the library's default Cartesian path
    runs (the two bases truncate the same series)."""
    out = subprocess.run([basis_checker, "30"], check=True, capture_output=True, text=True).stdout
    v = out.split()
    t1, t2, m2l = float(v[1]), float(v[3]), float(v[5])
    assert t1 < 1e-13 and t2 < 1e-13 and m2l < 1e-13, out
