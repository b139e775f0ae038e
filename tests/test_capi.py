"""Host-side checks of the C ABI library (no GPU needed): libtfmm.so is
built for gfx950, loads, and exports every entry point include/tfmm.h
declares."""
import os
import re
import subprocess

import pytest

import tfmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "tfmm.h")).read()
    return sorted(set(re.findall(r"\b(tfmm_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_binding_surface():
    assert declared() == sorted(tfmm.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = tfmm.lib()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", tfmm.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = set(re.findall(r"\bT (tfmm_\w+)", out))
    assert set(declared()) <= syms


def test_library_carries_gfx950_code():
    with open(tfmm.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_params_struct_layout():
    import ctypes
    assert ctypes.sizeof(tfmm.Params) == 24
    assert ctypes.sizeof(tfmm.BodyDesc) == 32


def test_tuning_knobs_declared():
    """The tuning knobs and the autotuner are part of the C surface."""
    txt = open(os.path.join(ROOT, "include", "tfmm.h")).read()
    for name in ("TFMM_TUNE_M2L_UNIT 1", "TFMM_TUNE_M2L_XCD_RUN 2", "TFMM_TUNE_P2P_SCHEDULE 3",
                 "tfmm_autotune("):
        assert name in txt
    assert (tfmm.TUNE_M2L_UNIT, tfmm.TUNE_M2L_XCD_RUN, tfmm.TUNE_P2P_SCHEDULE) == (1, 2, 3)
