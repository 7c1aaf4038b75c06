"""The C ABI library without a GPU: it loads, exports every declared entry
point, runs its host-side functions, and fails loudly (status + message, no
CPU fallback) when asked for device work with no device present."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from ultimateepm_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uepm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uepm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("uepm_ctx_create", "uepm_plan_create", "uepm_solve", "uepm_solve_device", "uepm_last_error",
                     "uepm_basis_vectors", "uepm_kpath", "uepm_bz_sample"):
        assert required in names


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (uepm_[a-z0-9_]+)", out.stdout))
    assert set(declared_functions()) <= exported
    # nothing but the declared boundary is exported under the uepm_ prefix
    assert exported <= set(declared_functions())


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["ldd", _abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "oracle" not in out and "libuepm_ref" not in out and "torch" not in out
    syms = subprocess.run(["nm", "-D", _abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in syms and "ref_" not in syms


def _no_gpu():
    try:
        h = C.c_void_p()
        code = _abi.lib().uepm_ctx_create(0, C.byref(h))
    except OSError:
        return True
    if code == 0:
        _abi.lib().uepm_ctx_destroy(h)
        return False
    return True


def test_device_work_without_a_device_fails_loudly(uepm):
    if not _no_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(_abi.UepmError) as ei:
        uepm.Context(0)
    assert ei.value.code in (_abi.UEPM_ERR_DEVICE, _abi.UEPM_ERR_INVALID_ARG)
    assert _abi.lib().uepm_last_error().decode() != ""


def test_null_arguments_are_rejected():
    L = _abi.lib()
    plan = C.c_void_p()
    G = np.zeros((1, 3), dtype=np.int32)
    m = _abi.uepm_material()
    m.lattice_constant_angstrom = 5.43
    assert L.uepm_plan_create(None, C.byref(m), G.ctypes.data_as(C.POINTER(C.c_int32)), 1, 0, 1,
                              C.byref(plan)) == _abi.UEPM_ERR_INVALID_ARG
    assert "null" in L.uepm_last_error().decode()
    k = np.zeros(3)
    assert L.uepm_solve_device_noalloc(None, k.ctypes.data, 1, k.ctypes.data, None, None,
                                       None) == _abi.UEPM_ERR_INVALID_ARG
    assert "null plan" in L.uepm_last_error().decode()
    assert L.uepm_ctx_create(0, None) == _abi.UEPM_ERR_INVALID_ARG
    assert L.uepm_solve(None, None, 0, None, None, None) == _abi.UEPM_ERR_INVALID_ARG


@pytest.mark.parametrize("n,N", [(2, 9), (3, 15), (6, 59), (10, 137), (12, 181), (14, 259)])
def test_basis_counts(uepm, n, N):
    assert len(uepm.basis_vectors(n)) == N


def test_basis_out_of_table(uepm):
    """n outside [2, 116] (the size of the shell table): GenerateBasisVectors returns false (BandStructure.cpp:34-37)."""
    assert len(uepm.basis_vectors(116)) > 0
    for n in (0, 1, 117):
        with pytest.raises(_abi.UepmError) as ei:
            uepm.basis_vectors(n)
        assert ei.value.code == _abi.UEPM_ERR_RUNTIME
    # a short buffer receives the first `capacity` vectors, *count the full size
    cnt = C.c_int(0)
    G = np.zeros((3, 3), dtype=np.int32)
    assert _abi.lib().uepm_basis_vectors(6, G.ctypes.data_as(C.POINTER(C.c_int32)), 3, C.byref(cnt)) == 0
    assert cnt.value == 59 and np.array_equal(G, uepm.basis_vectors(6)[:3])


def test_kpath_unknown_label(uepm):
    with pytest.raises(_abi.UepmError):
        uepm.kpath("GQX", 30)


def test_kpath_nodes_are_exact(uepm):
    pts, pos = uepm.kpath("LGXWKULWXKG", 2000)
    assert len(pts) == 2000
    # fcc symmetry points in units of 2 pi / a (SymmetryPoints.cpp:21-28)
    nodes = {"L": (0.5, 0.5, 0.5), "G": (0, 0, 0), "X": (1, 0, 0), "W": (1, 0.5, 0), "K": (0.75, 0.75, 0),
             "U": (1, 0.25, 0.25)}
    for lab, p in zip("LGXWKULWXKG", pos):
        assert tuple(pts[p]) == nodes[lab], lab
    assert pos[0] == 0 and pos[-1] == 1999
    assert (np.diff(pos.astype(int)) > 0).all()


def test_bz_sample_zero_and_validation(uepm):
    assert uepm.bz_sample(3, 0).shape == (0, 3)
    with pytest.raises(_abi.UepmError):
        _abi.check(_abi.lib().uepm_bz_sample(C.c_uint64(1), 5, None))
