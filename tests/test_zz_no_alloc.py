"""uepm_solve_device_noalloc, the solve entry point that allocates nothing,
against uepm_solve_device (the entry point whose kernels it shares).

The no-allocation entry point carves its chunk workspace out of the
library's fixed device scratch array; every k-point's stages are the same
kernels with the same arguments, so the results are the same bits.  Two
calls in a row, the second on fewer k-points, each against the allocating
entry point; the array keeps the first call's contents when the second runs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from ultimateepm_amd import api
from ultimateepm_amd._abi import UEPM_OK, lib
from ultimateepm_amd.material import load

pytestmark = pytest.mark.gpu


def _solve(fn, plan, ks, nb, dim, vectors, gradients, stream):
    """Device solve through one of the two entry points; host copies of the outputs."""
    ctx = plan._ctx
    dk = api.DeviceArray.from_host(ctx, ks)
    nan = np.full((len(ks), nb), np.nan)
    de = api.DeviceArray.from_host(ctx, nan)
    dv = api.DeviceArray.from_host(ctx, np.full((len(ks), nb, dim, 2), np.nan)) if vectors else None
    dg = api.DeviceArray.from_host(ctx, np.full((len(ks), nb, 3), np.nan)) if gradients else None
    rc = fn(plan.handle, C.c_void_p(dk.ptr), len(ks), C.c_void_p(de.ptr), C.c_void_p(dv.ptr if dv else None),
            C.c_void_p(dg.ptr if dg else None), C.c_void_p(stream.handle))
    assert rc == UEPM_OK, lib().uepm_last_error().decode()
    stream.synchronize()
    return de.numpy(), dv.numpy() if dv else None, dg.numpy() if dg else None


@pytest.mark.parametrize("sym,n,flags,nb", [
    ("Si", 12, api.NONLOCAL | api.MIXED | api.VECTORS, 12),  # N = 181, precision form, vectors
    ("Si", 12, api.NONLOCAL | api.VECTORS | api.GRADIENTS, 12),  # FP64 with gradients
    ("Ge", 10, api.NONLOCAL | api.SOC | api.MIXED, 16),  # dim 274: register grids from the 288-wide one
    ("Ge", 11, api.NONLOCAL | api.SOC | api.MIXED, 16),  # dim 338: blocked leading steps, then the grids
])
def test_noalloc_equals_solve_device(gpu_ctx, sym, n, flags, nb):
    print(f"case: no-allocation solve vs uepm_solve_device, {sym} n={n} flags {flags} nb {nb}", flush=True)
    plan = api.Plan(gpu_ctx, load(sym), api.basis_vectors(n), flags, nb)
    plan._ctx = gpu_ctx
    vectors = bool(flags & (api.VECTORS | api.GRADIENTS))
    gradients = bool(flags & api.GRADIENTS)
    stream = gpu_ctx.stream()
    L = lib()
    for seed, nk in ((11, 1000), (12, 300)):  # the second call is smaller
        ks = api.bz_sample(seed, nk)
        got = _solve(L.uepm_solve_device_noalloc, plan, ks, nb, plan.dim, vectors, gradients, stream)
        assert plan.status(stream.handle) == (0, -1)
        ref = _solve(L.uepm_solve_device, plan, ks, nb, plan.dim, vectors, gradients, stream)
        for a, b in zip(got, ref):
            if b is not None:
                assert np.all(np.isfinite(a))
                assert np.array_equal(a, b)
    stream.close()
    plan.close()


def test_noalloc_reports_failed_k_points(gpu_ctx):
    print("case: no-allocation solve with a NaN k-point, Si n=10 mixed", flush=True)
    plan = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(10), api.NONLOCAL | api.MIXED, 8)
    plan._ctx = gpu_ctx
    ks = api.bz_sample(5, 400)
    ks[333, 1] = np.nan
    stream = gpu_ctx.stream()
    ev, _, _ = _solve(lib().uepm_solve_device_noalloc, plan, ks, 8, plan.dim, False, False, stream)
    assert plan.status(stream.handle) == (1, 333)
    assert np.all(np.isnan(ev[333])) and np.all(np.isfinite(np.delete(ev, 333, axis=0)))
    stream.close()
    plan.close()
