"""Failure reporting and workspace lifetime on the GPU.

* A non-finite k-point is refused by the host-buffer solve before any device
  work (UEPM_ERR_INVALID_ARG).
* In a device batch, k-points whose H(k) is not finite get NaN eigenvalues
  and vectors and are counted by uepm_plan_status (UEPM_ERR_NOT_CONVERGED,
  the reference's "Eigenvalue decomposition failed", Hamiltonian.cpp:130-133)
  while every other k-point is solved as it is alone -- checked against the
  oracle.  The condition is arithmetic only: NaN and infinity propagate
  through the kernels, whose loops are all bounded.
* A finite k-point that overflows H(k) takes the same path through the
  host-buffer solve.
* A batch graph holds its plans' workspaces: the plans cannot be destroyed,
  reconfigured or grown while it exists.
* Solves of one plan on two streams are ordered by the library.
"""
from __future__ import annotations

import numpy as np
import pytest

from ultimateepm_amd import api
from ultimateepm_amd._abi import UEPM_ERR_INVALID_ARG, UEPM_ERR_NOT_CONVERGED, UepmError
from ultimateepm_amd.material import load

pytestmark = pytest.mark.gpu


def _dev(ctx, a):
    return api.DeviceArray.from_host(ctx, np.ascontiguousarray(a))


def test_host_solve_refuses_non_finite_k(gpu_ctx):
    print("case: uepm_solve with NaN / inf k, Si n=6", flush=True)
    p = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(6), api.NONLOCAL, 8)
    for bad in (np.nan, np.inf, -np.inf):
        ks = np.zeros((3, 3))
        ks[1, 2] = bad
        with pytest.raises(UepmError) as e:
            p.solve(ks)
        assert e.value.code == UEPM_ERR_INVALID_ARG and "k-point 1 is not finite" in str(e.value)
    ev, _, _ = p.solve(np.zeros((1, 3)))  # the plan is still usable
    assert np.all(np.isfinite(ev))
    p.close()


@pytest.mark.parametrize("flags,nb,shells", [(api.NONLOCAL, 8, 8), (api.NONLOCAL | api.GRADIENTS, 8, 8),
                                             (api.NONLOCAL | api.MIXED, 12, 8),
                                             (api.NONLOCAL | api.MIXED | api.VECTORS, 12, 8),
                                             (api.NONLOCAL | api.SOC | api.MIXED, 16, 8),
                                             (api.NONLOCAL | api.SOC | api.GRADIENTS, 16, 10)])
def test_device_batch_with_non_finite_k(gpu_ctx, oracle, flags, nb, shells):
    """NaN and infinite k-points inside a device batch spanning two chunks (two
    lanes); at Ge + SOC n = 10 (dim 274) the FP64 form's last stage is the
    compact-WY back-transform, which reports the failed k-points."""
    sym = "Ge" if flags & api.SOC else "Si"
    print(f"case: device batch with NaN/inf k, {sym} n={shells} flags {flags} nb {nb}", flush=True)
    mat = load(sym)
    G = api.basis_vectors(shells)
    ks = api.bz_sample(31, 90)
    bad = [5, 61, 62]
    ks[5, 0] = np.nan
    ks[61, 1] = np.inf
    ks[62, 2] = -np.inf
    p = api.Plan(gpu_ctx, mat, G, flags, nb)
    p.set_config(max_chunk=48)  # two chunks: the failures land in both lanes
    dim = p.dim
    want_vec = bool(flags & (api.VECTORS | api.GRADIENTS))
    dk = _dev(gpu_ctx, ks)
    de = api.DeviceArray(gpu_ctx, (len(ks), nb))
    dv = api.DeviceArray(gpu_ctx, (len(ks), nb, dim, 2)) if want_vec else None
    dg = api.DeviceArray(gpu_ctx, (len(ks), nb, 3)) if flags & api.GRADIENTS else None
    p.solve_device(dk.ptr, len(ks), de.ptr, dv.ptr if dv else 0, dg.ptr if dg else 0)
    n_failed, first = p.status()
    assert (n_failed, first) == (len(bad), bad[0])
    with pytest.raises(UepmError) as e:
        api.check(api.lib().uepm_plan_status(p.handle, None, None, None))
    assert e.value.code == UEPM_ERR_NOT_CONVERGED
    ev = de.numpy()
    good = np.setdiff1d(np.arange(len(ks)), bad)
    assert np.all(np.isnan(ev[bad])) and np.all(np.isfinite(ev[good]))
    w, _, _ = oracle.solve(mat, G, flags & (api.NONLOCAL | api.SOC), nb, ks[good], nthreads=8)
    tol = 1e-4 if flags & api.MIXED else 1e-9
    assert np.abs(ev[good] - w).max() <= tol
    if dv is not None:
        v = dv.numpy()
        assert np.all(np.isnan(v[bad])) and np.all(np.isfinite(v[good]))
    if dg is not None:
        g = dg.numpy()
        assert np.all(np.isnan(g[bad])) and np.all(np.isfinite(g[good]))
    # the next solve clears the status
    dk2 = _dev(gpu_ctx, ks[good])
    p.solve_device(dk2.ptr, len(good), de.ptr, dv.ptr if dv else 0, dg.ptr if dg else 0)
    assert p.status() == (0, -1)
    p.close()


def test_overflowing_k_reports_not_converged(gpu_ctx, oracle):
    """A finite k-point whose kinetic term overflows: the host-buffer solve returns
    UEPM_ERR_NOT_CONVERGED with every output written, the other k-points solved."""
    print("case: uepm_solve with |k| = 1e160, Si n=6 gradients", flush=True)
    mat = load("Si")
    G = api.basis_vectors(6)
    ks = api.bz_sample(8, 6)
    ks[3] = (1e160, 0.0, 0.0)
    p = api.Plan(gpu_ctx, mat, G, api.NONLOCAL | api.VECTORS | api.GRADIENTS, 8)
    with pytest.raises(UepmError) as e:
        p.solve(ks)
    assert e.value.code == UEPM_ERR_NOT_CONVERGED and "first at index 3" in str(e.value)
    ev, vec, gr = e.value.result
    good = [0, 1, 2, 4, 5]
    assert np.all(np.isnan(ev[3]))
    w, _, _ = oracle.solve(mat, G, api.NONLOCAL, 8, ks[good], nthreads=4)
    assert np.abs(ev[good] - w).max() <= 1e-9
    assert np.all(np.isfinite(vec[good])) and np.all(np.isfinite(gr[good]))
    assert np.all(np.isnan(vec[3])) and np.all(np.isnan(gr[3]))
    p.close()


def test_batch_graph_holds_its_plans(gpu_ctx):
    """uepm_batch_* keeps raw pointers into each plan's workspace: while the
    graph exists the plans refuse destroy, reconfiguration, tuning and a solve
    that would need a larger workspace; solves within the reserved size and
    relaunches of the graph still give the same results."""
    print("case: batch graph pins Si n=6 and Ge n=6 SOC plans", flush=True)
    ks = api.bz_sample(3, 200)
    dk = _dev(gpu_ctx, ks)
    p1 = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(6), api.NONLOCAL | api.MIXED, 8)
    p2 = api.Plan(gpu_ctx, load("Ge"), api.basis_vectors(6), api.NONLOCAL | api.SOC, 8)
    e1 = api.DeviceArray(gpu_ctx, (200, 8))
    e2 = api.DeviceArray(gpu_ctx, (200, 8))
    b = api.Batch(gpu_ctx, [(p1, dk.ptr, 200, e1.ptr), (p2, dk.ptr, 200, e2.ptr)], branches=2)
    b.launch()
    gpu_ctx.synchronize()
    r1, r2 = e1.numpy(), e2.numpy()
    for call in (lambda: p1.set_config(max_chunk=64), lambda: p1.tune(dk.ptr, 200)):
        with pytest.raises(UepmError) as e:
            call()
        assert e.value.code == UEPM_ERR_INVALID_ARG
    rc = api.lib().uepm_plan_destroy(p2.handle)
    assert rc == UEPM_ERR_INVALID_ARG
    big = api.bz_sample(4, 5000)  # a larger workspace than the graph's
    with pytest.raises(UepmError) as e:
        p1.solve(big)
    assert e.value.code == UEPM_ERR_INVALID_ARG and "batch graph" in str(e.value)
    ev, _, _ = p1.solve(ks[:50])  # within the reserved workspace
    assert np.abs(ev - r1[:50]).max() <= 1e-12
    b.launch()  # the graph's buffers are the same ones
    gpu_ctx.synchronize()
    assert np.array_equal(e1.numpy(), r1) and np.array_equal(e2.numpy(), r2)
    assert p1.status() == (0, -1) and p2.status() == (0, -1)
    b.close()
    ev, _, _ = p1.solve(big)  # released: the workspace grows
    assert np.all(np.isfinite(ev))
    p1.close()
    p2.close()


def test_solves_on_two_streams_are_ordered(gpu_ctx):
    """Two solves of one plan enqueued on different streams with no
    synchronisation between them share the plan's workspace; the library
    orders them, so each equals the same solve alone."""
    print("case: one plan, two streams back to back, Si n=10 vectors", flush=True)
    p = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(10), api.NONLOCAL | api.VECTORS, 8)
    ka, kb = api.bz_sample(41, 700), api.bz_sample(42, 700)
    ref_a, vec_a, _ = p.solve(ka)
    ref_b, vec_b, _ = p.solve(kb)
    s1, s2 = gpu_ctx.stream(), gpu_ctx.stream()
    dka, dkb = _dev(gpu_ctx, ka), _dev(gpu_ctx, kb)
    ea, eb = api.DeviceArray(gpu_ctx, (700, 8)), api.DeviceArray(gpu_ctx, (700, 8))
    va, vb = api.DeviceArray(gpu_ctx, (700, 8, p.dim, 2)), api.DeviceArray(gpu_ctx, (700, 8, p.dim, 2))
    for _ in range(3):
        p.solve_device(dka.ptr, 700, ea.ptr, va.ptr, 0, s1.handle)
        p.solve_device(dkb.ptr, 700, eb.ptr, vb.ptr, 0, s2.handle)
    s1.synchronize()
    s2.synchronize()
    assert np.array_equal(ea.numpy(), ref_a) and np.array_equal(eb.numpy(), ref_b)
    va_c = va.numpy()[..., 0] + 1j * va.numpy()[..., 1]
    vb_c = vb.numpy()[..., 0] + 1j * vb.numpy()[..., 1]
    assert np.array_equal(va_c, vec_a) and np.array_equal(vb_c, vec_b)
    s1.close()
    s2.close()
    p.close()
