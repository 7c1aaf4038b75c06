"""Runtime on the GPU: plan configuration and autotuner, batch graphs over
uneven plans, and the continuous-batching server (uepm_server_*)."""
from __future__ import annotations

import numpy as np
import pytest

from ultimateepm_amd import api, serving
from ultimateepm_amd._abi import UepmError
from ultimateepm_amd.material import load

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-4  # eV: the precision form's bound against FP64


def _dev(ctx, a):
    return api.DeviceArray.from_host(ctx, np.ascontiguousarray(a))


def test_config_validation_and_workspace(gpu_ctx):
    print("case: config round trip Si n=12 mixed", flush=True)
    p = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(12), api.NONLOCAL | api.MIXED, 12)
    assert p.config() == {"tridiag": 0, "tri_block": 0, "bt_bands": 0, "max_chunk": 0, "lanes": 0}
    p.set_config(tridiag=1, tri_block=256, bt_bands=8, max_chunk=512)
    assert p.config() == {"tridiag": 1, "tri_block": 256, "bt_bands": 8, "max_chunk": 512, "lanes": 0}
    for bad in ({"tridiag": 3}, {"tri_block": 128}, {"bt_bands": 5}, {"max_chunk": -1}, {"lanes": 3}):
        with pytest.raises(UepmError):
            p.set_config(**bad)
    assert p.config()["max_chunk"] == 512  # unchanged by a rejected config
    # one chunk in flight up to max_chunk points, two above it
    assert p.workspace_bytes(100) < p.workspace_bytes(512) and p.workspace_bytes(5000) == 2 * p.workspace_bytes(512)
    p.set_config(tridiag=1, tri_block=256, bt_bands=8, max_chunk=512, lanes=1)
    assert p.workspace_bytes(5000) == p.workspace_bytes(512)
    q = api.Plan(gpu_ctx, load("Ge"), api.basis_vectors(10), api.NONLOCAL | api.SOC | api.MIXED, 16)  # dim 274
    q.set_config(tridiag=2)  # the register-resident reduction takes dims up to 288
    r = api.Plan(gpu_ctx, load("Ge"), api.basis_vectors(11), api.NONLOCAL | api.SOC | api.MIXED, 16)  # dim 338
    with pytest.raises(UepmError):
        r.set_config(tridiag=2)
    # dims to 576 are taken (n = 15 with SOC: 566); n = 16 with SOC (614) is refused at creation
    big = api.Plan(gpu_ctx, load("Ge"), api.basis_vectors(15), api.NONLOCAL | api.SOC | api.MIXED, 16)
    assert big.dim == 566
    big.close()
    with pytest.raises(UepmError) as ei:
        api.Plan(gpu_ctx, load("Ge"), api.basis_vectors(16), api.NONLOCAL | api.SOC, 16)
    assert ei.value.code == 3 and "576" in str(ei.value)  # UEPM_ERR_OUT_OF_RANGE
    p.close()
    q.close()
    r.close()


@pytest.mark.parametrize("flags", [api.NONLOCAL | api.VECTORS, api.NONLOCAL | api.VECTORS | api.MIXED])
def test_two_lanes_equal_one_lane(gpu_ctx, flags):
    """A solve over several chunks runs alternate chunks on a second stream with
    its own workspace; every k-point's result is bit-identical to the one-lane
    solve, on the default stream and on a caller's stream, with stage timing on."""
    print(f"case: lanes 2 vs 1, Si n=10 flags {flags}, 1100 k in chunks of 256", flush=True)
    ks = api.bz_sample(13, 1100)
    nb = 10
    p = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(10), flags, nb)
    dk = _dev(gpu_ctx, ks)
    out = {}
    stream = gpu_ctx.stream()
    for lanes, sh, timing in ((1, 0, False), (2, 0, False), (2, stream.handle, True), (0, stream.handle, False)):
        p.set_config(max_chunk=256, lanes=lanes)
        p.set_timing(timing)
        de = api.DeviceArray(gpu_ctx, (len(ks), nb))
        dv = api.DeviceArray(gpu_ctx, (len(ks), nb, p.dim, 2))
        de.zero()
        dv.zero()
        gpu_ctx.synchronize()
        p.solve_device(dk.ptr, len(ks), de.ptr, dv.ptr, 0, sh)
        if sh:
            stream.synchronize()  # the caller's stream alone covers both lanes
        e, v = de.numpy(), dv.numpy()
        if timing:
            assert all(n == 5 for _, n in p.last_timing().values() if n)  # ceil(1100 / 256) launches
        if out:
            assert np.array_equal(e, out["e"]) and np.array_equal(v, out["v"])
        else:
            out = {"e": e, "v": v}
            assert np.all(np.isfinite(e)) and np.all(np.diff(e, axis=1) >= 0)
        de.free()
        dv.free()
    p.set_config(max_chunk=256)  # the host-buffer solve: two lanes, results copied back after the join
    p.set_timing(False)
    ev, _, _ = p.solve(ks)
    assert np.array_equal(ev, out["e"])
    stream.close()
    p.close()


def test_tuned_plan_keeps_results(gpu_ctx):
    print("case: autotune Si n=12 mixed, 1024 k", flush=True)
    ks = api.bz_sample(7, 1024)
    p = api.Plan(gpu_ctx, load("Si"), api.basis_vectors(12), api.NONLOCAL | api.MIXED, 12)
    ev0, _, _ = p.solve(ks)
    dk = _dev(gpu_ctx, ks)
    chosen = p.tune(dk.ptr, len(ks))
    assert chosen["tridiag"] in (1, 2) and chosen["bt_bands"] in (0, 8, 16)
    ev1, _, _ = p.solve(ks)
    assert np.abs(ev1 - ev0).max() <= 1e-8
    p.close()


@pytest.mark.parametrize("sym,n,soc,nb", [("Si", 14, False, 12), ("Ge", 10, True, 16), ("Ge", 11, True, 24),
                                          ("Ge", 14, True, 16)])
def test_wy_back_transform_matches_per_reflector(gpu_ctx, sym, n, soc, nb):
    """Precision form above dim 192: the compact-WY MFMA back-transform (automatic,
    bt_bands 0; one band tile at 16 Ritz vectors, two at 20 and 28, and at dim 518
    two one-tile passes, since two tiles of X do not fit the LDS there) and the
    reflector-by-reflector kernel (bt_bands 16) give the same eigenvalues and
    eigenvectors to the precision form's Ritz step."""
    print(f"case: WY vs per-reflector back-transform, {sym} n={n} soc={soc} nb={nb}", flush=True)
    flags = api.NONLOCAL | api.MIXED | api.VECTORS | (api.SOC if soc else 0)
    p = api.Plan(gpu_ctx, load(sym), api.basis_vectors(n), flags, nb)
    ks = api.bz_sample(808, 64)
    ev0, v0, _ = p.solve(ks)  # automatic: compact WY
    p.set_config(bt_bands=16)
    ev1, v1, _ = p.solve(ks)
    assert np.abs(ev0 - ev1).max() <= 1e-8
    # eigenvectors as subspaces: bands split into clusters at gaps above 50 meV
    # (Kramers pairs with spin-orbit stay together), and each cluster's two bases
    # span the same space, every principal cosine within 1e-5 of 1.  A Ritz vector
    # is within (residual / gap) of its eigenspace, residuals ~1e-4 eV from the
    # complex64 subspace.  The top cluster is left out: its gap to band nb is unknown.
    d = np.diff(ev1, axis=1)
    checked = 0
    for i in range(len(ks)):
        cuts = [0] + [b + 1 for b in range(nb - 1) if d[i, b] > 0.05]
        for a, b in zip(cuts[:-1], cuts[1:]):
            s = np.linalg.svd(v0[i, a:b].conj() @ v1[i, a:b].T, compute_uv=False)
            assert np.abs(s - 1.0).max() <= 1e-5, (i, a, b, s)
            checked += b - a
    assert checked >= len(ks) * nb // 2
    p.close()


@pytest.mark.parametrize("sym,n,soc,nb", [("Ge", 10, True, 16), ("Si", 20, False, 24), ("Ge", 13, True, 12),
                                          ("Ge", 14, True, 16)])
def test_fp64_wy_back_transform_matches_per_reflector(gpu_ctx, sym, n, soc, nb):
    """FP64 form above dim 192: the compact-WY back-transform on FP64 MFMA
    (automatic, bt_bands 0; dims 274, 411 in two passes of 16 bands, 458, and 518
    with X at the unpadded LDS stride) and the reflector-by-reflector kernel
    (bt_bands 16) apply the same Q to the same tridiagonal eigenvectors, so the
    eigenvectors agree elementwise to rounding and the Hellmann-Feynman gradients
    with them; the eigenvalues come before the back-transform and are identical."""
    print(f"case: FP64 WY vs per-reflector back-transform, {sym} n={n} soc={soc} nb={nb}", flush=True)
    flags = api.VECTORS | api.GRADIENTS | (api.SOC if soc else 0)
    p = api.Plan(gpu_ctx, load(sym), api.basis_vectors(n), flags, nb)
    ks = api.bz_sample(909, 48)
    ev0, v0, g0 = p.solve(ks)  # automatic
    p.set_config(bt_bands=16)
    ev1, v1, g1 = p.solve(ks)
    assert np.array_equal(ev0, ev1)
    assert np.isfinite(v0).all() and np.isfinite(g0).all()
    assert np.abs(v0 - v1).max() <= 1e-11
    assert np.abs(g0 - g1).max() <= 1e-9 * max(1.0, np.abs(g1).max())
    # unit eigenvectors (the reflectors are unitary)
    assert np.abs(np.linalg.norm(v0, axis=2) - 1.0).max() <= 1e-12
    p.close()


def test_tuned_plan_above_the_register_grids(gpu_ctx):
    """Above dim 288 the tuner chooses between the LDS-blocked reduction alone
    and its leading steps followed by the register grids; every candidate
    gives the same eigenvalues to the precision form's bar."""
    print("case: autotune Ge SOC n=11 mixed (dim 338), 256 k", flush=True)
    ks = api.bz_sample(11, 256)
    p = api.Plan(gpu_ctx, load("Ge"), api.basis_vectors(11), api.NONLOCAL | api.SOC | api.MIXED, 16)
    assert p.dim == 338
    ev0, _, _ = p.solve(ks)
    dk = _dev(gpu_ctx, ks)
    chosen = p.tune(dk.ptr, len(ks))
    assert chosen["tridiag"] in (0, 1) and chosen["tri_block"] in (256, 512)
    ev1, _, _ = p.solve(ks)
    assert np.abs(ev1 - ev0).max() <= 1e-8
    p.set_config(tridiag=1, tri_block=512, bt_bands=8, max_chunk=0)  # blocked only
    ev2, _, _ = p.solve(ks)
    assert np.abs(ev2 - ev0).max() <= 1e-8
    p.close()


@pytest.mark.parametrize("sym,n,flags,tol", [("Ge", 11, api.NONLOCAL | api.SOC | api.MIXED, 1e-8),
                                          ("Si", 14, api.NONLOCAL, 1e-10)])
def test_blocked_reduction_workgroup_sizes(gpu_ctx, sym, n, flags, tol):
    """The LDS-blocked reduction with 256- and 512-lane workgroups (its global
    leading columns streamed, the LDS block split in items over 4 or 8 waves)
    gives the eigenvalues of the automatic configuration: the precision form at
    dim 338 with the blocked kernel alone for every step, the FP64 form at dim
    259 for its leading steps."""
    print(f"case: blocked reduction 256 / 512 lanes, {sym} n={n} flags {flags}", flush=True)
    p = api.Plan(gpu_ctx, load(sym), api.basis_vectors(n), flags, 16)
    ks = api.bz_sample(77, 128)
    ev0, _, _ = p.solve(ks)
    for tb in (256, 512):
        p.set_config(tridiag=1, tri_block=tb)
        ev, _, _ = p.solve(ks)
        assert np.abs(ev - ev0).max() <= tol, (tb, np.abs(ev - ev0).max())
        again, _, _ = p.solve(ks)  # fixed-order sums, no atomics: the same bits every run
        assert np.array_equal(ev, again), tb
    p.close()


def test_batch_graph_matches_individual_solves(gpu_ctx):
    print("case: batch graph over 4 uneven plans, 2 replays", flush=True)
    specs = [("Si", 8, api.NONLOCAL | api.MIXED, 12), ("Ge", 8, api.NONLOCAL | api.SOC | api.MIXED, 16),
             ("GaAs", 6, api.NONLOCAL | api.SOC, 16), ("InP", 10, api.NONLOCAL, 8), ("Si", 12, api.MIXED, 12)]
    ctx = gpu_ctx
    plans, ks, dks, devs, ref = [], [], [], [], []
    for i, (sym, n, flags, nb) in enumerate(specs):
        p = api.Plan(ctx, load(sym), api.basis_vectors(n), flags, nb)
        k = api.bz_sample(100 + i, 300 + 50 * i)
        ref.append(p.solve(k)[0])
        plans.append(p)
        ks.append(k)
        dks.append(_dev(ctx, k))
        devs.append(api.DeviceArray(ctx, (len(k), nb)))
    b = api.Batch(ctx, [(p, dk.ptr, len(k), de.ptr) for p, dk, k, de in zip(plans, dks, ks, devs)], branches=3)
    for replay in range(2):
        for de in devs:
            de.zero()
        b.launch()
        ctx.synchronize()
        for p, k in zip(plans, ks):  # eager work between the replays
            p.solve(k[:5])
        for i, de in enumerate(devs):
            assert np.abs(de.numpy() - ref[i]).max() <= 1e-12, (replay, specs[i])
    b.close()
    for p in plans:
        p.close()


def test_batch_graph_with_vectors_matches_individual_solves(gpu_ctx):
    """Batch items with eigenvectors (and gradients): the captured graph holds the
    FP64 compact-WY back-transform (Ge + SOC dim 274), the complex64 one with the
    precision form's refinement (dim 274) and the per-reflector kernel (Si dim 89);
    every replay equals the plans' own solves (to 1e-12)."""
    print("case: batch graph with eigenvectors, 3 plans, 2 replays", flush=True)
    specs = [("Ge", 10, api.NONLOCAL | api.SOC | api.GRADIENTS, 16), ("Ge", 10, api.NONLOCAL | api.SOC | api.MIXED |
                                                                      api.VECTORS, 16),
             ("Si", 8, api.NONLOCAL | api.GRADIENTS, 8)]
    ctx = gpu_ctx
    plans, ks, dks, outs, ref = [], [], [], [], []
    for i, (sym, n, flags, nb) in enumerate(specs):
        p = api.Plan(ctx, load(sym), api.basis_vectors(n), flags, nb)
        k = api.bz_sample(300 + i, 150 + 40 * i)
        ev, vec, gr = p.solve(k, vectors=True)  # a UEPM_GRADIENTS plan computes the vectors too
        ref.append((ev, vec, gr))
        plans.append(p)
        ks.append(k)
        dks.append(_dev(ctx, k))
        de = api.DeviceArray(ctx, (len(k), nb))
        dv = api.DeviceArray(ctx, (len(k), nb, p.dim, 2))
        dg = api.DeviceArray(ctx, (len(k), nb, 3)) if flags & api.GRADIENTS else None
        outs.append((de, dv, dg))
    items = [(p, dk.ptr, len(k), de.ptr, dv.ptr, dg.ptr if dg else 0)
             for p, dk, k, (de, dv, dg) in zip(plans, dks, ks, outs)]
    b = api.Batch(ctx, items, branches=3)
    for replay in range(2):
        for de, dv, dg in outs:
            de.zero()
            dv.zero()
            if dg:
                dg.zero()
        b.launch()
        ctx.synchronize()
        for p, k in zip(plans, ks):  # eager work between the replays
            p.solve(k[:5])
        for i, (de, dv, dg) in enumerate(outs):
            ev, vec, gr = ref[i]
            assert np.abs(de.numpy() - ev).max() <= 1e-12, (replay, specs[i])
            v = dv.numpy()
            assert np.abs(v[..., 0] + 1j * v[..., 1] - vec).max() <= 1e-12, (replay, specs[i])
            if dg:
                assert np.abs(dg.numpy() - gr).max() <= 1e-12 * np.abs(gr).max(), (replay, specs[i])
    b.close()
    for p in plans:
        p.close()


def _workload(seed, count):
    return serving.poisson_workload(seed, count, rate=400.0, n_min=6, n_max=10, k_max=256)


def test_server_result_alone_equals_result_in_the_mix(gpu_ctx):
    print("case: server alone vs mix, 24 requests", flush=True)
    wl = _workload(11, 24)
    srv = api.Server(gpu_ctx, 8 << 30, max_inflight=6)
    mix, summary = serving.run(srv, wl, mixed=True, realtime=False)
    assert summary["requests"] == 24
    for i in (0, 5, 17, 23):
        alone, _ = serving.run(srv, [wl[i]], mixed=True, realtime=False)
        assert np.abs(alone[0][0] - mix[i][0]).max() <= 1e-12, i
    st = srv.stats()
    assert st["completed"] == 28 and st["queued"] == 0 and st["inflight"] == 0
    srv.close()


@pytest.mark.parametrize("flags", [api.NONLOCAL, api.NONLOCAL | api.MIXED])
def test_server_chunks_within_slot_arena(gpu_ctx, flags):
    """A slot's workspace arena (budget / slots) smaller than a request: the
    request runs in chunks carved out of the arena, and every k-point comes out
    bit-identical to a plain solve (each k-point's work does not depend on the
    chunk it runs in).  Two slots, four requests of 1000 k-points."""
    print(f"case: server arena chunking, Si n=12 flags {flags}", flush=True)
    mat, nb = load("Si"), 12
    G = api.basis_vectors(12)
    p = api.Plan(gpu_ctx, mat, G, flags, nb)
    per_k = p.workspace_bytes(1)
    ks = [api.bz_sample(900 + i, 1000) for i in range(4)]
    ref = [p.solve(k)[0] for k in ks]
    p.close()
    srv = api.Server(gpu_ctx, 2 * 96 * per_k, max_inflight=2)  # under 96 k-points per chunk
    tickets = [srv.submit(mat, 12, flags, nb, k) for k in ks]
    for t, r in zip(tickets, ref):
        ev, _, _ = srv.wait(t)
        assert np.array_equal(ev, r)
    st = srv.stats()
    assert st["completed"] == 4 and st["used_bytes"] <= 2 * 96 * per_k
    srv.close()


def test_server_mixed_vs_fp64_under_load(gpu_ctx):
    print("case: server mixed vs FP64, 40 requests at 400 req/s", flush=True)
    wl = _workload(12, 40)
    srv = api.Server(gpu_ctx, 8 << 30, max_inflight=8)
    r64, s64 = serving.run(srv, wl, mixed=False, realtime=True)
    rmx, smx = serving.run(srv, wl, mixed=True, realtime=True)
    worst = max(np.abs(a[0] - b[0]).max() for a, b in zip(r64, rmx))
    print(f"mixed vs FP64 max |dE| = {worst * 1e3:.3e} meV; p99 {s64['latency_p99_ms']} / {smx['latency_p99_ms']} ms")
    assert worst <= MIXED_TOL
    assert smx["latency_p50_ms"] > 0 and smx["latency_p99_ms"] >= smx["latency_p50_ms"]
    srv.close()


def test_server_errors_and_budget(gpu_ctx):
    print("case: server errors (InP SOC, budget)", flush=True)
    srv = api.Server(gpu_ctx, 8 << 30, max_inflight=2)
    k = api.bz_sample(3, 10)
    t_bad = srv.submit(load("InP"), 6, api.NONLOCAL | api.SOC, 8, k)
    t_ok = srv.submit(load("Si"), 6, api.NONLOCAL, 8, k)
    with pytest.raises(UepmError) as ei:
        srv.wait(t_bad)
    assert "Unsupported period" in str(ei.value)
    ev, lat, _ = srv.wait(t_ok)
    assert ev.shape == (10, 8) and lat > 0
    srv.close()
    # 64 KiB: less than one k-point's workspace at N=181 (a 1 MiB budget still fits one point per chunk)
    tiny = api.Server(gpu_ctx, 64 << 10, max_inflight=1)
    t = tiny.submit(load("Si"), 12, api.NONLOCAL, 8, k)
    with pytest.raises(UepmError) as ei:
        tiny.wait(t)
    assert ei.value.code == 6  # UEPM_ERR_NO_MEMORY
    tiny.close()
