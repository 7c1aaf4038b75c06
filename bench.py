"""Benchmark: k-points/s of the batched EPM band solve on one MI355X.

Workload (one "step"): Si with non-local corrections (materials.yaml), basis
n = 12 nearest shells (N = 181, complex128 Hermitian H(k)), 200,000 k-points
drawn with the seeded first-BZ sampler (seed 20261202), lowest 12 eigenvalues
plus eigenvectors, one batched solve call with inputs resident in HBM.

Prints one JSON line (see the metric / roofline / cpu_baseline fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "k-pts/s Si nonlocal N≈181 FP64 vs mixed @≤0.1 meV; serving p50/p99 latency"
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
# which datapath each reduction kernel issues its arithmetic on.  FP32 VALU
# instructions issue at twice the FP64 rate (the complex64 grids' VALU
# instruction counts would exceed their SIMD-cycles at four cycles each), so
# the FP32 vector peak is the 157.3 TF/s that FP32 MFMA also has.
DATAPATH = {
    ("mixed", "tridiag_reg_kernel"): "VALU FP32 (157.3 TF/s, the FP32 MFMA rate too): no MFMA in the register-grid chain",
    ("fp64", "tridiag_reg_kernel"): "VALU FP64 (78.6 TF/s, equal to the FP64 MFMA rate)",
    ("mixed", "tridiag_blocked_kernel"): "VALU FP32 symmetric products, MFMA FP32 panel updates (157.3 TF/s either way)",
    ("mixed", "tridiag_blocked+reg_kernel"): "VALU FP32 symmetric products, MFMA FP32 panel updates, then the register grids",
    ("fp64", "tridiag_blocked_kernel"): "VALU FP64 symmetric products, MFMA FP64 panel updates (78.6 TF/s either way)",
    ("fp64", "tridiag_blocked+reg_kernel"): "VALU FP64 symmetric products, MFMA FP64 panel updates, then the register grids",
}
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def log(msg: str) -> None:
    """A progress line on stderr (the JSON result is the one stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def tridiag_flops(n: int) -> float:
    """Algorithmic real flops of the Hermitian Householder reduction: (16/3) n^3."""
    return 16.0 / 3.0 * n ** 3


def host_cpu() -> dict:
    """CPU model and the cores this process may use (the job's share of the host)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    return {"model": model, "host_cpus": os.cpu_count(), "affinity_cpus": affinity, "threads": max(1, threads)}


def cpu_baseline(mat, G, nb, ks, seconds: float):
    """The oracle (C++ restatement of Compute_parallel, OpenMP, one solver object per
    thread, dynamic schedule) on a bounded sample of the headline k-set: on every
    core of the job's CPU share and on one core, with the assembly (Bessel-heavy
    non-local term) vs eigensolve split of the one-core run."""
    from oracle import oracle as O
    from ultimateepm_amd import api

    cpu = host_cpu()
    threads = cpu["threads"]
    flags = api.NONLOCAL | api.VECTORS
    probe = ks[: 2 * threads]
    t0 = time.perf_counter()
    O.solve(mat, G, flags, nb, probe, nthreads=threads)
    rate = len(probe) / (time.perf_counter() - t0)
    m = int(max(2 * threads, min(len(ks), rate * seconds)))
    t0 = time.perf_counter()
    O.solve(mat, G, flags, nb, ks[:m], nthreads=threads)
    dt = time.perf_counter() - t0
    # one core: a quarter of the time budget
    m1 = int(max(4, min(len(ks), rate / threads * seconds / 4)))
    t0 = time.perf_counter()
    O.solve(mat, G, flags, nb, ks[:m1], nthreads=1)
    dt1 = time.perf_counter() - t0
    # assembly alone (SetMatrix + hermitisation) on the same one-core sample
    t0 = time.perf_counter()
    for k in ks[:m1]:
        O.assemble(mat, G, k, nl=True, soc=False, herm=True)
    dt_asm = time.perf_counter() - t0
    return {
        "value": round(m / dt, 2),
        "unit": "k-pts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {m} k-points of the same seeded BZ set, Si non-local N={len(G)}, {nb} bands + "
                  f"eigenvectors (all eigenvectors, as Eigen computes them), oracle built -O3 "
                  f"-march=x86-64-v3 -ffp-contract=off, {threads} OpenMP threads, {dt:.1f} s",
        "cpu_model": cpu["model"],
        "host_cpus": cpu["host_cpus"],
        "affinity_cpus": cpu["affinity_cpus"],
        "one_core_kpts_per_s": round(m1 / dt1, 3),
        "one_core_sample": f"first {m1} k-points, 1 thread, {dt1:.1f} s",
        "one_core_split": {"assembly_ms_per_k": round(dt_asm / m1 * 1e3, 2),
                           "eigensolve_ms_per_k": round(max(dt1 - dt_asm, 0.0) / m1 * 1e3, 2),
                           "assembly_frac": round(dt_asm / dt1, 3)},
        "parallel_speedup": round((m / dt) / (m1 / dt1), 2),
    }


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_si_nl_n181.txt")


def measured_traffic(mode: str, kernel: str, N: int):
    """HBM bytes per k-point of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    for line in open(TRAFFIC_SUMMARY):
        if line.startswith("TRAFFIC "):
            rec = json.loads(line[len("TRAFFIC "):])
            if rec["mode"] == mode and rec["kernel"] == kernel and rec["N"] == N:
                return rec["bytes_per_k"]
    return None


def copy_calibration(mib: int, reps: int) -> int:
    """Copies `mib` MiB device-to-device `reps` times (known traffic: read + write of that size)."""
    from ultimateepm_amd import api

    ctx = api.Context(0)
    n = mib << 20
    a = api.DeviceArray(ctx, (n // 8,))
    b = api.DeviceArray(ctx, (n // 8,))
    a.zero()
    ctx.copy(b.ptr, a.ptr, n)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.copy(b.ptr, a.ptr, n)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"calibration": "copy_kernel", "bytes_read": n, "bytes_written": n, "launches": reps + 1,
                      "GB_per_s": round(2 * n * reps / dt / 1e9, 1)}), flush=True)
    a.free()
    b.free()
    ctx.close()
    return 0


def serving_leg(ctx, requests: int, seed: int = SEED, budget_gib: int = 64, inflight: int = 8) -> dict:
    """Serving form: Poisson arrivals (seeded); material uniform over Si/Ge/GaAs/InP,
    n in 6..14 shells, SOC on for half (off for InP), log-uniform 1..4096
    k-points, 8..24 bands; the same request stream in FP64 and in the precision
    form on a continuous-batching server (HBM budget, in-flight slots).  Three
    operating points per form: the configuration's 200 req/s; saturation (every
    request submitted at once: the throughput the server sustains); and a
    Poisson stream at half the saturation request rate, where latency is
    service time plus modest queueing."""
    from ultimateepm_amd import api, serving

    wl = serving.poisson_workload(seed, requests, rate=200.0, n_min=6, n_max=14, k_max=4096)
    srv = api.Server(ctx, budget_gib << 30, max_inflight=inflight)
    out = {"requests": requests, "rate_req_s": 200.0, "hbm_budget_gib": budget_gib, "max_inflight": inflight,
           "workload": "Poisson 200 req/s (seed %d); {Si,Ge,GaAs,InP} x n 6..14 x SOC half; k log-uniform "
                       "1..4096; nb 8..24; non-local" % seed}
    res = {}
    for mode in ("fp64", "mixed"):
        mixed = mode == "mixed"
        serving.run(srv, wl[: min(8, len(wl))], mixed=mixed, realtime=False)  # plan warm-up
        r, summ = serving.run(srv, wl, mixed=mixed, realtime=True)
        res[mode] = r
        out[mode] = summ
        log(f"serving {mode}: 200 req/s done, saturation run")
        _, sat = serving.run(srv, wl, mixed=mixed, realtime=False)
        out[f"{mode}_saturation"] = sat
        cap_req_s = sat["requests"] / sat["elapsed_s"]
        rate = 0.5 * cap_req_s
        wl_half = serving.poisson_workload(seed + 1, requests, rate=rate, n_min=6, n_max=14, k_max=4096)
        log(f"serving {mode}: half-load run at {rate:.1f} req/s")
        _, half = serving.run(srv, wl_half, mixed=mixed, realtime=True)
        half["offered_rate_req_s"] = round(rate, 2)
        out[f"{mode}_half_load"] = half
    worst = max(float(np.abs(a[0] - b[0]).max()) for a, b in zip(res["fp64"], res["mixed"]))
    out["mixed_vs_fp64_max_meV"] = worst * 1e3
    srv.close()
    return out


def ge_soc_leg(ctx, nk: int = 100_000, seed: int = SEED) -> dict:
    """Precision form at the spinor size: Ge with spin-orbit (materials.yaml,
    non-local on), n = 10 shells (dim 2 x 137 = 274), nk seeded first-BZ
    k-points, lowest 16 eigenvalues; both forms, one untimed and one timed
    solve each (device-resident k), max / RMS difference in meV."""
    from ultimateepm_amd import api
    from ultimateepm_amd.material import load

    mat = load("Ge")
    G = api.basis_vectors(10)
    nb = 16
    ks = api.bz_sample(seed + 3, nk)
    dk = api.DeviceArray.from_host(ctx, ks)
    out = {"workload": f"Ge + SOC (materials.yaml, non-local), n=10 (dim {2 * len(G)}), {nk} seeded first-BZ "
                       f"k-points, lowest {nb} eigenvalues", "nk": nk, "dim": 2 * len(G), "nb": nb}
    ev = {}
    for mode in ("fp64", "mixed"):
        flags = api.NONLOCAL | api.SOC | (api.MIXED if mode == "mixed" else 0)
        p = api.Plan(ctx, mat, G, flags, nb)
        de = api.DeviceArray(ctx, (nk, nb))
        p.solve_device(dk.ptr, nk, de.ptr)
        ctx.synchronize()
        t0 = time.perf_counter()
        p.solve_device(dk.ptr, nk, de.ptr)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        p.set_timing(True)
        p.solve_device(dk.ptr, nk, de.ptr)
        ctx.synchronize()
        out[f"{mode}_kpts_per_s"] = round(nk / dt, 1)
        out[f"{mode}_stage_ms"] = {k: round(v[0], 3) for k, v in p.last_timing().items()}
        ev[mode] = de.numpy()
        p.close()
        de.free()
    d = ev["mixed"] - ev["fp64"]
    out["mixed_vs_fp64_max_meV"] = float(np.abs(d).max() * 1e3)
    out["mixed_vs_fp64_rms_meV"] = float(np.sqrt(np.mean(d ** 2)) * 1e3)
    out["mixed_over_fp64"] = round(out["mixed_kpts_per_s"] / out["fp64_kpts_per_s"], 3)
    dk.free()
    return out


def dims_leg(ctx, seed: int = SEED, only=None, vectors: bool = False) -> list:
    """Throughput by basis dimension: Si (n = 6..14 shells, dims 59..259) and Ge with
    spin-orbit (n = 6..14, dims 118..518), non-local, 16 bands, eigenvalues only
    (`vectors`: and eigenvectors), 2048 seeded BZ k-points (512 above dim 200),
    both forms; one untimed solve, then one timed solve (wall clock over a
    device-resident batch) and the stage split of a third.  `only`: the
    dimensions to run (default: all)."""
    from ultimateepm_amd import api
    from ultimateepm_amd.material import load

    rows = []
    for sym, soc in (("Si", False), ("Ge", True)):
        for n in (6, 8, 10, 12, 14):
            G = api.basis_vectors(n)
            if only and len(G) * (2 if soc else 1) not in only:
                continue
            for mode in ("fp64", "mixed"):
                flags = api.NONLOCAL | (api.SOC if soc else 0) | (api.MIXED if mode == "mixed" else 0)
                if vectors:
                    flags |= api.VECTORS
                nb = 16
                p = api.Plan(ctx, load(sym), G, flags, nb)
                nk = 2048 if p.dim <= 200 else 512
                dk = api.DeviceArray.from_host(ctx, api.bz_sample(seed + 5, nk))
                de = api.DeviceArray(ctx, (nk, nb))
                dv = api.DeviceArray(ctx, (nk, nb, p.dim, 2)) if vectors else None
                vptr = dv.ptr if vectors else 0
                p.solve_device(dk.ptr, nk, de.ptr, vptr)
                ctx.synchronize()
                t0 = time.perf_counter()
                p.solve_device(dk.ptr, nk, de.ptr, vptr)
                ctx.synchronize()
                dt = time.perf_counter() - t0
                p.set_timing(True)
                p.solve_device(dk.ptr, nk, de.ptr, vptr)
                ctx.synchronize()
                rows.append({"material": sym, "shells": n, "soc": soc, "dim": p.dim, "mode": mode, "nk": nk,
                             "vectors": vectors, "kpts_per_s": round(nk / dt, 1),
                             "stage_ms": {k: round(v[0], 3) for k, v in p.last_timing().items()}})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
                p.close()
                dk.free()
                de.free()
                if vectors:
                    dv.free()
    return rows


def batching_leg(ctx, nk_per_item: int, branches: int = 8, ws_cap_gib: float = 2.0, seed: int = SEED) -> dict:
    """Batching form: {Si, Ge, GaAs, InP} x n in {6, 8, 10, 12} shells x SOC off/on
    (InP + SOC excluded: its anion period has no spin-orbit radial model), nk
    seeded first-BZ k-points each, all captured into one hipGraph (concurrent
    branches) and launched once; reduction variant and band grouping chosen per
    dimension by the autotuner.  Run in the precision form and in FP64."""
    from ultimateepm_amd import api
    from ultimateepm_amd.material import load

    combos = [(m, n, soc) for m in ("Si", "Ge", "GaAs", "InP") for n in (6, 8, 10, 12) for soc in (False, True)
              if not (m == "InP" and soc)]
    ks = {}
    out = {"items": len(combos), "nk_per_item": nk_per_item, "branches": branches}
    evals = {}
    for mode in ("mixed", "fp64"):
        tuned = {}  # dimension bucket -> configuration chosen by the autotuner
        plans, bufs, items = [], [], []
        t_tune = time.perf_counter()
        for i, (m, n, soc) in enumerate(combos):
            nb = 16 if soc else 8
            flags = api.NONLOCAL | (api.SOC if soc else 0) | (api.MIXED if mode == "mixed" else 0)
            p = api.Plan(ctx, load(m), api.basis_vectors(n), flags, nb)
            per_k = max(1, p.workspace_bytes(1))
            cap = int(max(1, min(4096, ws_cap_gib * (1 << 30) // per_k)))
            key = (p.dim, mode)
            if key not in tuned:
                sample = api.bz_sample(seed + 7, min(1024, nk_per_item))
                dsample = api.DeviceArray.from_host(ctx, sample)
                p.set_config(max_chunk=cap)
                tuned[key] = p.tune(dsample.ptr, len(sample))
                dsample.free()
            cfg = dict(tuned[key])
            cfg["max_chunk"] = cap
            p.set_config(**cfg)
            if i not in ks:
                ks[i] = api.bz_sample(seed + 1000 + i, nk_per_item)
            dk = api.DeviceArray.from_host(ctx, ks[i])
            de = api.DeviceArray(ctx, (nk_per_item, nb))
            plans.append(p)
            bufs.append((dk, de))
            items.append((p, dk.ptr, nk_per_item, de.ptr))
        t_tune = time.perf_counter() - t_tune
        log(f"batching {mode}: tuned and set up in {t_tune:.1f} s, capturing the graph")
        b = api.Batch(ctx, items, branches=branches)
        b.launch()
        ctx.synchronize()
        t0 = time.perf_counter()
        b.launch()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        evals[mode] = [de.numpy() for _, de in bufs]
        b.close()
        # the same plans solved one after another, each alone on the chip (one
        # untimed solve first): what the graph's concurrent branches save, and
        # where the time goes by dimension
        alone, by_dim = 0.0, {}
        for p, (dk, de) in zip(plans, bufs):
            p.solve_device(dk.ptr, nk_per_item, de.ptr)
            ctx.synchronize()
            t1 = time.perf_counter()
            p.solve_device(dk.ptr, nk_per_item, de.ptr)
            ctx.synchronize()
            t1 = time.perf_counter() - t1
            alone += t1
            by_dim[p.dim] = by_dim.get(p.dim, 0.0) + t1
        out[mode] = {"kpts_per_s": round(len(combos) * nk_per_item / dt, 1), "graph_launch_s": round(dt, 4),
                     "plans_alone_s": round(alone, 4), "graph_over_alone": round(dt / alone, 3),
                     "alone_s_by_dim": {f"dim{d}": round(t, 4) for d, t in sorted(by_dim.items())},
                     "tune_and_setup_s": round(t_tune, 2),
                     "tuned": {f"dim{d}": c for (d, _), c in sorted(tuned.items())}}
        for p in plans:
            p.close()
        for dk, de in bufs:
            dk.free()
            de.free()
    out["mixed_over_fp64"] = round(out["mixed"]["kpts_per_s"] / out["fp64"]["kpts_per_s"], 3)
    out["mixed_vs_fp64_max_meV"] = max(float(np.abs(a - b).max()) for a, b in zip(evals["mixed"], evals["fp64"])) * 1e3
    return out


def mesh_leg(ctx, steps: int, n_energies: int = 1000, cpu: bool = True, cpu_sample: int = 512) -> dict:
    """BandsOnBZ + DOS_MeshBZ on the reference's example mesh (examples/meshes/
    bz_mesh.msh: 14,705 vertices, 73,728 tetrahedra): Si local Cohen-Bergstresser
    factors, n = 10 shells (N = 137), 4 + 12 + 4 levels with Hellmann-Feynman
    gradients at every vertex (device-resident k), then the tetrahedron DOS of
    the 16 exported bands on n_energies energies each."""
    import gzip
    import shutil
    import tempfile

    from ultimateepm_amd import api, mesh
    from ultimateepm_amd.material import PARAMS_DIR, load

    mat = load("Si", os.path.join(PARAMS_DIR, "materials-local-cohen.yaml"))
    tmp = tempfile.mkdtemp(prefix="uepm_mesh_")
    path = os.path.join(tmp, "bz_mesh.msh")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "bz_mesh.msh.gz"), "rb") as src, open(path, "wb") as dst:
        shutil.copyfileobj(src, dst)
    m = mesh.MeshBZ.read_mesh_geometry_from_msh_file(ctx, path, mat.get_lattice_constant_meter())
    shutil.rmtree(tmp, ignore_errors=True)
    nv, nt, _ = m.counts()
    G = api.basis_vectors(10)
    nb = 20
    k = m.kpoints(False)
    dk = api.DeviceArray.from_host(ctx, k)
    stream = ctx.stream()
    out = {"workload": f"Si local (materials-local-cohen.yaml), n=10 (N={len(G)}), {nb} levels + gradients at "
                       f"the {nv} vertices of bz_mesh.msh; tetrahedron DOS of 16 bands x {n_energies} energies "
                       f"over its {nt} tetrahedra", "vertices": nv, "tetrahedra": nt}
    evals = {}
    for mode in ("fp64", "mixed"):
        flags = api.GRADIENTS | (api.MIXED if mode == "mixed" else 0)
        plan = api.Plan(ctx, mat, G, flags, nb)
        dE, dG = api.DeviceArray(ctx, (nv, nb)), api.DeviceArray(ctx, (nv, nb, 3))
        plan.solve_device(dk.ptr, nv, dE.ptr, 0, dG.ptr, stream.handle)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.solve_device(dk.ptr, nv, dE.ptr, 0, dG.ptr, stream.handle)
        stream.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[f"{mode}_bands_kpts_per_s"] = round(nv / dt, 1)
        out[f"{mode}_bands_ms"] = round(dt * 1e3, 3)
        evals[mode] = (dE, dG)
        plan.close()
    E64 = evals["fp64"][0].numpy()
    out["mixed_vs_fp64_max_meV"] = float(np.abs(evals["mixed"][0].numpy() - E64).max() * 1e3)
    # DOS of the 16 bands BandsOnBZ exports (4 valence, 12 conduction), on the device bands
    dE, dG = evals["fp64"]
    m.set_bands_device(dE.ptr, dG.ptr, nb, 0, 16, stream.handle)
    lo, hi = E64[:, :16].min(axis=0) - 0.1, E64[:, :16].max(axis=0) + 0.1
    den, ddos = api.DeviceArray(ctx, (16, n_energies)), api.DeviceArray(ctx, (16, n_energies))
    row = n_energies * 8

    def dos_all():
        for b in range(16):
            m.dos_device(b, n_energies, float(lo[b]), float(hi[b]), 1.0, 2.0, False, den.ptr + b * row,
                         ddos.ptr + b * row, stream.handle)

    dos_all()
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dos_all()
    stream.synchronize()
    dt = (time.perf_counter() - t0) / steps
    out["dos_ms_16_bands"] = round(dt * 1e3, 3)
    out["dos_tetra_energy_pairs_per_s"] = round(16 * n_energies * nt / dt, 1)
    if cpu:
        from oracle import oracle

        sel = np.arange(0, nv, max(1, nv // cpu_sample))[:cpu_sample]
        t0 = time.perf_counter()
        oracle.solve(mat, G, api.GRADIENTS, nb, k[sel], nthreads=16)
        cpu_k = len(sel) / (time.perf_counter() - t0)
        geo = m.geometry()
        Eh, Gh = dE.numpy()[:, :16], dG.numpy()[:, :16]
        t0 = time.perf_counter()
        oracle.tetra_dos(geo["xyz"], geo["tetrahedra"], Eh, Gh, 4, n_energies, float(lo[4]), float(hi[4]), 1.0, 2.0,
                         nthreads=16)
        cpu_dos = time.perf_counter() - t0
        out["cpu_baseline"] = {"bands_kpts_per_s": round(cpu_k, 1), "dos_ms_per_band": round(cpu_dos * 1e3, 1),
                               "cores": 16, "kind": "port",
                               "sample": f"oracle solve at {len(sel)} mesh vertices (FP64, gradients); oracle "
                                         f"tetrahedron DOS of band 4 on {n_energies} energies"}
        out["gpu_over_cpu_bands"] = round(out["fp64_bands_kpts_per_s"] / cpu_k, 1)
        out["gpu_over_cpu_dos"] = round(cpu_dos * 1e3 / (out["dos_ms_16_bands"] / 16), 1)
    stream.close()
    m.close()
    return out


def epsilon_grid(n: int) -> np.ndarray:
    """DielectricFunction::generate_k_points_grid(n, n, n, 0, false)
    (DielectricFunction.cpp:59-74): the grid points of [-1, 1]^3 in the first BZ."""
    pts = []
    for i in range(n - 1):
        for j in range(n - 1):
            for k in range(n - 1):
                x, y, z = (-1.0 + 2.0 * i / (n - 1), -1.0 + 2.0 * j / (n - 1), -1.0 + 2.0 * k / (n - 1))
                if abs(x) <= 1 and abs(y) <= 1 and abs(z) <= 1 and abs(x) + abs(y) + abs(z) <= 1.5:
                    pts.append((x, y, z))
    return np.array(pts)


def dielectric_leg(ctx, nq: int = 2, cpu: bool = True, cpu_sample: int = 32) -> dict:
    """The reference's epsilon example (examples/si.epsilon.yaml): Si local,
    n = 10 (N = 137), 16 bands, 40^3 grid k-points in the first BZ, energies
    0..20 eV in 1 meV steps (20,001), eta = 0.05 eV; per q-point: eigenpairs at
    k + q, 48 overlaps per k, Lorentzian sums (k eigenpairs solved once)."""
    from ultimateepm_amd import api
    from ultimateepm_amd.material import load

    mat = load("Si")
    G = api.basis_vectors(10)
    ks = epsilon_grid(40)
    en = np.round(np.arange(0.0, 20.0 + 1e-9, 0.001), 12)
    eta = 0.05
    qs = np.array([[0.01 * (i + 1), 0.0, 0.0] for i in range(nq)])
    plan = api.Plan(ctx, mat, G, api.VECTORS, 16)
    api.dielectric(plan, ks[:256], qs[:1], en[:64], eta)  # warm-up: workspaces
    t0 = time.perf_counter()
    S = api.dielectric(plan, ks, qs, en, eta)
    dt = time.perf_counter() - t0
    plan.close()
    nk = len(ks)
    out = {"workload": f"Si local (materials.yaml), n=10 (N={len(G)}), 16 bands, {nk} grid k-points (40^3 in the "
                       f"first BZ), {len(en)} energies, eta {eta}, {nq} q-points along x",
           "kpoints": nk, "qpoints": nq, "energies": len(en), "seconds": round(dt, 3),
           "kq_pairs_per_s": round(nk * nq / dt, 1),
           "lorentz_terms_per_s": round(nk * nq * 48 * len(en) / dt, 1),
           "finite": bool(np.all(np.isfinite(S)))}
    if cpu:
        from oracle import oracle

        sel = ks[:: max(1, nk // cpu_sample)][:cpu_sample]
        t0 = time.perf_counter()
        oracle.dielectric(mat, G, False, 16, sel, qs[:1], en, eta, nthreads=16)
        cpu_dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"kq_pairs_per_s": round(len(sel) / cpu_dt, 2), "cores": 16, "kind": "port",
                               "sample": f"oracle at {len(sel)} grid k-points x 1 q, same energies (eigen-solves "
                                         "at k and k+q included)"}
        out["gpu_over_cpu"] = round(out["kq_pairs_per_s"] / out["cpu_baseline"]["kq_pairs_per_s"], 1)
    return out


def states_leg(ctx, steps: int, cpu: bool = True, cpu_tetra: int = 200) -> dict:
    """The reference's fullstates app (apps/fullstates.cpp): BZ_States over
    bz_mesh.msh with Si (materials-cohen.yaml), n = 10 (N = 137), 12 bands,
    eigenpairs at the 14,705 vertices and at vertex + q (q = (3e-14, 0, 0)),
    then the tetrahedron-weighted dielectric function on 0..20 eV in 0.01 eV
    steps, eta = 0.05 eV."""
    import gzip
    import shutil
    import tempfile

    from ultimateepm_amd import api, mesh
    from ultimateepm_amd.material import PARAMS_DIR, load

    mat = load("Si", os.path.join(PARAMS_DIR, "materials-cohen.yaml"))
    tmp = tempfile.mkdtemp(prefix="uepm_states_")
    path = os.path.join(tmp, "bz_mesh.msh")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "bz_mesh.msh.gz"), "rb") as src, open(path, "wb") as dst:
        shutil.copyfileobj(src, dst)
    m = mesh.BZStates.read_mesh_geometry_from_msh_file(ctx, path, mat.get_lattice_constant_meter())
    shutil.rmtree(tmp, ignore_errors=True)
    nv, nt, _ = m.counts()
    G = api.basis_vectors(10)
    q = np.array([3e-14, 0.0, 0.0])
    en = []
    e = 0.0
    while e <= 20.0 + 0.01:  # fullstates.cpp:88-90
        en.append(e)
        e += 0.01
    en = np.array(en)
    out = {"workload": f"Si (materials-cohen.yaml), n=10 (N={len(G)}), 12 bands + eigenvectors at the {nv} vertices "
                       f"of bz_mesh.msh and at vertex + q; dielectric function over its {nt} tetrahedra on "
                       f"{len(en)} energies", "vertices": nv, "tetrahedra": nt, "energies": len(en)}
    for mode in ("fp64", "mixed"):
        plan = api.Plan(ctx, mat, G, api.VECTORS | (api.MIXED if mode == "mixed" else 0), 12)
        m.compute_eigenstates(plan)  # warm-up: workspaces
        t0 = time.perf_counter()
        for _ in range(steps):
            m.compute_eigenstates(plan)
            m.compute_shifted_eigenstates(q, plan)
        dt = (time.perf_counter() - t0) / steps
        out[f"{mode}_states_kpts_per_s"] = round(2 * nv / dt, 1)
        out[f"{mode}_states_ms"] = round(dt * 1e3, 3)
        plan.close()
    m.compute_dielectric_function(en, 0.05)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.compute_dielectric_function(en, 0.05)
    dt = (time.perf_counter() - t0) / steps
    out["dielectric_ms"] = round(dt * 1e3, 3)
    out["lorentz_terms_per_s"] = round(nv * 32 * len(en) / dt, 1)
    if cpu:
        from oracle import oracle

        geo = m.geometry()
        tet = geo["tetrahedra"][:: max(1, nt // cpu_tetra)][:cpu_tetra]
        ids, inv = np.unique(tet.ravel(), return_inverse=True)
        t0 = time.perf_counter()
        oracle.bz_states_dielectric(mat, G, False, 12, geo["xyz"][ids], inv.reshape(-1, 4), q, en, 0.05, nthreads=16)
        cpu_dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"vertex_pairs_per_s": round(len(ids) / cpu_dt, 1), "cores": 16, "kind": "port",
                               "sample": f"oracle over {len(tet)} tetrahedra ({len(ids)} vertices): eigenpairs at k "
                                         "and k + q, then the tetrahedron sums, same energies"}
        gpu_pairs = nv / (2 * nv / out["fp64_states_kpts_per_s"] + out["dielectric_ms"] / 1e3)
        out["gpu_over_cpu"] = round(gpu_pairs / out["cpu_baseline"]["vertex_pairs_per_s"], 1)
    m.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nk", type=int, default=200_000)
    ap.add_argument("--nb", type=int, default=12)
    ap.add_argument("--shells", type=int, default=12)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--timing-steps", type=int, default=2, help="extra steps with per-stage HIP-event timing")
    ap.add_argument("--ge-soc-nk", type=int, default=100_000,
                    help="k-points of the Ge + SOC (dim 274) leg of the default run (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed", help="comma list of fp64, mixed; the value is mixed's")
    ap.add_argument("--lanes", type=int, default=0, choices=[0, 1, 2],
                    help="chunks in flight (plan configuration; 0 = automatic: 2 for a multi-chunk solve)")
    ap.add_argument("--form", default="throughput", choices=["throughput", "serving", "batching", "dims"],
                    help="throughput: the headline batched solve (default); serving / batching: those forms "
                         "alone; dims: throughput by basis dimension")
    ap.add_argument("--requests", type=int, default=200, help="requests of --form serving")
    ap.add_argument("--dims", default="", help="comma list of basis dimensions for --form dims (default: all)")
    ap.add_argument("--dims-vectors", action="store_true", help="--form dims: eigenvectors too")
    ap.add_argument("--batch-nk", type=int, default=20000, help="k-points per plan of --form batching")
    ap.add_argument("--batch-leg-nk", type=int, default=20000,
                    help="k-points per plan of the batching leg of the default run (0: skip)")
    ap.add_argument("--serving-requests", type=int, default=100,
                    help="requests of the serving leg of the default run (0: skip)")
    ap.add_argument("--mesh-steps", type=int, default=3,
                    help="timed repetitions of the mesh leg of the default run (0: skip)")
    ap.add_argument("--dielectric-q", type=int, default=2,
                    help="q-points of the dielectric leg of the default run (0: skip)")
    ap.add_argument("--states-steps", type=int, default=3,
                    help="timed repeats of the BZ_States (fullstates) leg of the default run (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel over MIB MiB (profiler counter calibration)")
    args = ap.parse_args()
    args.timing_steps = max(1, args.timing_steps)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 10)

    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the k-point batch would shard across GPUs by block "
                                   "partition, one process per GPU, but no multi-GPU run is implemented",
                          "n_gpus": args.gpus}))
        return 2

    from ultimateepm_amd import api
    from ultimateepm_amd.material import load

    if args.form == "dims":
        ctx = api.Context(0)
        only = {int(x) for x in args.dims.split(",") if x.strip()}
        rows = dims_leg(ctx, only=only, vectors=args.dims_vectors)
        head = next((r for r in rows if r["dim"] == 181 and r["mode"] == "mixed"), None)
        what = "eigenvalues + eigenvectors" if args.dims_vectors else "eigenvalues only"
        line = {"metric": f"k-pts/s by basis dimension (16 bands, {what}); value: Si n=12 (dim 181) precision form",
                "value": head["kpts_per_s"] if head else None, "unit": "k-pts/s", "n_gpus": 1, "higher_is_better": True,
                "data": "synthetic", "dtype": "c64+f64",
                "config": {"workload": "Si n=6..14 and Ge+SOC n=6..14, non-local, nb 16, seeded BZ k-points",
                           "form": "dims"},
                "dims": rows}
        print(json.dumps(line), flush=True)
        ctx.close()
        return 0
    if args.form == "batching":
        ctx = api.Context(0)
        bt = batching_leg(ctx, args.batch_nk)
        line = {"metric": "k-pts/s batching form: 28 (material, cutoff, SOC) plans in one graph launch",
                "value": bt["mixed"]["kpts_per_s"], "unit": "k-pts/s", "n_gpus": 1, "higher_is_better": True,
                "data": "synthetic", "dtype": "c64+f64",
                "config": {"workload": "{Si,Ge,GaAs,InP} x n{6,8,10,12} x SOC{off,on} (no InP SOC), "
                                       f"{args.batch_nk} seeded BZ k-points each, nb 8 (16 with SOC)",
                           "form": "batching"},
                "batching": bt}
        print(json.dumps(line), flush=True)
        ctx.close()
        return 0
    if args.form == "serving":
        ctx = api.Context(0)
        sv = serving_leg(ctx, args.requests)
        line = {"metric": "serving p50/p99 latency (precision form), Poisson 200 req/s", "value": sv["mixed"]["latency_p99_ms"],
                "unit": "ms p99", "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
                "dtype": "c64+f64", "config": {"workload": sv.pop("workload"), "form": "serving"}, "serving": sv}
        print(json.dumps(line), flush=True)
        ctx.close()
        return 0

    mat = load("Si")
    G = api.basis_vectors(args.shells)
    N = len(G)
    nb = args.nb
    ctx = api.Context(0)
    ks = api.bz_sample(SEED, args.nk)
    dk = api.DeviceArray.from_host(ctx, ks)
    stream = ctx.stream()

    def run(mode: str):
        flags = api.NONLOCAL | api.VECTORS | (api.MIXED if mode == "mixed" else 0)
        plan = api.Plan(ctx, mat, G, flags, nb)
        plan.set_config(lanes=1)
        one = plan.workspace_bytes(args.nk)
        plan.set_config(lanes=args.lanes)
        lanes = plan.workspace_bytes(args.nk) // one  # chunks in flight
        dev = api.DeviceArray(ctx, (args.nk, nb))
        dvec = api.DeviceArray(ctx, (args.nk, nb, plan.dim, 2))

        def step():
            plan.solve_device(dk.ptr, args.nk, dev.ptr, dvec.ptr, 0, stream.handle)

        for _ in range(args.warmup):
            step()
        stream.synchronize()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):  # stage timing off: no host synchronisation inside a step
            step()
        stream.synchronize()
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        evals = dev.numpy()
        # the stage split, from HIP events around each stage launch, in separate steps
        plan.set_timing(True)
        stage_ms = {}
        for _ in range(args.timing_steps):
            step()
            for name, (ms, cnt) in plan.last_timing().items():
                acc = stage_ms.setdefault(name, [0.0, 0])
                acc[0] += ms
                acc[1] += cnt
        plan.set_timing(False)
        plan.close()
        dev.free()
        dvec.free()
        return {"elapsed": elapsed, "stage_ms": stage_ms, "evals": evals, "dim": plan.dim, "lanes": lanes,
                "timing_steps": args.timing_steps}

    res = {}
    for m in args.modes.split(","):
        log(f"headline {m}: {args.warmup} warm-up + {args.steps} timed steps of {args.nk} k-points")
        res[m] = run(m)
        log(f"headline {m}: {args.nk / (res[m]['elapsed'] / args.steps):.0f} k-pts/s")
    head = "mixed" if "mixed" in res else "fp64"
    r = res[head]
    ms_per_step = r["elapsed"] / args.steps * 1e3
    value = args.nk / (r["elapsed"] / args.steps)
    stage_ms = r["stage_ms"]
    ts = r["timing_steps"]
    dominant = max(stage_ms, key=lambda k: stage_ms[k][0])
    dom_ms, dom_launches = stage_ms[dominant]
    if dominant.startswith("tridiag"):
        flops = tridiag_flops(r["dim"]) * args.nk * ts
        peak = FP32_MFMA_PEAK_TFLOPS if head == "mixed" else FP64_PEAK_TFLOPS
    else:
        flops = 0.0
        peak = FP64_PEAK_TFLOPS
    achieved = flops / (dom_ms * 1e-3) / 1e12
    # the reduction's flops of a whole step over that step's wall time (timing off):
    # the chip-level rate with both lanes in flight and every other stage included
    step_s = r["elapsed"] / args.steps
    chip = tridiag_flops(r["dim"]) * args.nk / step_s / 1e12
    roofline = {
        "kernel": dominant,
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4),
        "traffic": None,
        "avg_launch_ms": round(dom_ms / max(dom_launches, 1), 4),
        "k_per_launch": round(args.nk * ts / max(dom_launches, 1), 1),
        "flop_per_k": tridiag_flops(r["dim"]),
        "datapath": DATAPATH.get((head, dominant.split("<")[0]), "not listed for this kernel"),
        "stage_ms_per_step": {k: round(v[0] / ts, 3) for k, v in stage_ms.items()},
        "launches_in_flight": r["lanes"],
        "chip_reduction_TFLOPs_over_step": round(chip, 3),
        "chip_frac_over_step": round(chip / peak, 4),
        "stage_sum_over_lanes_ms": round(sum(v[0] for v in stage_ms.values()) / ts / r["lanes"], 3),
    }
    bpk = measured_traffic(head, dominant, N)
    if bpk is not None:
        roofline["traffic"] = round(bpk * roofline["k_per_launch"])
        roofline["traffic_unit"] = "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    extra = {}
    for m, rr in res.items():
        extra[f"{m}_kpts_per_s"] = round(args.nk / (rr["elapsed"] / args.steps), 1)
        extra[f"{m}_stage_ms_per_step"] = {k: round(v[0] / rr["timing_steps"], 3) for k, v in rr["stage_ms"].items()}
    if "mixed" in res and "fp64" in res:
        d = res["mixed"]["evals"] - res["fp64"]["evals"]
        extra["mixed_vs_fp64_max_meV"] = float(np.abs(d).max() * 1e3)
        extra["mixed_vs_fp64_rms_meV"] = float(np.sqrt(np.mean(d ** 2)) * 1e3)
        # the precision form's bar: every eigenvalue within 0.1 meV of the FP64 form
        extra["mixed_within_0p1_meV"] = bool(np.abs(d).max() <= 1e-4)
        extra["mixed_over_fp64"] = round(extra["mixed_kpts_per_s"] / extra["fp64_kpts_per_s"], 3)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "k-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"Si non-local (materials.yaml), n={args.shells} shells (N={N}), {args.nk} seeded "
                        f"first-BZ k-points, lowest {nb} eigenvalues + eigenvectors",
            "mode": head,
            "nk": args.nk,
            "N": N,
            "nb": nb,
            **extra,
        },
        "roofline": roofline,
    }
    if args.ge_soc_nk > 0:
        log("Ge + SOC (dim 274) leg")
        out["ge_soc_n10"] = ge_soc_leg(ctx, args.ge_soc_nk)
    if args.batch_leg_nk > 0:
        log("batching leg (28 plans, tuned, one graph launch per form)")
        out["batching"] = batching_leg(ctx, args.batch_leg_nk)
    if args.serving_requests > 0:
        log("serving leg")
        sv = serving_leg(ctx, args.serving_requests)
        out["serving"] = {
            "workload": sv["workload"], "requests": sv["requests"],
            "mixed_p50_ms": sv["mixed"]["latency_p50_ms"], "mixed_p99_ms": sv["mixed"]["latency_p99_ms"],
            "fp64_p50_ms": sv["fp64"]["latency_p50_ms"], "fp64_p99_ms": sv["fp64"]["latency_p99_ms"],
            "mixed_kpts_per_s": sv["mixed"]["kpts_per_s"], "fp64_kpts_per_s": sv["fp64"]["kpts_per_s"],
            "mixed_vs_fp64_max_meV": sv["mixed_vs_fp64_max_meV"],
            "hbm_budget_gib": sv["hbm_budget_gib"], "max_inflight": sv["max_inflight"]}
        for mode in ("mixed", "fp64"):
            sat, half = sv[f"{mode}_saturation"], sv[f"{mode}_half_load"]
            out["serving"][f"{mode}_saturation_kpts_per_s"] = sat["kpts_per_s"]
            out["serving"][f"{mode}_saturation_req_per_s"] = round(sat["requests"] / sat["elapsed_s"], 2)
            out["serving"][f"{mode}_half_load_req_per_s"] = half["offered_rate_req_s"]
            out["serving"][f"{mode}_half_load_p50_ms"] = half["latency_p50_ms"]
            out["serving"][f"{mode}_half_load_p99_ms"] = half["latency_p99_ms"]
    if args.mesh_steps > 0:
        log("mesh leg")
        out["mesh"] = mesh_leg(ctx, args.mesh_steps, cpu=not args.no_cpu_baseline)
    if args.dielectric_q > 0:
        log("dielectric leg")
        out["dielectric"] = dielectric_leg(ctx, args.dielectric_q, cpu=not args.no_cpu_baseline)
    if args.states_steps > 0:
        log("BZ_States leg")
        out["bz_states"] = states_leg(ctx, args.states_steps, cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, all cores of the job's share, then one core)")
        out["cpu_baseline"] = cpu_baseline(mat, G, nb, ks, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    stream.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
