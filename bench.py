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
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def tridiag_flops(n: int) -> float:
    """Algorithmic real flops of the Hermitian Householder reduction: (16/3) n^3."""
    return 16.0 / 3.0 * n ** 3


def cpu_baseline(mat, G, nb, ks, seconds: float):
    """The oracle (C++ restatement of Compute_parallel, OpenMP) on a bounded sample."""
    from oracle import oracle as O
    from ultimateepm_amd import api

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    flags = api.NONLOCAL | api.VECTORS
    probe = ks[: 2 * threads]
    t0 = time.perf_counter()
    O.solve(mat, G, flags, nb, probe, nthreads=threads)
    rate = len(probe) / (time.perf_counter() - t0)
    m = int(max(2 * threads, min(len(ks), rate * seconds)))
    sample = ks[:m]
    t0 = time.perf_counter()
    O.solve(mat, G, flags, nb, sample, nthreads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(m / dt, 2),
        "unit": "k-pts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {m} k-points of the same seeded BZ set, Si non-local N={len(G)}, {nb} bands + "
                  f"eigenvectors (all eigenvectors, as Eigen computes them), oracle built -O3 "
                  f"-march=x86-64-v3 -ffp-contract=off, {threads} OpenMP threads, {dt:.1f} s",
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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nk", type=int, default=200_000)
    ap.add_argument("--nb", type=int, default=12)
    ap.add_argument("--shells", type=int, default=12)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed", help="comma list of fp64, mixed; the value is mixed's")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel over MIB MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 10)

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the k-point batch shards across GPUs by "
                                   "block partition (see DESIGN.md), but no multi-GPU run is part of it",
                          "n_gpus": args.gpus}))
        return 2

    from ultimateepm_amd import api
    from ultimateepm_amd.material import load

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
        dev = api.DeviceArray(ctx, (args.nk, nb))
        dvec = api.DeviceArray(ctx, (args.nk, nb, plan.dim, 2))

        def step():
            plan.solve_device(dk.ptr, args.nk, dev.ptr, dvec.ptr, 0, stream.handle)

        for _ in range(args.warmup):
            step()
        stream.synchronize()
        plan.set_timing(True)
        stage_ms = {}
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            for name, (ms, cnt) in plan.last_timing().items():
                acc = stage_ms.setdefault(name, [0.0, 0])
                acc[0] += ms
                acc[1] += cnt
        stream.synchronize()
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        evals = dev.numpy()
        plan.close()
        dev.free()
        dvec.free()
        return {"elapsed": elapsed, "stage_ms": stage_ms, "evals": evals, "dim": plan.dim}

    res = {m: run(m) for m in args.modes.split(",")}
    head = "mixed" if "mixed" in res else "fp64"
    r = res[head]
    ms_per_step = r["elapsed"] / args.steps * 1e3
    value = args.nk / (r["elapsed"] / args.steps)
    stage_ms = r["stage_ms"]
    dominant = max(stage_ms, key=lambda k: stage_ms[k][0])
    dom_ms, dom_launches = stage_ms[dominant]
    if dominant == "tridiag_kernel":
        flops = tridiag_flops(r["dim"]) * args.nk * args.steps
        peak = FP32_MFMA_PEAK_TFLOPS if head == "mixed" else FP64_PEAK_TFLOPS
    else:
        flops = 0.0
        peak = FP64_PEAK_TFLOPS
    achieved = flops / (dom_ms * 1e-3) / 1e12
    roofline = {
        "kernel": dominant,
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4),
        "traffic": None,
        "avg_launch_ms": round(dom_ms / max(dom_launches, 1), 4),
        "k_per_launch": round(args.nk * args.steps / max(dom_launches, 1), 1),
        "flop_per_k": tridiag_flops(r["dim"]),
        "stage_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in stage_ms.items()},
    }
    bpk = measured_traffic(head, dominant, N)
    if bpk is not None:
        roofline["traffic"] = round(bpk * roofline["k_per_launch"])
        roofline["traffic_unit"] = "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    extra = {}
    for m, rr in res.items():
        extra[f"{m}_kpts_per_s"] = round(args.nk / (rr["elapsed"] / args.steps), 1)
        extra[f"{m}_stage_ms_per_step"] = {k: round(v[0] / args.steps, 3) for k, v in rr["stage_ms"].items()}
    if "mixed" in res and "fp64" in res:
        d = res["mixed"]["evals"] - res["fp64"]["evals"]
        extra["mixed_vs_fp64_max_meV"] = float(np.abs(d).max() * 1e3)
        extra["mixed_vs_fp64_rms_meV"] = float(np.sqrt(np.mean(d ** 2)) * 1e3)
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
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mat, G, nb, ks, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    stream.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
