"""Load generator and runner for the serving engine (uepm_server_* in the C ABI).

The workload is the request stream of the serving configuration: Poisson
arrivals (seeded, mean rate `rate` requests/s); each request takes a material
uniformly from {Si, Ge, GaAs, InP}, a basis of n in [n_min, n_max] nearest
shells, spin-orbit on or off (off for InP, whose anion period the spin-orbit
model does not support), a log-uniform number of seeded first-BZ k-points in
[1, k_max] and a band count in [8, 24] (doubled rows with spin-orbit are
counted in that band count, as nb counts rows of the spectrum).
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import api
from .material import load

MATERIALS = ("Si", "Ge", "GaAs", "InP")


@dataclasses.dataclass
class RequestSpec:
    arrival_s: float
    material: str
    shells: int
    soc: bool
    nb: int
    k: np.ndarray


def poisson_workload(seed: int, count: int, rate: float = 200.0, n_min: int = 6, n_max: int = 14,
                     k_max: int = 4096, nb_range=(8, 24)) -> list[RequestSpec]:
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        mat = MATERIALS[rng.integers(0, len(MATERIALS))]
        shells = int(rng.integers(n_min, n_max + 1))
        soc = bool(rng.integers(0, 2)) and mat != "InP"
        nk = int(round(math.exp(rng.uniform(0.0, math.log(k_max)))))
        nb = int(rng.integers(nb_range[0], nb_range[1] + 1))
        k = api.bz_sample(int(rng.integers(0, 2 ** 63)), max(nk, 1))
        out.append(RequestSpec(t, mat, shells, soc, nb, k))
    return out


def run(server: api.Server, workload: list[RequestSpec], mixed: bool, nonlocal_: bool = True,
        realtime: bool = True):
    """Submits the workload at its arrival times (or as fast as possible) and
    collects every result.  Returns (results, summary): results[i] = (evals,
    latency_s, queued_s); summary holds p50/p99 latency and k-points/s."""
    mats = {m: load(m) for m in MATERIALS}
    flags_base = (api.NONLOCAL if nonlocal_ else 0) | (api.MIXED if mixed else 0)
    tickets = []
    t0 = time.perf_counter()
    for spec in workload:
        if realtime:
            delay = spec.arrival_s - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        flags = flags_base | (api.SOC if spec.soc else 0)
        tickets.append(server.submit(mats[spec.material], spec.shells, flags, spec.nb, spec.k))
    results = [server.wait(t) for t in tickets]
    elapsed = time.perf_counter() - t0
    lat = np.array([r[1] for r in results])
    nk = sum(len(s.k) for s in workload)
    summary = {
        "requests": len(workload),
        "kpoints": nk,
        "elapsed_s": round(elapsed, 3),
        "kpts_per_s": round(nk / elapsed, 1),
        "latency_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "latency_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "latency_max_ms": round(float(lat.max()) * 1e3, 3),
        "queued_p50_ms": round(float(np.percentile([r[2] for r in results], 50)) * 1e3, 3),
        "offered_rate_req_s": round(len(workload) / workload[-1].arrival_s, 1) if workload else 0.0,
    }
    return results, summary
