"""bench.py's host side on the CPU: the one-GPU refusal and the HBM traffic
figures it reads from the committed rocprofv3 summary (each stage's bytes per
k-point must be the sum of that stage's kernel rows in the summary's table)."""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def _table_bytes_per_k():
    """(form, kernel template name) -> bytes per k-point, from the summary's table rows."""
    rows = {}
    pat = re.compile(r"^\| (fp64|mixed) \| `uepm::([^`]+)` \| [\d,]+ \| [\d,]+ \| ([\d,]+) \|$")
    for line in open(bench.TRAFFIC_SUMMARY):
        m = pat.match(line.strip())
        if m:
            rows[(m.group(1), m.group(2))] = int(m.group(3).replace(",", ""))
    return rows


def test_traffic_records_match_the_kernel_table():
    rows = _table_bytes_per_k()
    assert rows, "no kernel rows parsed from the traffic summary"
    # the reduction stage: every grid of the chain (and, in the precision form, the one-wave tail)
    groups = {
        ("mixed", "tridiag_reg_kernel<c64>"): lambda k: k.startswith("tridiag_reg_kernel") and "float" in k
        or k.startswith("tridiag_wave64_kernel"),
        ("fp64", "tridiag_reg_kernel<c128>"): lambda k: k.startswith("tridiag_reg_kernel") and "double" in k,
        ("mixed", "tri_eig_wave_kernel"): lambda k: k.startswith("tri_eig_wave_kernel"),
        ("fp64", "tri_eig_wave_kernel"): lambda k: k.startswith("tri_eig_wave_kernel"),
        ("mixed", "refine_mfma_kernel"): lambda k: k.startswith("refine_mfma_kernel"),
    }
    for (form, stage), member in groups.items():
        want = sum(b for (f, k), b in rows.items() if f == form and member(k))
        assert want > 0, (form, stage)
        got = bench.measured_traffic(form, stage, 181)
        assert got == want, (form, stage, got, want)
    # a stage the summary does not cover
    assert bench.measured_traffic("mixed", "tridiag_reg_kernel<c64>", 274) is None
