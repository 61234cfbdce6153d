"""Workloads for rocprofv3 runs, and the summaries kept under profiles/.

  python -m acoustics_amd.profiling copy --mib 1024
      one plain device-to-device copy of known size (read N bytes, write N
      bytes) per iteration: the FETCH_SIZE / WRITE_SIZE counters of this
      kernel calibrate the counters' units against known bytes.
  python -m acoustics_amd.profiling step --n 256 --steps 20 [--tile T] [--boundary coupled]
      the headline 256^3 scene stepped eagerly (one dispatch per kernel per
      step) for per-dispatch counters of the step kernels; with coupled
      ghost cells, the same scene with the fully-coupled boundary.
  python -m acoustics_amd.profiling traffic --step-fetch D1 --step-write D2 \\
          --copy-fetch D3 --copy-write D4 --mib 1024 --out profiles/X.json
      reads the counter CSVs of the four --pmc runs above (rocprofv3
      --output-format csv), scales the step kernels' FETCH_SIZE and
      WRITE_SIZE by what the same counters read for the copy of known size,
      and writes HBM bytes per step (sum over the three step kernels) with
      the per-kernel split.
  python -m acoustics_amd.profiling stats --csv D/run_kernel_stats.csv
      prints a markdown table of a --kernel-trace --stats run.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import sys

STEP_KERNELS = ("step_kernel<0>", "step_kernel<1>", "step_kernel<2>", "step_kernel<3>")
KERNEL_ROLE = {"step_kernel<0>": "interior + object-adjacent cells", "step_kernel<1>": "PML slabs",
               "step_kernel<2>": "object-adjacent cells",
               "step_kernel<3>": "one step, all tile classes (PML, interior, object-adjacent) + listener terms",
               "advance_kernel": "block advance (step counter, amplitude ring for the next 64 steps)"}


def copy(mib, iters):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(iters):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy: {iters} x {mib} MiB read + {mib} MiB written", flush=True)


def step(n, steps, tile, scheme="collocated", boundary="rasterize"):
    from . import scenes
    from .capi import FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE
    from .solver import Solver
    if scheme == "staggered":
        sb, _ = scenes.staggered_free_field(seed=7, div=(n, n, n), h=0.64 / n, pml=16, n_listeners=8)
    else:
        bc = FDTD_BOUNDARY_FULLY_COUPLED if boundary == "coupled" else FDTD_BOUNDARY_RASTERIZE
        sb, _ = scenes.config3(seed=3, n=n, h=0.64 / n, boundary=bc)
    with Solver(sb, tile_config=tile) as s:
        # eager (one dispatch per kernel per step, one advance per 64 steps),
        # no timing events between the kernels
        s.step(steps)
        s.synchronize()
        inf = s.info()
    print(f"step: {n}^3, {steps} steps, {inf['bytes_per_step']:.0f} algorithmic bytes/step", flush=True)
    return inf


def _short(name):
    for k in list(STEP_KERNELS) + ["advance_kernel", "motion_kernel", "copyBuffer",
                                   "fillBuffer", "vectorized_elementwise"]:
        if k in name:
            return k
    return name[:40]


def read_counters(d, counter):
    """Mean counter value (KiB) per dispatch, by short kernel name."""
    path = os.path.join(d, "run_counter_collection.csv")
    agg = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                agg[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in agg.items()}


def traffic(a):
    sf, sw = read_counters(a.step_fetch, "FETCH_SIZE"), read_counters(a.step_write, "WRITE_SIZE")
    cf, cw = read_counters(a.copy_fetch, "FETCH_SIZE"), read_counters(a.copy_write, "WRITE_SIZE")
    known_kib = a.mib * 1024.0
    fetch_scale = known_kib / cf["copyBuffer"][0]
    write_scale = known_kib / cw["copyBuffer"][0]
    per = {}
    total = 0.0
    for k in STEP_KERNELS:
        if k not in sf:
            continue
        rd = sf[k][0] * fetch_scale * 1024.0
        wr = sw[k][0] * write_scale * 1024.0
        per[k] = {"role": KERNEL_ROLE[k], "read_bytes": round(rd), "write_bytes": round(wr), "dispatches": sf[k][1]}
        total += rd + wr
    out = {"n": a.n, "hbm_bytes_per_launch": round(total),
           "launch": "one step: the step kernel launches (the fused all-class kernel of the default plan)",
           "fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
           "calibration": f"rocclr copyBuffer of {a.mib} MiB: FETCH_SIZE and WRITE_SIZE scaled so the copy reads "
                          f"and writes exactly {a.mib} MiB",
           "algorithmic_bytes_per_step": a.algorithmic, "per_kernel": per}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def stats(a):
    rows = list(csv.DictReader(open(a.csv)))
    print("| Kernel | Role | Calls | Average (us) | Min (us) | Max (us) | Share |")
    print("|---|---|---|---|---|---|---|")
    for r in rows:
        k = _short(r["Name"])
        print(f"| `{k}` | {KERNEL_ROLE.get(k, '')} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f}% |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "step", "traffic", "stats"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--scheme", choices=["collocated", "staggered"], default="collocated")
    ap.add_argument("--boundary", choices=["rasterize", "coupled"], default="rasterize")
    ap.add_argument("--step-fetch")
    ap.add_argument("--step-write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--algorithmic", type=float, default=444989440.0)
    ap.add_argument("--out", default="profiles/step_kernel_256cube_traffic.json")
    ap.add_argument("--csv")
    a = ap.parse_args(argv)
    if a.what == "copy":
        copy(a.mib, a.iters)
    elif a.what == "step":
        step(a.n, a.steps, a.tile, a.scheme, a.boundary)
    elif a.what == "traffic":
        traffic(a)
    else:
        stats(a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
