"""Workloads for rocprofv3 counter runs.

  python -m acoustics_amd.profiling copy --mib 1024
      one plain device-to-device copy kernel of known size (read N bytes,
      write N bytes) per iteration: the FETCH_SIZE / WRITE_SIZE counters of
      this kernel calibrate the counters' units against known bytes.
  python -m acoustics_amd.profiling step --n 256 --steps 20
      the headline 256^3 scene stepped eagerly (one dispatch per kernel per
      step) for per-dispatch counters of the step kernels.
"""
from __future__ import annotations

import argparse
import sys


def copy(mib, iters):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(iters):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy: {iters} x {mib} MiB read + {mib} MiB written", flush=True)


def step(n, steps):
    from . import scenes
    from .solver import Solver
    sb, _ = scenes.config3(seed=3, n=n, h=0.64 / n)
    with Solver(sb) as s:
        s.step(steps)
        s.synchronize()
        inf = s.info()
    print(f"step: {n}^3, {steps} steps, {inf['bytes_per_step']:.0f} algorithmic bytes/step", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "step"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20)
    a = ap.parse_args(argv)
    if a.what == "copy":
        copy(a.mib, a.iters)
    else:
        step(a.n, a.steps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
