"""Writes the golden listener series of the full-size workloads with the FP64
oracle (the checker).  Run from the repository root (about 45 minutes on
eight cores):
    python tests/golden/make_golden_full.py [name ...]
Fixtures go to tests/golden/full/<name>.npz as plain arrays (no pickles):
  series       oracle output [steps][listeners], stored as float32 (the
               parity bar is a relative L2 error of 1e-4, far above float32's
               6e-8 rounding)
  listeners    the scene's listener positions, and
  centers      its object centres, both rebuilt by the test from the stored
               builder arguments and compared, so a changed scene builder
               cannot pass against a stale fixture
  meta         JSON: builder name, keyword arguments, steps, and for batch
               fixtures the scene's index in the batch.
Workloads: the 128^3 sphere (2000 steps, 64 shell listeners) with rasterized
and fully-coupled ghost cells; the 256^3 scene of five ellipsoids (4000
steps, 128 listeners) with both; four scenes of the 64-scene batch
(edges 32-128, 1000-3000 steps)."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "full")
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from acoustics_amd import scenes  # noqa: E402
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED  # noqa: E402

FC = FDTD_BOUNDARY_FULLY_COUPLED
FIXTURES = {
    "sphere128_raster": ("config2", dict(seed=2), None),
    "sphere128_coupled": ("config2", dict(seed=2, boundary=FC), None),
    "ellipsoids256_raster": ("config3", dict(seed=3), None),
    "ellipsoids256_coupled": ("config3", dict(seed=3, boundary=FC), None),
    "batch64_scene0": ("config4", dict(seed=4), 0),
    "batch64_scene1": ("config4", dict(seed=4), 1),
    "batch64_scene2": ("config4", dict(seed=4), 2),
    "batch64_scene3": ("config4", dict(seed=4), 3),
}


def build(name):
    """(SceneBuffers, steps) of a fixture, rebuilt from its builder."""
    fn, kw, idx = FIXTURES[name]
    if idx is None:
        return getattr(scenes, fn)(**kw)
    sc, steps = getattr(scenes, fn)(**kw)
    return sc[idx], steps[idx]


def fingerprint(sb):
    lst = np.asarray(sb.listeners_np, dtype=np.float64).reshape(-1, 3)
    ctr = np.asarray([o["center"] for o in sb.objects_list], dtype=np.float64).reshape(-1, 3)
    return lst, ctr


def main(names):
    from oracle.oracle import Oracle
    os.makedirs(OUT, exist_ok=True)
    for name in names:
        sb, steps = build(name)
        t = time.time()
        series = Oracle(sb, reclassify_each_step=False).step(steps)
        fn, kw, idx = FIXTURES[name]
        lst, ctr = fingerprint(sb)
        meta = {"builder": fn, "kwargs": kw, "steps": int(steps), "index": idx}
        np.savez_compressed(os.path.join(OUT, name + ".npz"), series=series.astype(np.float32), listeners=lst,
                            centers=ctr, meta=np.array(json.dumps(meta)))
        print(f"{name}: {series.shape} max {float(np.abs(series).max()):.3e} in {time.time() - t:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(FIXTURES))
