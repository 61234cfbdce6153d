"""Writes the golden listener series of small seeded scenes with the FP64
oracle (the checker).  Run from the repository root:
    python tests/golden/make_golden.py
Each fixture is an .npz of plain arrays (no pickles): the scene is rebuilt
from acoustics_amd.scenes with the stored arguments, and `series` holds the
oracle's FP64 [steps][listeners] output."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from acoustics_amd import scenes  # noqa: E402
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

FIXTURES = {
    "free_field_32": ("config1", dict(seed=101, n=32), 160),
    "sphere_raster_40": ("config2", dict(seed=102, n=40, radius=0.025, shell=0.06, pml=6, n_listeners=12), 180),
    "sphere_coupled_40": ("config2", dict(seed=103, n=40, radius=0.025, shell=0.06, pml=6, n_listeners=12,
                                          boundary=FDTD_BOUNDARY_FULLY_COUPLED), 150),
    "ellipsoids_raster_48": ("config3", dict(seed=104, n=48, h=0.005, pml=6, n_obj=2, n_listeners=12,
                                             axis_lo=0.01, axis_hi=0.02), 700),
}


def build(name):
    fn, kw, steps = FIXTURES[name]
    sb, _ = getattr(scenes, fn)(**kw)
    return sb, steps


def main():
    for name in FIXTURES:
        sb, steps = build(name)
        series = Oracle(sb, reclassify_each_step=True).step(steps)
        fn, kw, _ = FIXTURES[name]
        np.savez(os.path.join(HERE, name + ".npz"), series=series,
                 meta=np.array(json.dumps({"builder": fn, "kwargs": kw, "steps": steps})))
        print(name, series.shape, float(np.abs(series).max()))


if __name__ == "__main__":
    main()
