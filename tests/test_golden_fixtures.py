"""Committed golden series (tests/golden/*.npz, made by make_golden.py with
the FP64 oracle): the oracle reproduces them (CPU), and the HIP path matches
them to the 1e-4 relative-L2 contract (GPU)."""
import json
import os

import numpy as np
import pytest

from acoustics_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
NAMES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    sb, _ = getattr(scenes, meta["builder"])(**meta["kwargs"])
    return sb, meta["steps"], z["series"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    from oracle.oracle import Oracle
    sb, steps, series = load(name)
    got = Oracle(sb, reclassify_each_step=False).step(steps)
    assert np.allclose(got, series, rtol=0, atol=1e-12 * max(1.0, np.abs(series).max()))


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_matches_golden(hip_lib, name):
    from acoustics_amd.solver import Solver, rel_l2
    sb, steps, series = load(name)
    print("golden", name, sb.divisions, steps, flush=True)
    with Solver(sb, steps_per_graph=24) as s:
        s.step(steps)
        got = s.series()
    assert rel_l2(got, series) <= 1e-4
