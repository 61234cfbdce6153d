"""Plane constraints on the device against the FP64 oracle: floors along each
axis and a ceiling, crossing the PML (whose solid cells the PML tiles hold
at zero), with an object resting above the plane and a pressure pulse;
rasterized and fully-coupled ghost cells.  Tolerance: listener relative L2
error <= 1e-4; every tile plan gives the same series bit for bit."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = {
    "floor_y_sphere": dict(direction=1, sign=1),
    "floor_z_sphere": dict(direction=2, sign=1, frac=0.3),
    "ceiling_x_ellipsoid": dict(direction=0, sign=-1, obj="ellipsoid"),
    "floor_at_pml_edge": dict(direction=1, sign=1, layers=0),
    "floor_in_pml_no_object": dict(direction=1, sign=1, layers=-3, obj=None),
}


@pytest.mark.parametrize("boundary", [FDTD_BOUNDARY_RASTERIZE, FDTD_BOUNDARY_FULLY_COUPLED])
@pytest.mark.parametrize("case", list(CASES))
def test_constraint_scene_matches_oracle(hip_lib, case, boundary):
    sb, steps = scenes.floor_scene(boundary=boundary, **CASES[case])
    print(f"case: {case} boundary={boundary} divisions={sb.divisions} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb, steps_per_graph=16) as s:
        s.step(steps)
        got = s.series()
    err = rel_l2(got, ref)
    print(f"{case}: listener rel L2 {err:.3e}")
    assert np.abs(ref).max() > 0
    assert err <= TOL


def test_constraint_tile_plans_agree(hip_lib):
    sb, steps = scenes.floor_scene(direction=2, sign=1, frac=0.3)
    outs = []
    for plan in (0, 3, 4, 8):
        with Solver(sb, tile_config=plan) as s:
            s.step(150)
            outs.append(s.series())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
