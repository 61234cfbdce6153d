"""The seven-sample configuration of object cells (the reference built
without ONLY_SAMPLE_CENTER, Wavesolver_ConstantsAndTypes.h:46: occupancy by
the centre and six points h/4 from it, ghosts h/4 from the solid centre,
p_g = p_nbr - 0.75 h grad_p . n_cell, non-uniform Laplacian weights) on the
device against the FP64 oracle in the same configuration, over the object
paths: rasterized and fully-coupled static objects, a listener next to an
object, a spinning moving object, a floor constraint with an object, and
acceleration noise.  Tolerance: listener relative L2 error <= 1e-4.  The
two configurations must also differ: the switch changes the scene."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED, FDTD_SAMPLE_SEVEN
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = {
    "sphere_raster": lambda: scenes.config2(seed=31, n=48, radius=0.03, shell=0.07, pml=6, n_listeners=12, steps=250),
    "ellipsoids_coupled": lambda: scenes.config3(seed=32, n=56, h=0.005, pml=6, n_obj=3, n_listeners=12, axis_lo=0.01,
                                                 axis_hi=0.025, boundary=FDTD_BOUNDARY_FULLY_COUPLED, steps=250),
    "moving_spin": lambda: scenes.moving_object(seed=33, kind="ellipsoid", spin=600.0, speed=12.0, steps=250),
    "floor_and_sphere": lambda: scenes.floor_scene(seed=34, direction=1, sign=1),
    "acc_noise_coupled": lambda: scenes.vibration_scene(seed=35, harmonic=False, gaussian=True,
                                                        boundary=FDTD_BOUNDARY_FULLY_COUPLED),
}


@pytest.mark.parametrize("case", list(CASES))
def test_seven_sample_configuration_matches_oracle(hip_lib, case):
    sb0, steps = CASES[case]()
    sb = scenes.with_settings(sb0, sampling=FDTD_SAMPLE_SEVEN)
    print(f"case: {case} (seven samples) divisions={sb.divisions} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb, steps_per_graph=16) as s:
        s.step(steps)
        got = s.series()
    err = rel_l2(got, ref)
    print(f"{case}: listener rel L2 {err:.3e}")
    assert err <= TOL
    if case in ("sphere_raster", "ellipsoids_coupled"):
        with Solver(sb0) as s:
            s.step(steps)
            centre = s.series()
        assert rel_l2(centre, got) > 1e-3
