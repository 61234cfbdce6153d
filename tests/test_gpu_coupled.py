"""Fully-coupled ghost cells and one-cell cavities on the device against the
FP64 oracle: the device evaluates the coupled system's static solution
operator (x = M p + m a, built on the host in dependency order of the
coupling graph, with a dense inverse over each cycle), the oracle assembles
and solves the reference's sparse system with BiCGSTAB every step.
Tolerance: listener relative L2 error <= 1e-4."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE, FDTD_OBJ_ELLIPSOID, SceneBuffers
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

from test_oracle_coupled import cavity_scene

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _pair(sb, steps, **kw):
    print(f"case: divisions={sb.divisions} objects={len(sb.objects_list)} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb, **kw) as s:
        s.step(steps)
        return s.series(), ref


def cycle_scene(boundary=FDTD_BOUNDARY_FULLY_COUPLED):
    """A flat ellipsoid (semi-axes 4.0, 3.7 and 0.33 cells) whose coupling
    graph holds a pair of ghosts that reference each other (a 2-cycle)."""
    h = 0.005
    s = scenes.base_settings((40, 40, 40), h, 6, boundary)
    rot = np.array([0.74351049649793788, 0.49935277063641514, 0.44479090824243001, -0.20522105925106013,
                    0.80342012795310513, -0.55892791560244071, -0.6364561716546262, 0.32428831070137565,
                    0.69982900276080673]).reshape(3, 3)
    obj = dict(type=FDTD_OBJ_ELLIPSOID, center=(-0.0062476717570087281, 0.0040060227243257796, -0.0099774904083355757),
               semi_axes=(0.020068750013862154, 0.018534932573807859, 0.0016712742588349033), rotation=rot)
    lst = scenes.fibonacci_shell(np.zeros(3), 0.04, 12)
    return SceneBuffers(s, objects=[obj], impulse_sources=[dict(object=0, amplitude=1e3, width=1e-4, onset=3e-4)],
                        listeners=lst)


def test_coupled_cycle_scene(hip_lib):
    got, ref = _pair(cycle_scene(), 300)
    err = rel_l2(got, ref)
    print("coupled 2-cycle scene rel L2", err)
    assert err <= TOL


@pytest.mark.parametrize("boundary", [FDTD_BOUNDARY_RASTERIZE, FDTD_BOUNDARY_FULLY_COUPLED])
def test_cavity_scene(hip_lib, boundary):
    sb, cav, _ = cavity_scene()
    sb.settings_dict["boundary_handling"] = boundary
    sb = SceneBuffers(sb.settings_dict, sb.objects_list, sb.impulse_list, [], sb.listeners_np)
    got, ref = _pair(sb, 240)
    err = rel_l2(got, ref)
    print("cavity scene rel L2", boundary, err)
    assert err <= TOL


def test_coupled_ellipsoids_all_tile_plans_agree(hip_lib):
    sb, _ = scenes.config3(seed=33, n=56, h=0.004, pml=8, n_obj=3, n_listeners=16, axis_lo=0.01, axis_hi=0.025,
                           boundary=FDTD_BOUNDARY_FULLY_COUPLED)
    outs = []
    for plan in (0, 3, 6):
        with Solver(sb, tile_config=plan) as s:
            s.step(200)
            outs.append(s.series())
    ref = Oracle(sb, reclassify_each_step=False).step(200)
    assert rel_l2(outs[0], ref) <= TOL
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
