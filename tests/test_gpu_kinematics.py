"""Objects animated by kinematics-file keyframes on the device against the
FP64 oracle (FDTD_RigidObject_Blender_Animator, FDTD_RigidObject_Animator.cpp:
78-160): an ellipsoid swinging on a curved, accelerating path with a varying
turn, and a spinning ellipsoid's constant motion written as keyframes (which
must also agree with the same motion given as velocities).  Tolerance:
listener relative L2 error <= 1e-4; a keyframed scene in a batch gives its
single-scene series bit for bit."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.solver import Batch, Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _run(sb, steps, **kw):
    with Solver(sb, **kw) as s:
        s.step(steps)
        return s.series()


def test_keyframed_swing_matches_oracle(hip_lib):
    sb, steps = scenes.keyframed_object()
    print(f"keyframed swing: divisions={sb.divisions} steps={steps}", flush=True)
    ref = Oracle(sb).step(steps)
    got = _run(sb, steps, steps_per_graph=16)
    err = rel_l2(got, ref)
    print(f"listener rel L2 {err:.3e}")
    assert np.abs(ref).max() > 0
    assert err <= TOL


@pytest.mark.parametrize("rate", [2000.0, 20000.0])
def test_keyframes_of_constant_motion(hip_lib, rate):
    sb, steps = scenes.moving_object(seed=12, kind="ellipsoid", spin=900.0, speed=15.0, steps=300)
    kb = scenes.as_keyframes(sb, frame_rate=rate)
    print(f"keyframes at {rate} frames/s of a spinning ellipsoid: steps={steps}", flush=True)
    ref = Oracle(kb).step(steps)
    got = _run(kb, steps)
    const = _run(sb, steps)
    assert rel_l2(got, ref) <= TOL
    assert rel_l2(got, const) <= TOL


def test_keyframed_scene_in_batch(hip_lib):
    kb, st_k = scenes.keyframed_object(seed=23, n=52, steps=240)
    free, st_f = scenes.config1(seed=24, n=32, steps=150)
    sbs, steps = [free, kb], [st_f, st_k]
    with Batch(sbs, steps, steps_per_graph=16) as b:
        b.run()
        got = [b.series(i) for i in range(len(sbs))]
    for i, sb in enumerate(sbs):
        assert np.array_equal(got[i], _run(sb, steps[i])), i
