"""Pins of the oracle's kinematics-file animation (FDTD_RigidObject_Blender_
Animator, FDTD_RigidObject_Animator.cpp:78-160; Quaternion::slerp,
Quaternion.hpp:365-420; FDTD_RigidObject::SetRigidBodyTransform,
FDTD_RigidObject.cpp:550-566) by what the mathematics fixes:

* frames sampled from a constant velocity and a constant-rate turn are
  reproduced exactly between frames: linear blending of a linear path is the
  path, and the slerp of two samples of a constant-rate turn is the turn at
  the blended time.  With more than about 7 degrees between frames (the
  slerp branch) the pose matches to 1e-15 and a whole moving-object run to
  1e-15 relative L2; with less (the normalized-lerp branch the reference
  takes above a quaternion dot of 0.9995) to the lerp's own error (1e-8);
* a frame whose quaternion is negated (the same rotation) changes nothing
  (the shorter-arc flip);
* after the last frame the pose holds it.
Planted in the oracle on the CPU, each of these fails a pin: the z-up turn
applied on the left (corr * q) or with +90 degrees, swapped blend weights,
and a slerp without the shorter-arc flip."""
import numpy as np
import pytest

from acoustics_amd import scenes
from oracle.oracle import Oracle, keyframe_pose


def _sampled(w, v, c0, R0, step, n):
    frames, prev = [], None
    for f in range(n):
        t = f * step
        R = scenes.rodrigues(w, np.linalg.norm(w) * t) @ R0
        row = np.array(scenes.file_frame(c0 + v * t, R))
        if prev is not None and np.dot(row[3:], prev[3:]) < 0:
            row[3:] = -row[3:]
        frames.append(row)
        prev = row
    return np.array(frames)


@pytest.mark.parametrize("rate,tol", [(1000.0, 1e-14), (40000.0, 1e-6)])
def test_keyframe_pose_reproduces_constant_motion(rate, tol):
    rng = np.random.default_rng(5)
    R0 = scenes.random_rotation(rng)
    w, v, c0 = np.array([40.0, -80.0, 120.0]), np.array([3.0, -2.0, 5.0]), np.array([0.01, 0.0, -0.02])
    step, n = 1.0 / rate, 24
    frames = _sampled(w, v, c0, R0, step, n)
    for t in np.concatenate([rng.uniform(0, (n - 1) * step, 40), [0.0, 3 * step, (n - 1) * step]]):
        c, R = keyframe_pose(0, step, frames, t)
        assert np.abs(c - (c0 + v * t)).max() < 1e-15
        assert np.abs(R - scenes.rodrigues(w, np.linalg.norm(w) * t) @ R0).max() < tol, t


def test_negated_frame_and_hold_after_the_last_frame():
    rng = np.random.default_rng(6)
    R0 = scenes.random_rotation(rng)
    w, v, c0 = np.array([0.0, 150.0, -60.0]), np.array([-4.0, 1.0, 0.5]), np.zeros(3)
    step, n = 1e-3, 12
    frames = _sampled(w, v, c0, R0, step, n)
    flipped = frames.copy()
    flipped[5, 3:] = -flipped[5, 3:]          # the same rotation
    for t in (4.3e-3, 5.0e-3, 5.6e-3):
        c, R = keyframe_pose(0, step, flipped, t)
        assert np.abs(R - scenes.rodrigues(w, np.linalg.norm(w) * t) @ R0).max() < 1e-14, t
    last = (n - 1) * step
    cl, Rl = keyframe_pose(0, step, frames, last)
    for t in (last + 1e-4, last + 0.05):
        c, R = keyframe_pose(0, step, frames, t)
        assert np.array_equal(c, cl) and np.abs(R - Rl).max() < 1e-15


@pytest.mark.parametrize("rate,tol", [(2000.0, 1e-13), (20000.0, 1e-7)])
def test_keyframed_run_equals_constant_motion_run(rate, tol):
    sb, steps = scenes.moving_object(seed=12, kind="ellipsoid", spin=900.0, speed=15.0, steps=300)
    kb = scenes.as_keyframes(sb, frame_rate=rate)
    a = Oracle(sb).step(steps)
    b = Oracle(kb).step(steps)
    assert np.abs(a).max() > 0
    err = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert err < tol, err
