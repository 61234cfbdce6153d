"""Harmonic and acceleration-noise vibrational sources on the device against
the FP64 oracle (HarmonicVibrationalSource.cpp:5-28,
AccelerationNoiseVibrationalSource.cpp:63-90): rotated ellipsoids with
decayed, time-windowed pulsations and off-centre impacts with overlapping
contacts, alone and together with a Gaussian impulse, with rasterised and
fully-coupled ghost cells.  Tolerance: listener relative L2 error <= 1e-4.
A scene with acceleration noise (four amplitude channels per object) shares
batched and served launches with one-channel scenes and gives its
single-scene series bit for bit."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE
from acoustics_amd.solver import Batch, Server, Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = {
    "harmonic": dict(acc_noise=False),
    "acc_noise": dict(harmonic=False),
    "all_kinds": dict(gaussian=True, seed=18, n_obj=3),
    "chirp": dict(harmonic=False, acc_noise=False, chirp=True, seed=26),
}


@pytest.mark.parametrize("boundary", [FDTD_BOUNDARY_RASTERIZE, FDTD_BOUNDARY_FULLY_COUPLED])
@pytest.mark.parametrize("case", list(CASES))
def test_vibration_sources_match_oracle(hip_lib, case, boundary):
    sb, steps = scenes.vibration_scene(boundary=boundary, **CASES[case])
    print(f"case: {case} boundary={boundary} divisions={sb.divisions} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb, steps_per_graph=16) as s:
        s.step(steps)
        got = s.series()
    err = rel_l2(got, ref)
    print(f"{case}: listener rel L2 {err:.3e}")
    assert np.abs(ref).max() > 0
    assert err <= TOL


def _alone(sb, steps):
    with Solver(sb) as s:
        s.step(steps)
        return s.series()


def test_acc_noise_scene_in_batch_and_server(hip_lib):
    an, st_an = scenes.vibration_scene(seed=19, n=40, pml=6, steps=200, harmonic=False)
    hm, st_hm = scenes.vibration_scene(seed=20, n=36, pml=6, steps=150, acc_noise=False)
    free, st_free = scenes.config1(seed=21, n=32, steps=120)
    sbs, steps = [free, an, hm], [st_free, st_an, st_hm]
    print("batch:", [sb.divisions for sb in sbs], steps, flush=True)
    with Batch(sbs, steps, steps_per_graph=16) as b:
        b.run()
        got = [b.series(i) for i in range(len(sbs))]
    with Server(steps_per_block=16) as srv:
        ids = [srv.submit(sb, st) for sb, st in zip(sbs, steps)]
        for rid in ids:
            assert srv.wait(rid, 120000) == 0
        served = [srv.take(rid)[1] for rid in ids]
    for i, sb in enumerate(sbs):
        ref = _alone(sb, steps[i])
        assert np.array_equal(got[i], ref), i
        assert np.array_equal(served[i], ref), i
