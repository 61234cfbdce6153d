"""Pins of the oracle's staggered (MAC) leapfrog, PML_WaveSolver::stepLeapfrog
(src/wavesolver/PML_WaveSolver.cpp:710-773) with MAC_Grid::PML_velocityUpdate /
PML_pressureUpdate / UpdatePMLPressure (src/wavesolver/MAC_Grid.cpp:301-370,
948-1129) and the time-integrated Gaussian source of the non-collocated build
(src/wavesolver/GaussianPressureSource.cpp:24-27).

Away from the PML the velocities can be eliminated exactly:
    p^{n+1} - 2 p^n + p^{n-1} = (c dt/h)^2 L(p^n) + 3 dt (S_{n+1/2} - S_{n-1/2}),
with L the 7-point Laplacian and S the source evaluated at t_n + dt/2 (added to
each of the three split pressures).  The test recomputes the right-hand side
independently and compares it with the oracle's fields, so a wrong sign, a
missing face shift, a single instead of a triple source injection or a wrong
time factor fails.  A second test checks the split-field PML against a
larger domain."""
import math

import numpy as np
from scipy.special import erf

from acoustics_amd import scenes
from acoustics_amd.capi import SceneBuffers, FDTD_SCHEME_STAGGERED
from oracle import oracle as O

H = 0.005


def _source():
    src = np.array([0.0011, -0.0023, 0.0017])
    ps = scenes.gaussian_pressure_source(src, 2e-4)
    ps["width_space"] = 2 * H
    ps["normalize_constant"] = 1.0 / (math.sqrt(2 * math.pi) * 2 * H) ** 3
    return ps


def _settings(n, w):
    s = scenes.base_settings((n, n, n), H, w)
    s["scheme"] = FDTD_SCHEME_STAGGERED
    return s


def _scene(n, w, listeners):
    return SceneBuffers(_settings(n, w), pressure_sources=[_source()], listeners=listeners)


def _source_field(settings, ps, t):
    """S(x, t) at every cell centre: spatial Gaussian inside the +-2 w_s box,
    times erf((t - t0)/(sqrt2 wt)) - erf(-t0/(sqrt2 wt)), times K."""
    n = settings["divisions"]
    mn = np.array(settings["min_bound"])
    axes = [mn[d] + (0.5 + np.arange(n[d])) * H for d in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    p = np.array(ps["position"])
    ws, wt, t0, K = ps["width_space"], ps["width_time"], ps["offset_time"], ps["normalize_constant"]
    r2 = (X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2
    inside = (np.abs(X - p[0]) < 2 * ws) & (np.abs(Y - p[1]) < 2 * ws) & (np.abs(Z - p[2]) < 2 * ws)
    tf = erf((t - t0) / (math.sqrt(2) * wt)) - erf(-t0 / (math.sqrt(2) * wt))
    return np.where(inside, np.exp(-r2 / (2 * ws * ws)), 0.0) * tf * K


def test_staggered_interior_is_the_second_order_wave_equation():
    n, w = 28, 5
    s = _settings(n, w)
    sb = SceneBuffers(s, pressure_sources=[_source()], listeners=np.array([[0.02, 0.0, 0.0]]))
    dt, c = s["time_step"], s["sound_speed"]
    lam2 = (c * dt / H) ** 2
    o = O.Oracle(sb, reclassify_each_step=False)
    P = [o.pressure().reshape(n, n, n).copy()]
    for _ in range(70):
        o.step(1)
        P.append(o.pressure().reshape(n, n, n).copy())
    o.close()
    m = w + 2   # cells whose faces and neighbours all lie outside the PML
    sl = (slice(m, n - m),) * 3
    scale = max(np.abs(p).max() for p in P)
    assert scale > 0
    worst = 0.0
    for k in range(1, len(P) - 1):
        p = P[k]
        lap = (np.roll(p, 1, 0) + np.roll(p, -1, 0) + np.roll(p, 1, 1) + np.roll(p, -1, 1)
               + np.roll(p, 1, 2) + np.roll(p, -1, 2) - 6 * p)
        src = 3 * dt * (_source_field(s, _source(), (k + 0.5) * dt) - _source_field(s, _source(), (k - 0.5) * dt))
        lhs = P[k + 1] - 2 * p + P[k - 1]
        rhs = lam2 * lap + src
        worst = max(worst, np.abs(lhs - rhs)[sl].max())
    assert worst < 1e-10 * scale, (worst, scale)


def test_staggered_pml_absorbs_outgoing_wave():
    """A listener run in a 48^3 domain with an 8-cell PML matches the same run in
    a 96^3 domain, whose walls are too far for any reflection to arrive in time."""
    lst = np.array([[0.05, 0.0, 0.0], [0.03, 0.04, -0.04], [-0.05, -0.05, 0.0]])
    steps = 300
    a = O.Oracle(_scene(48, 8, lst), reclassify_each_step=False).step(steps)
    b = O.Oracle(_scene(96, 8, lst), reclassify_each_step=False).step(steps)
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert err < 0.03, err
