"""Pins of the oracle's air-viscosity bulk update (MAC_Grid.cpp:926-930):

    p^{n+1} = 2 p^n - p^{n-1} - c alpha dt L^{n-1} + ((c dt)^2 + c alpha dt) L^n,
    L = (sum of the six neighbours - 6 p) / h^2,

switched on by air_viscosity_alpha > 0 (ImpulseResponseParser.cpp:540-549).

The checks do not reuse the oracle's code: a compact pulse in the middle of a
free-field grid is stepped by a separate NumPy stencil on the whole array (the
pulse stays two cells away from the PML for the steps taken, so the PML never
sees it), and the dissipation check is a property of the scheme: viscosity
removes energy, more of it for a larger alpha, and alpha -> 0 gives the
inviscid update exactly."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import SceneBuffers
from oracle.oracle import Oracle

DIV = (26, 24, 28)
PML = 6


def _scene(alpha, div=DIV):
    s = scenes.base_settings(div, 0.005, PML)
    s["air_viscosity_alpha"] = alpha
    return SceneBuffers(s, listeners=np.zeros((0, 3)))


def _pulse(div=DIV):
    """Smooth bump of radius 3 cells at the grid centre (zero elsewhere)."""
    g = np.indices(div).astype(float)
    c = [(n - 1) / 2.0 for n in div]
    r2 = sum((g[d] - c[d]) ** 2 for d in range(3))
    f = np.where(r2 < 9.0, (1.0 - r2 / 9.0) ** 2, 0.0)
    return f


def _lap(p, h):
    out = np.zeros_like(p)
    out[1:-1, 1:-1, 1:-1] = (p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1] + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1] +
                             p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2] - 6.0 * p[1:-1, 1:-1, 1:-1]) / (h * h)
    return out


def _run(alpha, steps, div=DIV):
    sb = _scene(alpha, div)
    o = Oracle(sb, reclassify_each_step=False)
    o.set_state(np.zeros(sb.n_cells), _pulse(div).reshape(-1))
    out = []
    for _ in range(steps):
        o.step(1, record=False)
        out.append(o.pressure().copy())
    o.close()
    return sb, out


def _brute(alpha, steps):
    s = scenes.base_settings(DIV, 0.005, PML)
    h, dt, c = s["cell_size"], s["time_step"], s["sound_speed"]
    kcalp, c2k2 = c * alpha * dt, (c * dt) ** 2
    p_prev, p = np.zeros(DIV), _pulse()
    lap_prev = np.zeros(DIV)
    out = []
    for _ in range(steps):
        lap = _lap(p, h)
        p_next = 2.0 * p - p_prev - kcalp * lap_prev + (c2k2 + kcalp) * lap
        p_prev, p, lap_prev = p, p_next, lap
        out.append(p.copy())
    return out


def _energy(pn1, pn, lam2):
    """Discrete energy of the leapfrog between levels n and n+1,
    ||p^{n+1} - p^n||^2 + lam2 <D p^n, D p^{n+1}> (D: differences across every
    face).  Multiplying p^{n+1} - 2 p^n + p^{n-1} = lam2 Lap p^n by
    p^{n+1} - p^{n-1} and summing shows the inviscid scheme keeps it exactly
    while the pulse stays inside the bulk."""
    e = float(np.sum((pn1 - pn) ** 2))
    for d in range(3):
        e += lam2 * float(np.sum(np.diff(pn1, axis=d) * np.diff(pn, axis=d)))
    return e


@pytest.mark.parametrize("alpha", [0.0, 2e-4, 1e-3])
def test_viscous_steps_match_brute_force(alpha):
    _, got = _run(alpha, 3)
    ref = _brute(alpha, 3)
    for q in range(3):
        assert np.max(np.abs(got[q] - ref[q])) <= 1e-12 * np.max(np.abs(ref[q])), (alpha, q)


def test_alpha_zero_is_inviscid_and_viscosity_dissipates():
    s = scenes.base_settings(DIV, 0.005, PML)
    lam2 = (s["sound_speed"] * s["time_step"] / s["cell_size"]) ** 2
    big = (40, 38, 42)   # the pulse (radius 3) is 13 cells from the PML after 8 steps
    runs = {a: [_pulse(big)] + _run(a, 8, big)[1] for a in (0.0, 2e-4, 1e-3)}
    ens = {a: [_energy(r[q + 1], r[q], lam2) for q in range(8)] for a, r in runs.items()}
    e0 = ens[0.0]
    assert max(abs(e - e0[0]) for e in e0) <= 1e-10 * abs(e0[0])          # inviscid: conserved
    for a in (2e-4, 1e-3):
        assert all(ens[a][q + 1] < ens[a][q] for q in range(7)), a          # viscous: decreasing
    assert ens[1e-3][-1] < ens[2e-4][-1] < e0[-1]                          # more alpha, more loss
    # the first step differs from the inviscid one by exactly c alpha dt L^0
    d = runs[2e-4][1] - runs[0.0][1]
    want = s["sound_speed"] * 2e-4 * s["time_step"] * _lap(_pulse(big), s["cell_size"])
    assert np.max(np.abs(d - want)) <= 1e-12 * np.max(np.abs(want))
