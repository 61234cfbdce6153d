"""Pins of the oracle's Engquist absorbing boundary (MAC_Grid.cpp:521-919):
the update of the layer-1 cells of a grid without PML cells (use_mesh = 0,
MAC_Grid.cpp:2457-2473, or a PML width of zero), order 2 on faces and order 1
on edges and corners, with the exterior (layer-0) pressures extrapolated into
p^n for the next step.

The checks do not reuse the boundary formulas:
* absorption: a pulse leaving a 30^3 grid is compared with the same pulse in
  a 90^3 grid whose boundary it does not reach in the steps taken; what the
  boundary sends back into the small grid is a small fraction of what a
  pressure-release wall (p = 0 on layer 0, stepped by a NumPy stencil) sends
  back.  A wrong sign or a dropped term in the face, edge or corner update
  makes the boundary reflect strongly or blow up.
* a uniform field solves the wave equation and every boundary update (the
  coefficients of each sum to one): it stays uniform, and the extrapolated
  exterior pressures equal it.
* mirror symmetry: a pulse at the centre of an odd-sized grid stays symmetric
  under each axis flip, which fails when the low and high sides of an axis
  are treated differently (a wrong normal sign or neighbour index).
* axis symmetry: on a cube the field stays symmetric under axis swaps, which
  fails when a face or an edge takes its tangential axes wrongly.
* stability: after the pulse has left, the field decays instead of growing.
"""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import SceneBuffers
from oracle.oracle import Oracle

H = 0.005


def _pulse(div, radius=3.0):
    g = np.indices(div).astype(float)
    c = [(n - 1) / 2.0 for n in div]
    r2 = sum((g[d] - c[d]) ** 2 for d in range(3))
    return np.where(r2 < radius * radius, (1.0 - r2 / (radius * radius)) ** 2, 0.0)


def _run(div, steps, p0=None, p_last=None, use_mesh=0, keep=None):
    s = scenes.base_settings(div, H, 0)
    s["use_mesh"] = use_mesh
    sb = SceneBuffers(s, listeners=np.zeros((0, 3)))
    o = Oracle(sb, reclassify_each_step=False)
    p0 = _pulse(div) if p0 is None else p0
    p_last = p0 if p_last is None else p_last
    o.set_state(p_last.reshape(-1), p0.reshape(-1))
    out = {}
    for q in range(1, steps + 1):
        o.step(1, record=False)
        if keep is None or q in keep:
            out[q] = o.pressure().reshape(div).copy()
    o.close()
    return out, s


def _lam(s):
    return s["sound_speed"] * s["time_step"] / s["cell_size"]


def _pressure_release(div, steps, lam, keep):
    p = _pulse(div)
    pl = p.copy()
    out = {}
    for q in range(1, steps + 1):
        pn = np.zeros_like(p)
        c = p[1:-1, 1:-1, 1:-1]
        lap = (p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1] + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1] +
               p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2] - 6.0 * c)
        pn[1:-1, 1:-1, 1:-1] = 2.0 * c - pl[1:-1, 1:-1, 1:-1] + lam * lam * lap
        pl, p = p, pn
        if q in keep:
            out[q] = p.copy()
    return out


def test_boundary_absorbs():
    small_div, big_div, steps = (30, 30, 30), (90, 90, 90), 100
    keep = {70, 85, 100}
    small, s = _run(small_div, steps, keep=keep)
    # the pulse (radius 3) needs 44 / lam > 150 steps to reach the big grid's boundary
    big, _ = _run(big_div, steps, keep=keep)
    wall = _pressure_release(small_div, steps, _lam(s), keep)
    for q in sorted(keep):
        free = big[q][32:-32, 32:-32, 32:-32]
        back_abc = np.linalg.norm(small[q][2:-2, 2:-2, 2:-2] - free)
        back_wall = np.linalg.norm(wall[q][2:-2, 2:-2, 2:-2] - free)
        print(q, "reflected: Engquist", back_abc, "pressure release", back_wall)
        assert back_wall > 0.1 * np.linalg.norm(free)       # the wave has reached the boundary
        assert back_abc < 0.15 * back_wall, (q, back_abc, back_wall)
    # before the wave arrives the small grid is the free field (up to the
    # stencil's tail, which spreads a cell per step)
    early, _ = _run(small_div, 20, keep={20})
    early_big, _ = _run(big_div, 20, keep={20})
    e = early[20][1:-1, 1:-1, 1:-1] - early_big[20][31:-31, 31:-31, 31:-31]
    assert np.max(np.abs(e)) <= 1e-6 * np.max(np.abs(early_big[20]))


def test_uniform_field_stays_uniform():
    div = (12, 13, 14)
    u = np.full(div, 0.75)
    out, _ = _run(div, 6, p0=u, p_last=u)
    for q, p in out.items():
        inner = p[1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(inner - 0.75)) <= 1e-12, q
    # the face cells' exterior pressures (written into p^n) equal the field
    # too: the step after them still sees a uniform neighbourhood
    assert np.max(np.abs(out[6][0, 2:-2, 2:-2] - 0.75)) <= 1e-12


@pytest.mark.parametrize("use_mesh", [0, 1])
def test_mirror_symmetry(use_mesh):
    div = (25, 27, 29)
    out, _ = _run(div, 90, use_mesh=use_mesh, keep={45, 90})
    for q, p in out.items():
        inner = p[1:-1, 1:-1, 1:-1]
        scale = np.max(np.abs(inner))
        for ax in range(3):
            assert np.max(np.abs(inner - np.flip(inner, axis=ax))) <= 1e-12 * scale, (q, ax)


def test_axis_symmetry():
    div = (23, 23, 23)
    out, _ = _run(div, 90, keep={40, 90})
    for q, p in out.items():
        inner = p[1:-1, 1:-1, 1:-1]
        scale = np.max(np.abs(inner))
        for perm in [(1, 0, 2), (0, 2, 1), (2, 1, 0), (1, 2, 0)]:
            assert np.max(np.abs(inner - np.transpose(inner, perm))) <= 1e-12 * scale, (q, perm)


def test_field_decays_after_the_pulse_leaves():
    div = (24, 26, 28)
    out, _ = _run(div, 900, keep={30, 300, 600, 900})
    amp = {q: np.max(np.abs(p[1:-1, 1:-1, 1:-1])) for q, p in out.items()}
    print("max |p|", amp)
    assert np.all(np.isfinite(out[900]))
    assert amp[300] < 0.2 * amp[30]
    assert amp[900] < amp[600] < amp[300]
