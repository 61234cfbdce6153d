"""Pins of the oracle's plane constraints (FDTD_PlaneConstraint; the
classification at MAC_Grid.cpp:3511-3577 and the perfectly reflecting ghost
cells at MAC_Grid.cpp:1198-1204), by what geometry and the wave equation fix:

* a cell is solid exactly when the constraint holds its centre: the solid
  count is the number of cell layers with (x_d - H) s < 0 times the cells of
  a layer, for every axis and sign, and with the plane 0.3 h past a face
  (where a seven-sample test would take one layer more);
* a uniform field stays exactly uniform next to the plane (a rigid wall adds
  no gradient; a pressure-release wall would pull the layer above it down at
  the first step);
* a pulse above a plane at a cell face equals, over its direct and
  reflected arrivals and to within the wall's placement error (the
  non-uniform ghost stencil puts the effective wall h/8 off the face), the
  same pulse plus its mirror image in the same grid without the plane
  (measured 1.9% relative L2; bound 5%).  Planted in the oracle on the CPU,
  a pressure-release ghost value (0 instead of the neighbour's pressure)
  fails both this and the uniform-field pin, and a seven-sample test for
  constraint cells fails the count at 0.3 h past a face."""
import math

import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import SceneBuffers
from oracle.oracle import Oracle


def _free(n=40, h=0.005, pml=6):
    return scenes.base_settings((n, n, n), h, pml)


@pytest.mark.parametrize("direction", [0, 1, 2])
@pytest.mark.parametrize("sign", [1, -1])
@pytest.mark.parametrize("frac", [0.0, 0.3])
def test_solid_cells_are_those_whose_centre_lies_behind_the_plane(direction, sign, frac):
    n, h = 40, 0.005
    s = _free(n, h)
    mb = s["min_bound"][direction]
    height = mb + (17 + frac) * h
    sb = SceneBuffers(s, listeners=np.zeros((1, 3)), constraints=[dict(direction=direction, sign=sign, height=height)])
    c = Oracle(sb, reclassify_each_step=False).counts()
    centres = mb + (np.arange(n) + 0.5) * h
    layers = int(np.sum((centres - height) * sign < 0))
    assert c["n_solid"] == layers * n * n
    assert c["n_ghost"] == n * n   # the last solid layer's cells, one fluid neighbour each


def test_uniform_field_stays_uniform_next_to_the_plane():
    n, h, pml = 40, 0.005, 6
    s = _free(n, h, pml)
    height = s["min_bound"][1] + (pml + 8) * h      # a cell face inside the interior
    sb = SceneBuffers(s, listeners=np.zeros((1, 3)), constraints=[dict(direction=1, sign=1, height=height)])
    o = Oracle(sb, reclassify_each_step=False)
    p = np.ones(sb.n_cells)
    o.set_state(p, p)
    steps = 3
    o.step(steps, record=False)
    q = o.pressure()
    # the PML's split state starts at rest and pulls its cells to zero; cells
    # further than `steps` + 1 cells from it have not seen that yet
    a, b = pml + steps + 1, n - pml - steps - 1
    j0 = pml + 8                                       # the first fluid layer
    inner = q[a:b, j0:b, a:b]
    assert np.all(inner == 1.0), np.abs(inner - 1.0).max()


def _pulse_scene(with_plane, image, n=64, h=0.005, pml=10):
    s = _free(n, h, pml)
    height = s["min_bound"][1] + (pml + 14) * h        # a cell face
    src = np.array([0.0013, height + 8 * h, -0.0021])
    ps = []
    for y in ([src[1], 2 * height - src[1]] if image else [src[1]]):
        q = scenes.gaussian_pressure_source((src[0], y, src[2]), 1e-4, offset=4e-4)
        q["width_space"] = 2 * h
        q["normalize_constant"] = 1.0 / (math.sqrt(2 * math.pi) * 2 * h) ** 3
        ps.append(q)
    lst = np.array([[0.03, height + 3.3 * h, 0.01], [-0.02, height + 12.2 * h, 0.04], [0.05, height + 20.5 * h, -0.03],
                    [-0.04, height + 6.1 * h, -0.045]])
    cons = [dict(direction=1, sign=1, height=height)] if with_plane else []
    return SceneBuffers(s, pressure_sources=ps, listeners=lst, constraints=cons)


def test_plane_reflects_like_a_mirror_image_source():
    # the direct and the reflected arrivals (200 steps): later, a mode
    # trapped where the plane enters the PML (whose solid cells act as a
    # pressure-release wall) leaks back slowly and decays (peak 7e-3 of the
    # pulse near step 2000, 2e-4 by step 5500), with no mirror counterpart
    steps = 200
    wall = Oracle(_pulse_scene(True, False), reclassify_each_step=False).step(steps)
    mirror = Oracle(_pulse_scene(False, True), reclassify_each_step=False).step(steps)
    alone = Oracle(_pulse_scene(False, False), reclassify_each_step=False).step(steps)
    err = np.linalg.norm(wall - mirror) / np.linalg.norm(mirror)
    assert err < 0.05, err
    # the image matters: without it the series differ by far more
    assert np.linalg.norm(wall - alone) / np.linalg.norm(mirror) > 0.3
