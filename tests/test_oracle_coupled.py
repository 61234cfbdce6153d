"""Pins of the oracle's fully-coupled ghost-cell solve and one-cell-cavity
freeze (MAC_Grid.cpp:1131-1441, 924-925, 294-296), by what the mathematics
fixes rather than by restating the oracle's formulas:

* the assembled sparse system's BiCGSTAB solution is the LAPACK solution of
  the same system, with a residual at rounding level;
* on the exact field p = a0 + b (|x - c| - R) around a sphere of radius R
  centred at c, with the Neumann data that field satisfies (dp/dn = b, so
  a_n = -b / rho), the ghost values reproduce the field at the points the
  Laplacian takes them for -- the ghost position, except that a rasterized
  ghost of the centre-sampled configuration sits on the face but stands for
  the solid cell's centre (its value p_nbr - h dp/dn_cell extrapolates a
  whole cell) -- to second order: max error <= 0.06 b h at h = 4 mm
  (measured 0.0005-0.032) and the error falls by at least 0.65x when h
  halves (measured 0.24-0.53x).  Planted in the oracle on the CPU, each of these
  raised the error to 0.48-370 b h: the sign of the image-point correction,
  the sign of the -1 diagonal, a Neumann row without its h scaling, the
  factor 2 of a ghost inside the object, the sign of the unknown-ghost
  weights, and (rasterize) the sign of the 0.75 h term, or h replaced by
  0.75 h in the centre-sampled rasterized value;
* a fluid cell whose two neighbours along each of two axes are solid keeps
  p^{n+1} = p^n, while a cell solid on both sides along one axis only does
  not (the cavity needs two such axes)."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import (FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE, FDTD_OBJ_SPHERE,
                                FDTD_SAMPLE_CENTER, FDTD_SAMPLE_SEVEN, SceneBuffers)
from oracle.oracle import Oracle

A0, B = 1.0, 50.0


def _radial(boundary, h, R, n, sampling=FDTD_SAMPLE_CENTER):
    s = scenes.base_settings((n, n, n), h, 6, boundary)
    s["sampling"] = sampling
    c = np.array([0.0013, -0.0007, 0.0004])   # off the cell lattice
    rho = s["density"]
    obj = dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(R, R, R), rotation=np.eye(3))
    # a_n(t) = -b / rho for all t of interest: a Gaussian of width 1e6 s
    imp = dict(object=0, amplitude=-B / rho, width=1e6, onset=0.0)
    sb = SceneBuffers(s, objects=[obj], impulse_sources=[imp], listeners=np.array([[0.07, 0.07, 0.07]]))
    o = Oracle(sb, reclassify_each_step=False)
    idx = np.indices((n, n, n)).reshape(3, -1).T
    x = np.array(s["min_bound"]) + (idx + 0.5) * h
    p = A0 + B * (np.linalg.norm(x - c, axis=1) - R)
    val, pos, nbr = o.ghost_update_at(p, 1e-3)
    if sampling == FDTD_SAMPLE_CENTER and boundary == FDTD_BOUNDARY_RASTERIZE:
        pos = 2.0 * pos - x[nbr]     # the solid cell's centre, mirrored from its neighbour's
    exact = A0 + B * (np.linalg.norm(pos - c, axis=1) - R)
    return o, p, np.abs(val - exact).max() / (B * h)


@pytest.mark.parametrize("sampling", [FDTD_SAMPLE_CENTER, FDTD_SAMPLE_SEVEN])
@pytest.mark.parametrize("boundary", [FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE])
def test_ghost_values_reproduce_radial_field(boundary, sampling):
    o, _, coarse = _radial(boundary, 0.004, 0.04, 48, sampling)
    if boundary == FDTD_BOUNDARY_FULLY_COUPLED:
        assert o.counts()["n_coupled"] > 0.95 * o.counts()["n_ghost"]
    _, _, fine = _radial(boundary, 0.002, 0.04, 80, sampling)
    assert coarse <= 0.06, coarse
    assert fine <= 0.65 * coarse, (coarse, fine)


def test_coupled_system_solution_is_lapack_solution():
    o, p, _ = _radial(FDTD_BOUNDARY_FULLY_COUPLED, 0.004, 0.04, 48)
    rng = np.random.default_rng(5)
    q = p + 0.1 * rng.standard_normal(p.size)   # any field: the system is linear in it
    val, _, _ = o.ghost_update_at(q, 2e-3)
    A, rhs = o.ghost_system()
    assert A.shape[0] == val.size
    # rows with an 8-cell image stencil hold a -1 diagonal (coupled or fallen
    # back on cond > 25), the others the identity row of the rasterized value
    assert np.all(np.abs(np.diag(A)) == 1.0)
    assert np.sum(np.diag(A) == -1.0) >= o.counts()["n_coupled"]
    ref = np.linalg.solve(A, rhs)
    assert np.abs(val - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.linalg.norm(A @ val - rhs) <= 1e-12 * np.linalg.norm(rhs) * max(1.0, np.linalg.norm(A, 1))


def cavity_scene(h=0.005, n=32):
    """Four spheres of radius 0.9 h 1.5 h from the centre of cell (16, 16, 16)
    along -x, +x, -y, +y make its x and y neighbours solid and leave it
    fluid (its samples a quarter cell out stay 1.25 h from every centre): a
    one-cell cavity.  Two spheres along +-x around cell (16, 24, 16) make a
    channel cell, solid on both sides along x only."""
    s = scenes.base_settings((n, n, n), h, 6, FDTD_BOUNDARY_RASTERIZE)
    mb = np.array(s["min_bound"])
    cav = mb + (np.array([16, 16, 16]) + 0.5) * h
    chan = mb + (np.array([16, 24, 16]) + 0.5) * h
    objs = []
    for ctr, axes in ((cav, (0, 1)), (chan, (0,))):
        for d in axes:
            for sg in (-1.0, 1.0):
                c = ctr.copy()
                c[d] += sg * 1.5 * h
                objs.append(dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(0.9 * h,) * 3, rotation=np.eye(3)))
    lst = np.array([mb + (np.array([16, 16, 20]) + 0.3) * h, mb + (np.array([10, 10, 10]) + 0.5) * h])
    sb = SceneBuffers(s, objects=objs, impulse_sources=[dict(object=0, amplitude=1e3, width=1e-4, onset=1e-4)],
                      listeners=lst)
    flat = lambda ijk: (ijk[0] * n + ijk[1]) * n + ijk[2]
    return sb, flat((16, 16, 16)), flat((16, 24, 16))


def test_one_cell_cavity_holds_its_value():
    sb, cav, chan = cavity_scene()
    o = Oracle(sb, reclassify_each_step=False)
    rng = np.random.default_rng(9)
    n = sb.n_cells
    last, curr = rng.standard_normal(n), rng.standard_normal(n)
    o.set_state(last, curr)
    for it in range(5):
        o.step(1, record=False)
        assert o.counts()["n_cavity"] == 1   # flagged by the step's Laplacian
        p = o.pressure().reshape(-1)
        assert p[cav] == curr[cav]          # frozen, bit for bit
        assert p[chan] != curr[chan]        # one doubly-solid axis: updated
