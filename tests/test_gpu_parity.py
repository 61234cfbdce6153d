"""Parity of the HIP path (FP32 fields, FP64 ghosts/listeners) against the FP64
CPU oracle on the same seeded scenes.

Tolerance: listener relative L2 error <= 1e-4 over the whole time series
(the engine's stated accuracy contract).  Graph replay and eager stepping
must agree bit for bit, and reruns must be bitwise deterministic.
"""
import numpy as np
import pytest

from acoustics_amd import capi, scenes
from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _run_pair(sb, steps, **kw):
    print(f"case: divisions={sb.divisions} listeners={sb.n_listeners} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb, **kw) as s:
        s.step(steps)
        got = s.series()
    return got, ref


def test_free_field_small(hip_lib):
    sb, _ = scenes.config1(seed=11, n=32)
    got, ref = _run_pair(sb, 240)
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= TOL, rel_l2(got, ref)


def test_free_field_config1(hip_lib):
    sb, steps = scenes.config1(seed=1, n=64, steps=500)
    got, ref = _run_pair(sb, steps)
    err = rel_l2(got, ref)
    print("config1 listener rel L2", err)
    assert err <= TOL


def test_sphere_rasterize(hip_lib):
    sb, _ = scenes.config2(seed=21, n=48, radius=0.03, shell=0.07, pml=8, n_listeners=16)
    got, ref = _run_pair(sb, 300)
    err = rel_l2(got, ref)
    print("sphere rasterize rel L2", err)
    assert err <= TOL


def test_ellipsoids_rasterize(hip_lib):
    sb, _ = scenes.config3(seed=31, n=64, h=0.004, pml=8, n_obj=3, n_listeners=16, axis_lo=0.01, axis_hi=0.025)
    got, ref = _run_pair(sb, 300)
    err = rel_l2(got, ref)
    print("ellipsoids rasterize rel L2", err)
    assert err <= TOL


def test_graph_replay_matches_eager(hip_lib):
    sb, _ = scenes.config2(seed=22, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)
    with Solver(sb) as a:
        a.step(120)
        ea = a.series()
        pa = a.pressure()
    with Solver(sb, steps_per_graph=24) as b:
        b.step(5)      # unaligned prefix runs eagerly
        b.step(115)    # graph blocks + eager tail
        eb = b.series()
        pb = b.pressure()
    assert np.array_equal(ea, eb)
    assert np.array_equal(pa, pb)


def test_deterministic_rerun(hip_lib):
    sb, _ = scenes.config1(seed=12, n=40)
    outs = []
    for _ in range(2):
        with Solver(sb) as s:
            s.step(100)
            outs.append((s.series(), s.pressure()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("tile", list(range(11)))
def test_tile_configs_agree(hip_lib, tile):
    sb, _ = scenes.config2(seed=23, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)
    with Solver(sb, tile_config=0) as a:
        a.step(60)
        ea = a.pressure()
    with Solver(sb, tile_config=tile) as b:
        b.step(60)
        eb = b.pressure()
    assert np.array_equal(ea, eb)


def test_sphere_fully_coupled(hip_lib):
    sb, _ = scenes.config2(seed=24, n=48, radius=0.03, shell=0.07, pml=8, n_listeners=16,
                           boundary=FDTD_BOUNDARY_FULLY_COUPLED)
    got, ref = _run_pair(sb, 300)
    err = rel_l2(got, ref)
    print("sphere fully_coupled rel L2", err)
    assert err <= TOL


def test_ellipsoids_fully_coupled(hip_lib):
    sb, _ = scenes.config3(seed=32, n=64, h=0.004, pml=8, n_obj=3, n_listeners=16, axis_lo=0.01, axis_hi=0.025,
                           boundary=FDTD_BOUNDARY_FULLY_COUPLED)
    got, ref = _run_pair(sb, 250)
    err = rel_l2(got, ref)
    print("ellipsoids fully_coupled rel L2", err)
    assert err <= TOL


def test_listener_next_to_object(hip_lib):
    """Listeners whose enclosing cells include ghost cells (MLS over ghost
    children, PML_WaveSolver.cpp:246-262) and solid cells."""
    sb0, _ = scenes.config2(seed=25, n=40, radius=0.025, shell=0.07, pml=6, n_listeners=4)
    c = np.array(sb0.objects_list[0]["center"])
    lst = np.array([c + [0.026, 0.0, 0.0], c + [0.0, -0.0255, 0.001], c + [0.018, 0.018, 0.0], c + [0.0, 0.0, 0.0262]])
    from acoustics_amd.capi import SceneBuffers
    sb = SceneBuffers(sb0.settings_dict, sb0.objects_list, sb0.impulse_list, [], lst)
    got, ref = _run_pair(sb, 200)
    err = rel_l2(got, ref)
    print("near-object listeners rel L2", err)
    assert err <= TOL


@pytest.mark.parametrize("boundary", [0, FDTD_BOUNDARY_FULLY_COUPLED])
def test_odd_sized_grid(hip_lib, boundary):
    """Rows whose length is not a multiple of four: the device pads each row
    to a multiple of four floats, float4 row tiles end in a partial group, and
    object cells, ghost neighbours and listener cells are addressed through
    the padded rows."""
    rng = np.random.default_rng(41)
    div = (45, 39, 53)
    s = scenes.base_settings(div, 0.005, 8, boundary)
    lo, hi = scenes.interior_bounds(s, 2)
    c = (lo + hi) / 2 + rng.uniform(-0.005, 0.005, 3)
    obj = dict(type=1, center=tuple(c), semi_axes=(0.022, 0.015, 0.018), rotation=scenes.random_rotation(rng))
    src = dict(object=0, amplitude=1e3, width=1e-4, onset=4e-4)
    ps = scenes.gaussian_pressure_source(lo + 0.3 * (hi - lo), 2e-4)
    lst = np.vstack([scenes.fibonacci_shell(c, 0.032, 10), rng.uniform(lo, hi, size=(4, 3))])
    from acoustics_amd.capi import SceneBuffers
    sb = SceneBuffers(s, objects=[obj], impulse_sources=[src], pressure_sources=[ps], listeners=lst)
    got, ref = _run_pair(sb, 260)
    err = rel_l2(got, ref)
    print("odd-sized grid rel L2", err)
    assert err <= TOL
    with Solver(sb) as a:
        a.step(40)
        pa = a.pressure()
    o = Oracle(sb, reclassify_each_step=False)
    o.step(40, record=False)
    assert pa.shape == tuple(div)
    assert rel_l2(pa, o.pressure()) <= TOL
    o.close()


@pytest.mark.parametrize("n2", [168, 100, 257])
def test_row_widths(hip_lib, n2):
    """Float4 row tiles at widths that leave lanes idle: one row per
    wavefront with 38 of 64 lanes (n2 = 168), three rows of 21 lanes
    (n2 = 100), and 61 lanes ending in a partial group with a padded row
    (n2 = 257).  The whole field is compared, so a lane writing outside its
    row's interior cells shows up in the PML."""
    rng = np.random.default_rng(n2)
    div = (22, 20, n2)
    s = scenes.base_settings(div, 0.005, 8)
    lo, hi = scenes.interior_bounds(s, 1)
    src = scenes.gaussian_pressure_source((lo + hi) / 2, 2e-4)
    from acoustics_amd.capi import SceneBuffers
    sb = SceneBuffers(s, pressure_sources=[src], listeners=rng.uniform(lo, hi, size=(6, 3)))
    print(f"case: row widths divisions={div}", flush=True)
    o = Oracle(sb, reclassify_each_step=False)
    ref = o.step(120)
    p_ref = o.pressure()
    o.close()
    with Solver(sb) as a:
        a.step(120)
        got, p = a.series(), a.pressure()
    assert rel_l2(got, ref) <= TOL, rel_l2(got, ref)
    assert rel_l2(p, p_ref) <= TOL, rel_l2(p, p_ref)
    pml = np.ones(div, dtype=bool)
    pml[8:-8, 8:-8, 8:-8] = False
    assert rel_l2(p[pml], p_ref[pml]) <= TOL, rel_l2(p[pml], p_ref[pml])


@pytest.mark.parametrize("case", ["sphere_raster", "ellipsoids_coupled", "free_field_source"])
def test_air_viscosity(hip_lib, case):
    """air_viscosity_alpha > 0 (ImpulseResponseParser.cpp:540-549): the bulk
    cells keep the previous step's Laplacian (MAC_Grid.cpp:926-930), in the
    2D-chunk interior tiles, the tiles with a pressure source and the
    object-adjacent cells (FP64 ghost-aware Laplacian)."""
    alpha = 2e-4
    if case == "sphere_raster":
        sb0, _ = scenes.config2(seed=27, n=44, radius=0.025, shell=0.06, pml=8, n_listeners=10)
    elif case == "ellipsoids_coupled":
        sb0, _ = scenes.config3(seed=33, n=56, h=0.004, pml=8, n_obj=2, n_listeners=10, axis_lo=0.01,
                                axis_hi=0.02, boundary=FDTD_BOUNDARY_FULLY_COUPLED)
    else:
        sb0, _ = scenes.config1(seed=13, n=44)
    s = dict(sb0.settings_dict)
    s["air_viscosity_alpha"] = alpha
    from acoustics_amd.capi import SceneBuffers
    sb = SceneBuffers(s, sb0.objects_list, sb0.impulse_list, sb0.pressure_list, sb0.listeners_np)
    got, ref = _run_pair(sb, 260)
    err = rel_l2(got, ref)
    print(case, "viscous rel L2", err)
    assert err <= TOL
    # the viscous term is visible: the inviscid series differs from the
    # viscous one far more than the device differs from the oracle
    ref0 = Oracle(sb0, reclassify_each_step=False).step(260)
    assert rel_l2(ref0, ref) > max(10 * TOL, 1000 * err), (rel_l2(ref0, ref), err)
    # the reference's snapshot files lack the previous Laplacian: a viscous
    # scene resumes from the engine state file only
    import os
    import tempfile
    with Solver(sb, pml_state=capi.FDTD_PML_STATE_SPLIT) as a, tempfile.TemporaryDirectory() as d:
        a.step(2)
        a.save_snapshot(d + "/s_%s", "0", 0)
        os.remove(d + "/s_0_snapshot_0_engine_state.bin")
        with Solver(sb, pml_state=capi.FDTD_PML_STATE_SPLIT) as b:
            with pytest.raises(RuntimeError, match="Laplacian"):
                b.load_snapshot(d + "/s_%s", "0", 0)


def _engquist_scene(case):
    if case == "free_field":
        return scenes.engquist_free_field(seed=51, div=(40, 44, 48), n_listeners=8)[0]
    if case == "free_field_use_mesh":
        return scenes.engquist_free_field(seed=52, div=(37, 41, 45), n_listeners=8, use_mesh=1)[0]
    if case == "sphere_raster":
        return scenes.engquist_sphere(seed=53, n=44)[0]
    return scenes.engquist_sphere(seed=54, n=46, boundary=FDTD_BOUNDARY_FULLY_COUPLED)[0]


@pytest.mark.parametrize("case", ["free_field", "free_field_use_mesh", "sphere_raster", "sphere_coupled"])
def test_engquist_boundary(hip_lib, case):
    """Scenes without PML cells (use_mesh = 0, or a PML width of zero): the
    layer-1 cells take the Engquist boundary (MAC_Grid.cpp:521-919) in the
    boundary tiles, whose listener terms include the face and the (1, 1, 1)
    corner cells the free-field listeners read; the pulse reaches the
    boundary well inside the steps run."""
    sb = _engquist_scene(case)
    got, ref = _run_pair(sb, 400)
    err = rel_l2(got, ref)
    print(case, "Engquist rel L2", err)
    assert np.all(np.isfinite(got))
    assert err <= TOL
    with Solver(sb) as s:
        info = s.info()
        assert info["n_pml_cells"] == 0 and info["n_pml_merged_cells"] == 0
        s.step(400)
        p = s.pressure()
    o = Oracle(sb, reclassify_each_step=False)
    o.step(400, record=False)
    q = o.pressure().reshape(sb.divisions)
    o.close()
    # layer 0 holds boundary state on the device; every updated cell matches
    inner, qi = p[1:-1, 1:-1, 1:-1], q[1:-1, 1:-1, 1:-1]
    assert np.linalg.norm(inner - qi) <= 1e-4 * np.linalg.norm(qi)


@pytest.mark.parametrize("tile", list(range(11)))
def test_engquist_tile_plans_and_graphs_agree(hip_lib, tile):
    sb = _engquist_scene("sphere_raster")
    print(f"case: Engquist tile plan {tile}", flush=True)
    with Solver(sb, tile_config=0) as a:
        a.step(90)
        ea, sa = a.pressure(), a.series()
    with Solver(sb, tile_config=tile, steps_per_graph=16 if tile % 2 else 0) as b:
        b.step(7)
        b.step(83)
        eb, sb_ = b.pressure(), b.series()
    assert np.array_equal(ea, eb)
    assert np.array_equal(sa, sb_)
