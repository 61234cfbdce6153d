"""Host dry runs of the device planner (fdtd_plan_check, no GPU): for every
tile plan and scenes of every kind the engine runs (free field with a
pressure source, sphere and ellipsoids in both boundary modes, uneven batch
grids, stream requests), every tile's class matches the launch that runs it,
every read stays inside the grid, every cell lies in exactly one tile box (none of the outermost layer
with the Engquist boundary),
object cells only in boxes that skip them, and every object-adjacent record
in exactly one tile.  A plan that broke any of these would read or write
outside its arrays on the device."""
import ctypes as C

import numpy as np
import pytest

from acoustics_amd import capi, scenes

N_PLANS = 11


def _scenes():
    out = [("config1_40", scenes.config1(seed=1, n=40)[0]),
           ("sphere_raster_48", scenes.config2(seed=2, n=48, pml=8, radius=0.03, shell=0.07, n_listeners=16)[0]),
           ("sphere_coupled_48", scenes.config2(seed=2, n=48, pml=8, radius=0.03, shell=0.07, n_listeners=16,
                                                boundary=capi.FDTD_BOUNDARY_FULLY_COUPLED)[0]),
           ("ellipsoids_80", scenes.config3(seed=3, n=80, h=0.64 / 80, pml=8)[0])]
    bs, _ = scenes.config4(seed=4, n_scenes=4, edge_lo=32, edge_hi=70)
    out += [(f"batch{i}", b) for i, b in enumerate(bs)]
    stream = scenes.config5_stream(seed=5, n_requests=2, edge_lo=48, edge_hi=64)
    out += [(f"stream{i}", sb) for i, (_, sb, _) in enumerate(stream)]
    out.append(("staggered_36x40x44", scenes.staggered_free_field(seed=7, div=(36, 40, 44))[0]))
    out.append(("engquist_free_field", scenes.engquist_free_field(seed=8, div=(37, 41, 300))[0]))
    out.append(("moving_sphere", scenes.moving_object(seed=10)[0]))
    out.append(("speaker_sphere", scenes.speaker_scene(seed=15, with_sphere=True)[0]))
    out.append(("moving_ellipsoid_engquist", scenes.moving_object(seed=11, kind="ellipsoid", speed=12.0, spin=300.0,
                                                                  pml_free=True, n=46)[0]))
    out.append(("engquist_sphere_coupled", scenes.engquist_sphere(seed=9, n=44,
                                                                  boundary=capi.FDTD_BOUNDARY_FULLY_COUPLED)[0]))
    return out


SCENES = _scenes()


@pytest.mark.parametrize("name,sb", SCENES, ids=[n for n, _ in SCENES])
def test_every_plan_is_valid(name, sb):
    lib = capi.load_library()
    counts = []
    for plan in range(N_PLANS):
        n = C.c_int64(0)
        st = lib.fdtd_plan_check(C.byref(sb.scene), plan, C.byref(n))
        assert st == 0, (name, plan, lib.fdtd_last_error())
        assert n.value > 0
        counts.append(n.value)
    # the default plan: 1 for the staggered scheme, 10 for a grid of under
    # 400,000 cells, else 0
    cells = int(np.prod(sb.divisions))
    default = 1 if sb.settings_dict.get("scheme", 0) == 1 else (10 if cells < 400000 else 0)
    n = C.c_int64(0)
    assert lib.fdtd_plan_check(C.byref(sb.scene), -1, C.byref(n)) == 0 and n.value == counts[default]


def test_staggered_scheme_rejects_objects():
    """The staggered scheme runs free-field scenes only; a scene with an object
    is refused with FDTD_ERR_UNSUPPORTED and a message."""
    lib = capi.load_library()
    sb, _ = scenes.config2(seed=2, n=48, pml=8, radius=0.03, shell=0.07, n_listeners=4)
    sb.scene.settings.scheme = capi.FDTD_SCHEME_STAGGERED
    assert lib.fdtd_plan_check(C.byref(sb.scene), 0, None) == 6
    assert b"staggered" in lib.fdtd_last_error()


def test_engquist_boundary_refusals():
    """Without PML cells the layer-1 cells take the Engquist boundary; the
    setups it cannot serve are refused with FDTD_ERR_UNSUPPORTED and a
    message: objects with use_mesh = 0 (the reference classifies none), a
    solid cell within three cells of the boundary, a listener reading a
    cell of layer 0 (never updated), the staggered scheme, and a PML one
    cell wide (layer 1 next to a PML layer)."""
    import numpy as np
    from acoustics_amd.capi import SceneBuffers
    lib = capi.load_library()

    def status(settings, sb, objects=None, listeners=None):
        b = SceneBuffers(settings, sb.objects_list if objects is None else objects, sb.impulse_list,
                         sb.pressure_list, sb.listeners_np if listeners is None else listeners)
        return lib.fdtd_plan_check(C.byref(b.scene), 0, None), lib.fdtd_last_error()

    ff, _ = scenes.engquist_free_field(seed=8)
    sp, _ = scenes.engquist_sphere(seed=9)
    h = ff.settings_dict["cell_size"]
    assert status(ff.settings_dict, ff)[0] == 0 and status(sp.settings_dict, sp)[0] == 0
    s = dict(sp.settings_dict, use_mesh=0)
    assert status(s, sp) == (6, b"use_mesh=0 classifies no object cells; objects need use_mesh=1")
    o = dict(sp.objects_list[0])
    o["center"] = tuple(np.array(sp.settings_dict["min_bound"]) + np.array([0.025 + 2.2 * h, 0.11, 0.11]))
    st, msg = status(sp.settings_dict, sp, objects=[o], listeners=np.zeros((0, 3)))
    assert st == 6 and b"three cells" in msg
    mb = np.array(ff.settings_dict["min_bound"])
    st, msg = status(ff.settings_dict, ff, listeners=np.array([mb + 1.2 * h]))
    assert st == 6 and b"outermost layer" in msg
    s = dict(ff.settings_dict, scheme=capi.FDTD_SCHEME_STAGGERED)
    st, msg = status(s, ff)
    assert st == 6 and b"collocated" in msg
    s = dict(sp.settings_dict, pml_width=1.0)
    st, msg = status(s, sp)
    assert st == 6 and b"two outermost cell layers" in msg


def test_moving_object_refusals():
    """Moving objects need the collocated scheme with rasterized ghost cells
    and no air viscosity, a stop_time to plan their path up to, a path (with
    a two-cell margin) inside the non-PML interior, and listeners clear of
    every cell the path reaches."""
    import numpy as np
    from acoustics_amd.capi import SceneBuffers
    lib = capi.load_library()
    sb, _ = scenes.moving_object(seed=10)

    def status(settings=None, objects=None, listeners=None):
        b = SceneBuffers(sb.settings_dict if settings is None else settings, sb.objects_list if objects is None else objects,
                         sb.impulse_list, sb.pressure_list, sb.listeners_np if listeners is None else listeners)
        return lib.fdtd_plan_check(C.byref(b.scene), 0, None), lib.fdtd_last_error()

    assert status()[0] == 0
    s = dict(sb.settings_dict, boundary_handling=capi.FDTD_BOUNDARY_FULLY_COUPLED)
    assert status(settings=s) == (6, b"moving objects run with the collocated scheme, rasterized ghost cells and no air viscosity")
    assert status(settings=dict(sb.settings_dict, air_viscosity_alpha=1e-4))[0] == 6
    st, msg = status(settings=dict(sb.settings_dict, stop_time=0.0))
    assert st == 1 and b"stop_time" in msg
    fast = [dict(sb.objects_list[0], velocity=(400.0, 0.0, 0.0))]
    st, msg = status(objects=fast, listeners=np.zeros((0, 3)))
    assert st == 6 and b"leaves the non-PML interior" in msg
    st, msg = status(listeners=np.array([sb.objects_list[0]["center"]]))
    assert st == 6 and b"path reaches" in msg
