"""Host dry runs of the device planner (fdtd_plan_check, no GPU): for every
tile plan and scenes of every kind the engine runs (free field with a
pressure source, sphere and ellipsoids in both boundary modes, uneven batch
grids, stream requests), every tile's class matches the launch that runs it,
every read stays inside the grid, every cell lies in exactly one tile box,
object cells only in boxes that skip them, and every object-adjacent record
in exactly one tile.  A plan that broke any of these would read or write
outside its arrays on the device."""
import ctypes as C

import pytest

from acoustics_amd import capi, scenes

N_PLANS = 10


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
    # the default plan is plan 0
    n = C.c_int64(0)
    assert lib.fdtd_plan_check(C.byref(sb.scene), -1, C.byref(n)) == 0 and n.value == counts[0]


def test_staggered_scheme_rejects_objects():
    """The staggered scheme runs free-field scenes only; a scene with an object
    is refused with FDTD_ERR_UNSUPPORTED and a message."""
    lib = capi.load_library()
    sb, _ = scenes.config2(seed=2, n=48, pml=8, radius=0.03, shell=0.07, n_listeners=4)
    sb.scene.settings.scheme = capi.FDTD_SCHEME_STAGGERED
    assert lib.fdtd_plan_check(C.byref(sb.scene), 0, None) == 6
    assert b"staggered" in lib.fdtd_last_error()
