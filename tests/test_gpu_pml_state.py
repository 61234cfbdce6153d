"""The merged PML state (FDTD_PML_STATE_MERGED, the default) against the
reference's split state (FDTD_PML_STATE_SPLIT) and against the FP64 oracle.

A PML cell damped along one axis A only keeps the increment d of the sum q
of its two undamped split pressures instead of the splits and their face
velocities; with f1 = f3 = 1 on those axes the split updates reduce exactly
to d^{n+1} = d^n + transverse second differences and
p^{n+1} = p^n + (pDir_A^{n+1} - pDir_A^n) + d^{n+1}, so the two states differ
by FP32 rounding only.  Listeners placed inside the PML slabs,
edges and corners read the PML state directly.

Tolerances: merged vs split, relative L2 <= 2e-5 over the pressure field and
the listener series after a pulse's passage (FP32 rounding of two equivalent
updates); against the oracle, the engine's 1e-4 listener contract over the
series and for each listener on the scale of the series' RMS listener norm,
and over the whole field (the PML included) no more than 1.5 times the FP32
split state's own error (or 2e-5)."""
import numpy as np
import pytest

from acoustics_amd import capi, scenes
from acoustics_amd.capi import SceneBuffers
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL_STATE = 2e-5
TOL_ORACLE = 1e-4


def _pml_listener_scene(div, pml, seed):
    """Free field with a pulse near one corner of the interior and listeners
    in an x slab, a y slab, a z slab, an edge, a corner and the interior."""
    rng = np.random.default_rng(seed)
    h = 0.005
    s = scenes.base_settings(div, h, pml)
    bmin = np.array(s["min_bound"], dtype=float)
    n = np.array(div)
    ext = n * h
    w = pml * h
    src_pos = bmin + np.array([w + 4 * h, ext[1] / 2, ext[2] - w - 5 * h])
    src = scenes.gaussian_pressure_source(src_pos, 2e-4)
    mid = bmin + ext / 2
    listeners = np.array([
        [bmin[0] + 0.4 * w, mid[1] + 0.01, mid[2] - 0.007],            # x slab
        [mid[0] - 0.006, bmin[1] + ext[1] - 0.5 * w, mid[2] + 0.004],   # y slab
        [mid[0] + 0.003, mid[1] - 0.008, bmin[2] + ext[2] - 0.3 * w],   # z slab
        [bmin[0] + 0.6 * w, bmin[1] + 0.6 * w, mid[2]],                 # x-y edge
        [bmin[0] + 0.5 * w, mid[1], bmin[2] + ext[2] - 0.7 * w],        # x-z edge
        [bmin[0] + 0.5 * w, bmin[1] + 0.5 * w, bmin[2] + 0.5 * w],      # corner
        mid + rng.uniform(-0.01, 0.01, 3),                              # interior
    ])
    return SceneBuffers(s, pressure_sources=[src], listeners=listeners)


def _run(sb, steps, state, **kw):
    with Solver(sb, pml_state=state, **kw) as s:
        s.step(steps)
        return s.series(), s.pressure(), s.info()


@pytest.mark.parametrize("div,pml,steps", [((40, 40, 40), 8, 260), ((36, 44, 52), 7, 260), ((48, 32, 40), 10, 260),
                                            ((32, 36, 40), 6, 1500)])
def test_merged_matches_split_and_oracle(hip_lib, div, pml, steps):
    sb = _pml_listener_scene(div, pml, seed=sum(div))
    print(f"case: pml state divisions={div} pml={pml} steps={steps}", flush=True)
    ser_m, p_m, inf_m = _run(sb, steps, capi.FDTD_PML_STATE_MERGED)
    ser_s, p_s, inf_s = _run(sb, steps, capi.FDTD_PML_STATE_SPLIT)
    assert inf_s["n_pml_merged_cells"] == 0
    nx, ny, nz = div
    # cells damped on the slab axis only: every PML slab cell whose two
    # other coordinates are interior
    want = 2 * pml * ((ny - 2 * pml) * (nz - 2 * pml) + (nx - 2 * pml) * (nz - 2 * pml) + (nx - 2 * pml) * (ny - 2 * pml))
    assert inf_m["n_pml_merged_cells"] == want
    assert inf_m["state_bytes_per_step"] == inf_m["bytes_per_step"] - 24 * want
    assert np.all(np.isfinite(ser_m)) and np.abs(ser_m).max() > 0
    if steps <= 300:
        assert rel_l2(p_m, p_s) <= TOL_STATE, rel_l2(p_m, p_s)
        assert rel_l2(ser_m, ser_s) <= TOL_STATE, rel_l2(ser_m, ser_s)
    o = Oracle(sb, reclassify_each_step=False)
    ref = o.step(steps)
    p_ref = o.pressure()
    o.close()
    assert rel_l2(ser_m, ref) <= TOL_ORACLE, rel_l2(ser_m, ref)
    # the whole field (PML included) against the FP64 split state: the merged
    # state is as accurate as the FP32 split state
    fm, fs = rel_l2(p_m, p_ref), rel_l2(p_s, p_ref)
    print(f"field: merged {fm:.3e} split {fs:.3e}", flush=True)
    assert fm <= max(TOL_STATE, 1.5 * fs), (fm, fs)
    # per listener, on the scale of the whole series (a listener deep in the
    # layer sees a pressure far smaller than the splits it is the sum of, so
    # its error relative to its own signal reflects FP32 rounding of those
    # splits in either state: 1e-4 to 2e-3 for the split state there)
    rms = np.linalg.norm(ref) / np.sqrt(ref.shape[1])
    for c in range(ser_m.shape[1]):
        em, es = rel_l2(ser_m[:, c], ref[:, c]), rel_l2(ser_s[:, c], ref[:, c])
        am = np.linalg.norm(ser_m[:, c] - ref[:, c]) / rms
        print(f"listener {c}: merged {em:.3e} split {es:.3e} (merged vs split {rel_l2(ser_m[:, c], ser_s[:, c]):.3e}); "
              f"merged error / rms listener norm {am:.3e}", flush=True)
        assert am <= TOL_ORACLE, (c, am)


def test_merged_state_with_objects_graph_and_plans(hip_lib):
    """Merged state under graph replay and every tile plan, on an object scene:
    bit-identical to eager stepping with the default plan."""
    sb, _ = scenes.config2(seed=26, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)
    print(f"case: merged graph/plans divisions={sb.divisions}", flush=True)
    with Solver(sb) as a:
        a.step(64)
        ea, pa = a.series(), a.pressure()
    with Solver(sb, steps_per_graph=16) as b:
        b.step(64)
        assert np.array_equal(b.series(), ea)
        assert np.array_equal(b.pressure(), pa)
    with Solver(sb, pml_state=capi.FDTD_PML_STATE_SPLIT) as c:
        c.step(64)
        assert rel_l2(c.pressure(), pa) <= TOL_STATE
