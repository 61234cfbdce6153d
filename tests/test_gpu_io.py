"""GPU checks of the reference tool's files and of checkpoint/resume.

fdtd_run_xml runs an XML scene as the reference's tool does
(FDTD_AcousticSimulator::SaveSolverConfig / Run, SimWorld's audio output):
listening_position.dat must read back through scripts/binary_io.py's format,
all_audio.dat must hold one row of L doubles per step equal (bit for bit) to
the series of the same scene stepped through the single-scene API, and the
control file must be copied unchanged.  Snapshots in the layout of
FDTD_AcousticSimulator::_SaveSimulationSnapshot must resume a run bit for bit,
in both schemes."""
import os

import numpy as np
import pytest

from acoustics_amd import capi, scenes
from acoustics_amd.capi import SceneBuffers
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_matrix_xd_binary(path):
    """scripts/binary_io.py:readMatrixXdBinary: int32 rows, int32 cols, doubles in row order."""
    with open(path, "rb") as f:
        n1 = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        n2 = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        data = np.fromfile(f, dtype=np.float64, count=n1 * n2)
    return data.reshape(n1, n2)


def test_run_xml_writes_reference_files(hip_lib, tmp_path):
    import ctypes as C
    src = os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")
    text = open(src).read().replace('output_pattern="out_%s"', f'output_pattern="{tmp_path}/run_%s"')
    xml = tmp_path / "scene.xml"
    xml.write_text(text)
    print(f"case: run_xml {xml}", flush=True)
    n = C.c_int64(0)
    assert hip_lib.fdtd_run_xml(str(xml).encode(), b"7", None, C.byref(n)) == 0, hip_lib.fdtd_last_error()
    assert n.value == 300
    lp = read_matrix_xd_binary(str(tmp_path / "run_7_listening_position.dat"))
    assert np.allclose(lp, [[0.05, 0.01, -0.02], [-0.04, 0.03, 0.01]])
    audio = np.fromfile(str(tmp_path / "run_all_audio.dat"), dtype=np.float64).reshape(-1, 2)
    assert audio.shape == (300, 2)
    with Solver.from_xml(str(xml)) as s:
        s.step(300)
        ref = s.series()
    assert np.array_equal(audio, ref)
    assert open(tmp_path / "run_7_input_control_file.xml").read() == text
    assert "steps 300" in open(tmp_path / "run_7_solver_settings.txt").read()


@pytest.mark.parametrize("kind", ["collocated_sphere", "staggered_free_field", "engquist_free_field"])
def test_snapshot_resumes_bit_for_bit(hip_lib, tmp_path, kind):
    if kind == "collocated_sphere":
        sb, _ = scenes.config2(seed=21, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)
    elif kind == "engquist_free_field":
        # no PML: the boundary's kept inner values are rebuilt from p^{n-1} on load
        sb, _ = scenes.engquist_free_field(seed=63, div=(34, 36, 38))
    else:
        sb, _ = scenes.staggered_free_field(seed=9, div=(36, 32, 40), pml=6)
    print(f"case: snapshot {kind} divisions={sb.divisions}", flush=True)
    pat = str(tmp_path / "snap_%s")
    split = capi.FDTD_PML_STATE_SPLIT
    with Solver(sb, pml_state=split) as a:
        a.step(70)
        a.save_snapshot(pat, "0", 2)
        a.step(50)
        ser_a = a.series()
        p_a = a.pressure()
    # the files follow the reference's MATRIX::write layout
    N = int(np.prod(sb.divisions))
    assert read_matrix_xd_binary(str(tmp_path / "snap_0_snapshot_2_p_pml_0.dat")).shape == (N, 1)
    nx, ny, nz = sb.divisions
    assert read_matrix_xd_binary(str(tmp_path / "snap_0_snapshot_2_v_pml_1.dat")).shape == (nx * (ny + 1) * nz, 1)
    # the reference's files alone (without the engine state file) resume a split handle exactly
    os.remove(str(tmp_path / "snap_0_snapshot_2_engine_state.bin"))
    with Solver(sb, pml_state=split) as b:
        b.load_snapshot(pat, "0", 2)
        b.step(50)
        ser_b = b.series()
        p_b = b.pressure()
    assert np.array_equal(ser_b[70:], ser_a[70:])
    assert np.array_equal(p_b, p_a)
    assert np.all(ser_b[:70] == 0)


def _write_matrix(path, v):
    """IO::writeMatrixX / MATRIX::write layout: int32 rows, int32 cols, doubles."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 1)
    with open(path, "wb") as f:
        f.write(np.array([v.shape[0], 1], dtype=np.int32).tobytes())
        f.write(v.tobytes())


def _snap_scene():
    return scenes.config2(seed=24, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)[0]


def test_snapshot_files_hold_the_oracles_state(hip_lib, tmp_path):
    """The saved p_collocated (p^n, p^{n-1}), p_pml_d and v_pml_d equal the
    FP64 oracle's state after the same steps, to FP32 accuracy."""
    sb = _snap_scene()
    steps = 90
    print(f"case: snapshot vs oracle divisions={sb.divisions} steps={steps}", flush=True)
    pat = str(tmp_path / "s_%s")
    with Solver(sb, pml_state=capi.FDTD_PML_STATE_SPLIT) as a:
        a.step(steps)
        a.save_snapshot(pat, "0", 0)
    o = Oracle(sb, reclassify_each_step=False)
    o.step(steps, record=False)
    last, curr, pdir, vel = o.state()
    rd = lambda what: read_matrix_xd_binary(str(tmp_path / f"s_0_snapshot_0_{what}.dat"))[:, 0]
    scale = np.abs(curr).max()
    assert scale > 0
    assert np.abs(rd("p_collocated_1") - curr).max() <= 1e-5 * scale
    assert np.abs(rd("p_collocated_0") - last).max() <= 1e-5 * scale
    for d in range(3):
        ps = np.abs(pdir[d]).max()
        assert np.abs(rd(f"p_pml_{d}") - pdir[d]).max() <= 1e-5 * max(ps, 1e-30), d
        vs = np.abs(vel[d]).max()
        assert vs > 0
        assert np.abs(rd(f"v_pml_{d}") - vel[d]).max() <= 1e-5 * vs, d


@pytest.mark.parametrize("pml_state", [capi.FDTD_PML_STATE_SPLIT, capi.FDTD_PML_STATE_MERGED])
def test_resume_from_oracle_state(hip_lib, tmp_path, pml_state):
    """A state the FP64 oracle wrote in the reference's snapshot layout, loaded
    and stepped on, follows the oracle's own continuation (relative L2 <=
    1e-4 over the listener rows after the load).  A merged handle converts
    the split state: pDir_A, and the increment of the transverse sum from the
    face velocities."""
    sb = _snap_scene()
    k, m = 80, 120
    print(f"case: resume from oracle state pml_state={pml_state} divisions={sb.divisions}", flush=True)
    o = Oracle(sb, reclassify_each_step=False)
    o.step(k, record=False)
    last, curr, pdir, vel = o.state()
    pre = str(tmp_path / "o_0_snapshot_1_")
    _write_matrix(pre + "p_collocated_1.dat", curr)
    _write_matrix(pre + "p_collocated_0.dat", last)
    _write_matrix(pre + "p_collocated_2.dat", np.zeros_like(curr))
    _write_matrix(pre + "p_pml_full.dat", np.zeros_like(curr))
    for d in range(3):
        _write_matrix(pre + f"p_pml_{d}.dat", pdir[d])
        _write_matrix(pre + f"v_pml_{d}.dat", vel[d])
    open(pre + "p_collocated_ind.dat", "w").write("1")
    open(pre + "step.txt", "w").write(f"{k}\n")
    ref = o.step(m)
    with Solver(sb, pml_state=pml_state) as b:
        b.load_snapshot(str(tmp_path / "o_%s"), "0", 1)
        b.step(m)
        got = b.series()[k:]
    assert rel_l2(got, ref) <= 1e-4, rel_l2(got, ref)


@pytest.mark.parametrize("kind", ["merged_sphere", "viscous_sphere"])
def test_engine_state_resumes_bit_for_bit(hip_lib, tmp_path, kind):
    """The default (merged) PML state and air viscosity, whose state the
    reference's files cannot hold in full, resume bit for bit from the engine
    state file every snapshot also writes."""
    sb = _snap_scene()
    if kind == "viscous_sphere":
        sb.settings_dict["air_viscosity_alpha"] = 1e-5
        sb = SceneBuffers(sb.settings_dict, sb.objects_list, sb.impulse_list, [], sb.listeners_np)
    print(f"case: engine state {kind} divisions={sb.divisions}", flush=True)
    pat = str(tmp_path / "e_%s")
    with Solver(sb) as a:
        a.step(70)
        a.save_snapshot(pat, "0", 3)
        a.step(60)
        ser_a, p_a = a.series(), a.pressure()
    with Solver(sb) as b:
        b.load_snapshot(pat, "0", 3)
        b.step(60)
        ser_b, p_b = b.series(), b.pressure()
    assert np.array_equal(ser_b[70:], ser_a[70:])
    assert np.array_equal(p_b, p_a)
    if kind == "merged_sphere":
        # a split handle refuses the merged handle's reference files, whose
        # transverse face velocities of merged cells the merged state does not keep
        with Solver(sb, pml_state=capi.FDTD_PML_STATE_SPLIT) as c:
            with pytest.raises(RuntimeError, match="merged"):
                c.load_snapshot(pat, "0", 3)
