"""Time-parallel chunks on the device (fdtd_run_time_parallel): every chunk
of a run (SimWorld.cpp:141-221) in one batch on one GPU, summed by linear
superposition.

* With an overlap reaching past the end of the run, the summed series equals
  the run without chunks up to FP32 rounding (the chunks only regroup the
  source terms), for pressure sources and object impulses alike.
* One chunk on its own matches the oracle's chunk (source gate included) to
  the engine's listener tolerance.
* With a long overlap, chunks stop early once their listeners fall below the
  reference's termination threshold (SimWorld.cpp:556-590), and the sum stays
  within the small tail they drop."""
import os

import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.capi import SceneBuffers
from acoustics_amd.solver import Solver, rel_l2, time_chunk, time_parallel_run
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _with(sb, **kw):
    s = dict(sb.settings_dict)
    s.update(kw)
    return SceneBuffers(s, sb.objects_list, sb.impulse_list, sb.pressure_list, sb.listeners_np)


def _whole(sb, n):
    with Solver(_with(sb, time_parallel=0)) as s:
        s.step(n)
        return s.series()


@pytest.mark.parametrize("case", ["pressure_source", "sphere_impulse"])
def test_chunks_sum_to_the_whole_run(hip_lib, case):
    if case == "pressure_source":
        sb, _ = scenes.config1(seed=81, n=36)
    else:
        sb, _ = scenes.config2(seed=82, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)
    n = 600
    tp = _with(sb, time_parallel=1, num_time_chunks=4, overlap_time=1.0, early_termination_ratio=1e-3)
    print(f"case: time-parallel {case} divisions={sb.divisions} steps={n}", flush=True)
    got, run = time_parallel_run(tp, n, steps_per_graph=16)
    assert run == [n - 150 * c for c in range(4)]
    ref = _whole(sb, n)
    err = rel_l2(got, ref)
    print(case, "chunk sum vs whole run rel L2", err)
    assert err <= 2e-5


def test_one_chunk_matches_the_oracle(hip_lib):
    sb, _ = scenes.config2(seed=83, n=40, radius=0.025, shell=0.06, pml=8, n_listeners=8)
    s = dict(sb.settings_dict, time_parallel=1, num_time_chunks=3, overlap_time=1e-3, early_termination_ratio=1e-3)
    d, first, steps = time_chunk(s, 600, 1)
    ch = SceneBuffers(d, sb.objects_list, sb.impulse_list, sb.pressure_list, sb.listeners_np)
    print(f"case: one chunk start={d['start_time']} steps={steps}", flush=True)
    ref = Oracle(ch, reclassify_each_step=False).step(steps)
    with Solver(ch) as a:
        a.step(steps)
        got = a.series()
    err = rel_l2(got, ref)
    print("chunk vs oracle rel L2", err)
    assert err <= 1e-4


def test_early_termination(hip_lib):
    """The first chunk's pulse is cut by its stop time, so its listeners are
    loud after the gate and the monitor arms; 0.05 s after they fall below
    1e-3 of the chunk's peak the chunk stops, well before the run ends."""
    sb0, _ = scenes.config1(seed=84, n=24)
    n, k = 60000, 2
    per = n // k
    dt = sb0.settings_dict["time_step"]
    src = scenes.gaussian_pressure_source(sb0.pressure_list[0]["position"], 2e-4, offset=per * dt - 1e-4)
    sb = SceneBuffers(sb0.settings_dict, pressure_sources=[src], listeners=sb0.listeners_np)
    tp = _with(sb, time_parallel=1, num_time_chunks=k, overlap_time=0.5, early_termination_ratio=1e-3)
    print(f"case: early termination divisions={sb.divisions} steps={n}", flush=True)
    got, run = time_parallel_run(tp, n, steps_per_graph=32)
    print("steps run", run)
    window = int(0.05 / dt)
    assert per + window < run[0] < n - 2000     # armed after the gate, stopped a window later
    assert run[1] == n - per                     # the last chunk never falls quiet after its gate
    ref = _whole(sb, n)
    err = rel_l2(got, ref)
    print("early-terminated chunk sum vs whole run rel L2", err)
    assert err <= 1e-3


def test_run_xml_writes_chunk_files(hip_lib, tmp_path):
    """fdtd_run_xml on a time_parallel scene writes what each of the
    reference's chunk processes writes; the chunk streams, shifted to their
    first steps and summed, equal fdtd_run_time_parallel's series."""
    import ctypes as C
    from acoustics_amd import capi
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "tests", "golden", "scene_sphere.xml")).read()
    text = text.replace('output_pattern="out_%s"', f'output_pattern="{tmp_path}/run_%s"')
    text = text.replace("<solver ", '<solver time_parallel="1" num_time_chunks="3" overlap_time="0.0005" ')
    xml = tmp_path / "tp.xml"
    xml.write_text(text)
    print(f"case: run_xml time-parallel {xml}", flush=True)
    n = C.c_int64(0)
    assert hip_lib.fdtd_run_xml(str(xml).encode(), b"0", None, C.byref(n)) == 0, hip_lib.fdtd_last_error()
    p = C.c_void_p()
    assert hip_lib.fdtd_parse_xml(str(xml).encode(), C.byref(p)) == 0
    sc = hip_lib.fdtd_parsed_scene(p).contents
    steps = hip_lib.fdtd_parsed_num_steps(p)
    assert steps == 300
    L = sc.n_listeners
    total = np.zeros((steps, L))
    ran = 0
    for c in range(3):
        d, first, want = time_chunk({k: getattr(sc.settings, k) if k not in ("divisions", "min_bound")
                                     else tuple(getattr(sc.settings, k)) for k, _ in capi.Settings._fields_},
                                    steps, c)
        audio = np.fromfile(str(tmp_path / f"run_{c:05d}_all_audio.dat"), dtype=np.float64).reshape(-1, L)
        assert audio.shape[0] == min(want, steps - first)
        total[first:first + audio.shape[0]] += audio
        ran += audio.shape[0]
        assert float(open(tmp_path / f"run_{c:05d}_start_time").read()) == pytest.approx(d["start_time"], abs=1e-15)
        rng = [float(x) for x in open(tmp_path / f"run_{c:05d}_time_range").read().split()]
        assert rng[1] == pytest.approx(d["stop_boundary_acc_time"], abs=1e-15)
        assert os.path.exists(tmp_path / f"run_{c:05d}_sim-done")
    assert n.value == ran
    got, _ = time_parallel_run(SceneBuffers_from(sc), steps)
    hip_lib.fdtd_parsed_free(p)
    assert np.array_equal(total, got)


def SceneBuffers_from(sc):
    """A SceneBuffers holding a copy of a parsed scene."""
    from acoustics_amd import capi
    settings = {k: getattr(sc.settings, k) if k not in ("divisions", "min_bound") else tuple(getattr(sc.settings, k))
                for k, _ in capi.Settings._fields_}
    objs = []
    for i in range(sc.n_objects):
        o = sc.objects[i]
        objs.append(dict(type=o.type, center=tuple(o.center), semi_axes=tuple(o.semi_axes),
                         rotation=np.array(list(o.rotation)).reshape(3, 3)))
    imps = [dict(object=sc.impulse_sources[i].object, amplitude=sc.impulse_sources[i].amplitude,
                 width=sc.impulse_sources[i].width, onset=sc.impulse_sources[i].onset) for i in range(sc.n_impulse_sources)]
    ps = []
    for i in range(sc.n_pressure_sources):
        q = sc.pressure_sources[i]
        ps.append(dict(position=tuple(q.position), width_time=q.width_time, width_space=q.width_space,
                       offset_time=q.offset_time, normalize_constant=q.normalize_constant))
    lst = np.array([[sc.listeners[3 * i + d] for d in range(3)] for i in range(sc.n_listeners)])
    return SceneBuffers(settings, objs, imps, ps, lst)
