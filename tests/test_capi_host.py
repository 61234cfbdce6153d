"""Host-side checks of the C-ABI library (no GPU compute calls): every
symbol include/fdtd_acoustics.h declares is exported, the XML scene reader
follows the ImpulseResponseParser keys and defaults, errors come back as
status codes, and the C++ facade links and reports them."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from acoustics_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fdtd_acoustics.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(fdtd_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = capi.load_library()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_status_strings_and_defaults():
    lib = capi.load_library()
    assert lib.fdtd_status_string(0) == b"ok"
    assert lib.fdtd_status_string(6) == b"unsupported"
    s = capi.Settings()
    lib.fdtd_settings_default(C.byref(s))
    assert s.sound_speed == 343.0 and s.density == 1.0 and s.use_mesh == 1 and s.boundary_handling == 0


def test_xml_scene_parse():
    lib = capi.load_library()
    p = C.c_void_p()
    path = os.path.join(ROOT, "tests", "golden", "scene_sphere.xml").encode()
    assert lib.fdtd_parse_xml(path, C.byref(p)) == 0
    sc = lib.fdtd_parsed_scene(p).contents
    st = sc.settings
    assert lib.fdtd_parsed_num_steps(p) == 300
    assert st.cell_size == 0.005 and st.density == 1.2 and st.pml_width == 8
    assert abs(st.time_step - 1 / 237637.6) < 1e-18
    assert list(st.divisions) == [40, 40, 40]
    assert abs(st.min_bound[0] + 0.1) < 1e-15
    assert sc.n_objects == 1 and sc.objects[0].type == capi.FDTD_OBJ_SPHERE and sc.objects[0].semi_axes[0] == 0.025
    assert sc.n_impulse_sources == 1 and sc.impulse_sources[0].object == 0
    assert sc.n_pressure_sources == 1
    ps = sc.pressure_sources[0]
    # parser defaults: width_space = c * width_time, K = 1/(sqrt(2 pi) ws)^3
    assert abs(ps.width_space - 343 * 2e-4) < 1e-15
    assert abs(ps.normalize_constant - 1 / (np.sqrt(2 * np.pi) * ps.width_space) ** 3) < 1e-9 * ps.normalize_constant
    assert sc.n_listeners == 2 and sc.listeners[3] == -0.04
    lib.fdtd_parsed_free(p)


def test_xml_errors(tmp_path):
    lib = capi.load_library()
    p = C.c_void_p()
    bad = tmp_path / "bad.xml"
    bad.write_text('<impulse_response><solver cellsize="0.005"/></impulse_response>')
    assert lib.fdtd_parse_xml(str(bad).encode(), C.byref(p)) == 5  # FDTD_ERR_PARSE
    assert b"stop_time" in lib.fdtd_last_error()
    bad.write_text('<impulse_response><solver')
    assert lib.fdtd_parse_xml(str(bad).encode(), C.byref(p)) == 5
    assert lib.fdtd_parse_xml(b"/nonexistent.xml", C.byref(p)) == 5


def test_create_without_device_reports_status():
    """No GPU in this container: creation must fail with a status, not abort."""
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    from acoustics_amd import scenes
    lib = capi.load_library()
    sb, _ = scenes.config1(n=16)
    h = C.c_void_p()
    st = lib.fdtd_create(sb.ref(), None, C.byref(h))
    assert st == 2  # FDTD_ERR_NO_DEVICE


def test_facade_compiles_and_links(tmp_path):
    exe = tmp_path / "facade_probe"
    src = os.path.join(ROOT, "tests", "facade", "facade_probe.cpp")
    libdir = os.path.join(ROOT, "acoustics_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), src, "-L" + libdir, "-lfdtd_acoustics",
                    "-Wl,-rpath," + libdir], check=True)
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present; the probe runs in the GPU tests")
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 3 and "no device" in r.stdout


def test_noalloc_size_query_and_scratch_limit():
    """fdtd_solve_noalloc's host part: with fields NULL it only reports the
    state bytes (the two pressure slots plus the PML state); a scene whose
    tables exceed the 64 MiB scratch array is refused before any device call."""
    from acoustics_amd import scenes
    lib = capi.load_library()
    sb, _ = scenes.config1(seed=3, n=32)
    need = C.c_int64(0)
    assert lib.fdtd_solve_noalloc(sb.ref(), 10, None, 0, None, None, C.byref(need)) == 0
    n_pad = (32 * 32 * 32 + 63) // 64 * 64
    assert need.value >= 2 * 4 * n_pad
    # 416^3 with an object: the per-cell skip mask alone exceeds 64 MiB
    big, _ = scenes.config2(seed=3, n=416, pml=8, radius=0.05, n_listeners=4)
    need_big = C.c_int64(0)
    assert lib.fdtd_solve_noalloc(big.ref(), 1, None, 0, None, None, C.byref(need_big)) == 0
    st = lib.fdtd_solve_noalloc(big.ref(), 1, C.c_void_p(256), need_big.value, None,
                                (C.c_double * 4)(), None)
    assert st == 6
    assert b"scratch" in lib.fdtd_last_error()
    # too small a state buffer
    st = lib.fdtd_solve_noalloc(sb.ref(), 1, C.c_void_p(256), need.value - 4, None, (C.c_double * 8)(), None)
    assert st == 1


def test_xml_plane_constraints(tmp_path):
    """<constraint> elements of <scene> (ImpulseResponseParser.cpp:434-449):
    direction, sign and height, defaulting to 1, 1 and 0."""
    lib = capi.load_library()
    src = open(os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")).read()
    src = src.replace("<scene>", '<scene>\n    <constraint id="floor" direction="1" sign="1" height="-0.05"/>\n'
                                 '    <constraint direction="2" sign="-1" height="0.07"/>\n    <constraint/>')
    path = tmp_path / "floor.xml"
    path.write_text(src)
    p = C.c_void_p()
    assert lib.fdtd_parse_xml(str(path).encode(), C.byref(p)) == 0
    sc = lib.fdtd_parsed_scene(p).contents
    assert sc.n_constraints == 3
    got = [(sc.constraints[i].direction, sc.constraints[i].sign, sc.constraints[i].height) for i in range(3)]
    assert got == [(1, 1, -0.05), (2, -1, 0.07), (1, 1, 0.0)]
    assert sc.n_objects == 1
    lib.fdtd_parsed_free(p)


def test_xml_harmonic_and_acc_noise_sources(tmp_path):
    """has_harmonic_source / harmonic_source_f on an object
    (ImpulseResponseParser.cpp:266-275: omega = 2 pi f, phase 0, unit
    pulsation velocity, no decay) and an inline acceleration-noise source
    whose inverse inertia defaults to the uniform solid of the object's shape."""
    lib = capi.load_library()
    src = open(os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")).read()
    src = src.replace('radius="0.025">', 'radius="0.025" has_harmonic_source="1" harmonic_source_f="1500" '
                      'has_chirp_source="1" chirp_frequency_0="1000" chirp_frequency_1="3000" chirp_time_0="1e-4" '
                      'chirp_time_1="9e-4" chirp_amplitude_0="1" chirp_amplitude_1="1">\n'
                      '      <acc_noise_source mass="0.2" center_of_mass_z="0.001">\n'
                      '        <impact timestamp="1e-4" support="2e-4" impulse_x="0.001" impulse_y="0" impulse_z="-0.002"'
                      ' position_x="0.025"/>\n'
                      '        <impact timestamp="3e-4" support="1e-4" impulse_x="0" impulse_y="0.003" impulse_z="0"/>\n'
                      '      </acc_noise_source>')
    path = tmp_path / "vib.xml"
    path.write_text(src)
    p = C.c_void_p()
    assert lib.fdtd_parse_xml(str(path).encode(), C.byref(p)) == 0, capi.last_error()
    sc = lib.fdtd_parsed_scene(p).contents
    kinds = sorted(sc.impulse_sources[i].kind for i in range(sc.n_impulse_sources))
    assert kinds == [capi.FDTD_VIB_GAUSSIAN, capi.FDTD_VIB_HARMONIC, capi.FDTD_VIB_CHIRP]
    ch = [sc.impulse_sources[i] for i in range(sc.n_impulse_sources) if sc.impulse_sources[i].kind == capi.FDTD_VIB_CHIRP][0]
    assert abs(ch.omega - 2 * np.pi * 1000) < 1e-9 and abs(ch.omega1 - 2 * np.pi * 3000) < 1e-9
    assert ch.t0 == 1e-4 and ch.t1 == 9e-4
    hs = [sc.impulse_sources[i] for i in range(sc.n_impulse_sources) if sc.impulse_sources[i].kind == capi.FDTD_VIB_HARMONIC][0]
    assert hs.object == 0 and abs(hs.omega - 2 * np.pi * 1500) < 1e-9 and hs.phase == 0 and hs.amplitude == 1.0
    assert hs.decay_rate < 0
    assert sc.n_acc_noise_sources == 1
    an = sc.acc_noise_sources[0]
    assert an.object == 0 and an.mass == 0.2 and list(an.center_of_mass) == [0.0, 0.0, 0.001]
    inv = np.array(list(an.inv_inertia)).reshape(3, 3)
    assert np.allclose(inv, np.eye(3) / (0.4 * 0.2 * 0.025 ** 2), rtol=1e-12)
    assert an.n_impacts == 2
    assert an.impacts[0].timestamp == 1e-4 and an.impacts[0].support == 2e-4
    assert list(an.impacts[0].impulse) == [0.001, 0.0, -0.002] and list(an.impacts[0].position) == [0.025, 0.0, 0.0]
    assert list(an.impacts[1].impulse) == [0.0, 0.003, 0.0]
    lib.fdtd_parsed_free(p)


def test_acc_noise_refusals():
    """Acceleration noise runs on static objects in the collocated scheme."""
    from acoustics_amd import scenes
    lib = capi.load_library()
    sb, _ = scenes.vibration_scene(n=32, pml=4, harmonic=False)
    assert lib.fdtd_scene_device_bytes(sb.ref()) > 0
    objs = [dict(o) for o in sb.objects_list]
    objs[0]["velocity"] = (1.0, 0.0, 0.0)
    moving = capi.SceneBuffers(sb.settings_dict, objects=objs, listeners=sb.listeners_np,
                               acc_noise_sources=sb.acc_noise_list)
    assert lib.fdtd_scene_device_bytes(moving.ref()) == -1
    assert "static" in capi.last_error()


def test_xml_kinematics_file(tmp_path):
    """file_kinematics / kinematics_frame_rate on an object
    (ImpulseResponseParser.cpp:248-263): rows frame tx ty tz qw qx qy qz read
    next to the XML file, step 1 / rate."""
    lib = capi.load_library()
    (tmp_path / "swing.txt").write_text("0 0.0 0.01 0.0 1 0 0 0\n\n1 0.001 0.01 0.0 0.9 0.1 0.2 0.3\n"
                                        "2 0.002 0.011 -0.001 0.8 0.2 0.3 0.4\n")
    src = open(os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")).read()
    src = src.replace('radius="0.025">', 'radius="0.025" file_kinematics="swing.txt" kinematics_frame_rate="2000">')
    path = tmp_path / "kin.xml"
    path.write_text(src)
    p = C.c_void_p()
    assert lib.fdtd_parse_xml(str(path).encode(), C.byref(p)) == 0, capi.last_error()
    sc = lib.fdtd_parsed_scene(p).contents
    assert sc.n_kinematics == 1
    k = sc.kinematics[0]
    assert k.object == 0 and k.n_frames == 3 and k.step_size == 1 / 2000
    fr = np.array([k.frames[i] for i in range(21)]).reshape(3, 7)
    assert np.array_equal(fr[1], [0.001, 0.01, 0.0, 0.9, 0.1, 0.2, 0.3])
    lib.fdtd_parsed_free(p)
    (tmp_path / "swing.txt").write_text("0 0.0 0.01 0.0 1 0 0\n")
    assert lib.fdtd_parse_xml(str(path).encode(), C.byref(p)) == 5   # FDTD_ERR_PARSE: a short row
    assert "row needs" in capi.last_error()
