"""Host-side checks of the reference tool's file formats (no GPU):
output names as FDTD_AcousticSimulator::CompositeFilename
(src/wavesolver/FDTD_AcousticSimulator.cpp:110-124), binary matrices as
IO::writeMatrixX BINARY (src/utils/IO/IO.h:107-139) read back the way
scripts/binary_io.py:readMatrixXdBinary does (int32 rows, int32 cols, doubles
in row order), and the XML reader's output_pattern (ImpulseResponseParser.cpp:521)."""
import ctypes as C
import os

import numpy as np

from acoustics_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_matrix_xd_binary(path):
    """The reader of scripts/binary_io.py, restated with numpy."""
    with open(path, "rb") as f:
        n1 = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        n2 = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        data = np.fromfile(f, dtype=np.float64, count=n1 * n2)
    return data.reshape(n1, n2)


def composite(lib, pattern, sim_id, name):
    buf = C.create_string_buffer(512)
    st = lib.fdtd_composite_filename(pattern.encode() if pattern is not None else None,
                                     sim_id.encode() if sim_id is not None else None, name.encode(), buf, 512)
    assert st == 0
    return buf.value.decode()


def test_composite_filename():
    lib = capi.load_library()
    assert composite(lib, "out/test_%s", "3", "listening_position.dat") == "out/test_3_listening_position.dat"
    assert composite(lib, "out/test_%s", None, "all_audio.dat") == "out/test_all_audio.dat"
    assert composite(lib, "%s", "", "solver_settings.txt") == "solver_settings.txt"
    small = C.create_string_buffer(4)
    assert lib.fdtd_composite_filename(b"%s", None, b"all_audio.dat", small, 4) == 1


def test_binary_matrix_round_trip(tmp_path):
    lib = capi.load_library()
    a = np.arange(15, dtype=np.float64).reshape(5, 3) * 0.25 - 1.0
    path = str(tmp_path / "m.dat")
    assert lib.fdtd_write_matrix(path.encode(), a.ctypes.data_as(C.POINTER(C.c_double)), 5, 3) == 0
    assert os.path.getsize(path) == 8 + 15 * 8
    b = read_matrix_xd_binary(path)
    assert b.shape == (5, 3) and np.array_equal(a, b)


def test_xml_output_pattern_and_scheme(tmp_path):
    lib = capi.load_library()
    p = C.c_void_p()
    path = os.path.join(ROOT, "tests", "golden", "scene_sphere.xml").encode()
    assert lib.fdtd_parse_xml(path, C.byref(p)) == 0
    assert lib.fdtd_parsed_output_pattern(p) == b"out_%s"
    assert lib.fdtd_parsed_scene(p).contents.settings.scheme == capi.FDTD_SCHEME_COLLOCATED
    lib.fdtd_parsed_free(p)
    text = open(path.decode()).read().replace('use_mesh="1"', 'use_mesh="1" scheme="staggered"')
    q = tmp_path / "s.xml"
    q.write_text(text)
    assert lib.fdtd_parse_xml(str(q).encode(), C.byref(p)) == 0
    assert lib.fdtd_parsed_scene(p).contents.settings.scheme == capi.FDTD_SCHEME_STAGGERED
    lib.fdtd_parsed_free(p)
    q.write_text(text.replace('scheme="staggered"', 'scheme="leapfrog"'))
    assert lib.fdtd_parse_xml(str(q).encode(), C.byref(p)) == 5


def test_xml_time_parallel_keys(tmp_path):
    """time_parallel, num_time_chunks, overlap_time and early_termination_ratio
    (ImpulseResponseParser.cpp:525-533), with the reference's defaults."""
    lib = capi.load_library()
    text = open(os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")).read()
    p = C.c_void_p()
    q = tmp_path / "tp.xml"
    q.write_text(text.replace("<solver ", '<solver time_parallel="1" num_time_chunks="5" '))
    assert lib.fdtd_parse_xml(str(q).encode(), C.byref(p)) == 0
    s = lib.fdtd_parsed_scene(p).contents.settings
    assert (s.time_parallel, s.num_time_chunks, s.overlap_time, s.early_termination_ratio) == (1, 5, 6e-3, 1e-3)
    lib.fdtd_parsed_free(p)
    q.write_text(text.replace("<solver ", '<solver time_parallel="1" num_time_chunks="2" overlap_time="0.001" '
                                          'early_termination_ratio="0.01" '))
    assert lib.fdtd_parse_xml(str(q).encode(), C.byref(p)) == 0
    s = lib.fdtd_parsed_scene(p).contents.settings
    assert (s.num_time_chunks, s.overlap_time, s.early_termination_ratio) == (2, 0.001, 0.01)
    lib.fdtd_parsed_free(p)
    q.write_text(text.replace("<solver ", '<solver time_parallel="1" '))
    assert lib.fdtd_parse_xml(str(q).encode(), C.byref(p)) == 5       # num_time_chunks is required
    p = C.c_void_p()
    assert lib.fdtd_parse_xml(os.path.join(ROOT, "tests", "golden", "scene_sphere.xml").encode(), C.byref(p)) == 0
    assert lib.fdtd_parsed_scene(p).contents.settings.time_parallel == 0
    lib.fdtd_parsed_free(p)
