"""The XML scene config path (fdtd_create_from_xml) and the C++ facade on
the device, checked against the FP64 oracle on the same parsed scene."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from acoustics_amd import capi
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XML = os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")


class _ParsedScene:
    """Adapts a parsed fdtd_scene to what the oracle binding reads."""

    def __init__(self, lib, path):
        self.lib = lib
        self.p = C.c_void_p()
        capi.check(lib.fdtd_parse_xml(path.encode(), C.byref(self.p)), "fdtd_parse_xml")
        self.sc = lib.fdtd_parsed_scene(self.p)
        s = self.sc.contents
        self.n_listeners = s.n_listeners
        self.divisions = tuple(s.settings.divisions)
        self.n_cells = int(np.prod(self.divisions))
        self.steps = lib.fdtd_parsed_num_steps(self.p)

    def ref(self):
        return self.sc


def test_xml_scene_matches_oracle(hip_lib):
    ps = _ParsedScene(hip_lib, XML)
    print("xml scene", ps.divisions, ps.steps, flush=True)
    ref = Oracle(ps, reclassify_each_step=False).step(ps.steps)
    with Solver.from_xml(XML) as s:
        s.step(ps.steps)
        got = s.series()
    assert rel_l2(got, ref) <= 1e-4


def test_facade_drives_the_engine(hip_lib, tmp_path):
    exe = tmp_path / "facade_probe"
    libdir = os.path.join(ROOT, "acoustics_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "facade", "facade_probe.cpp"),
                    "-L" + libdir, "-lfdtd_acoustics", "-Wl,-rpath," + libdir], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "facade cells=32768 divs=32,32,32" in r.stdout
    # velocity field 0: one more layer along x, starting h/2 below the box
    assert "velocity0 divs=33,32,32 first=-0.080000,-0.077500,-0.077500" in r.stdout
    assert "dim 0 refused" in r.stdout
    r2 = subprocess.run([str(exe), XML, "boxed"], capture_output=True, text=True, timeout=120)
    print(r2.stdout)
    assert r2.returncode == 0 and "xml steps=300" in r2.stdout
    assert "name=out/probe_all_audio.dat" in r2.stdout
    assert "boxed cells=46656" in r2.stdout and "rows=" in r2.stdout
    last = float(r2.stdout.split("last=")[1].split()[0])
    ps = _ParsedScene(hip_lib, XML)
    ref = Oracle(ps, reclassify_each_step=False).step(ps.steps)
    assert abs(last - ref[-1, -1]) <= 1e-4 * np.abs(ref).max()
