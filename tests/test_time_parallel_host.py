"""Time-parallel chunks on the host (no GPU): the chunk settings follow
SimWorld's set-up (SimWorld.cpp:151-166) and the oracle's source gate
(SimWorld::UpdateObjectState, SimWorld.cpp:468-479) makes the chunks a linear
decomposition of the whole run.

The decisive pin is superposition: the scheme is linear in its sources, so
the chunks' listener series, each shifted to its chunk's first step and
summed, must give the series of the run without chunks (FP64 oracle, to
rounding).  A gate one step early or late drops or doubles one step's
source term and fails it by many orders of magnitude."""
import ctypes as C
import math

import numpy as np
import pytest

from acoustics_amd import capi, scenes
from acoustics_amd.capi import SceneBuffers
from acoustics_amd.solver import time_chunk
from oracle.oracle import Oracle


def _settings(k, overlap, div=(24, 26, 28)):
    s = scenes.base_settings(div, 0.005, 5)
    s.update(time_parallel=1, num_time_chunks=k, overlap_time=overlap, early_termination_ratio=1e-3)
    return s


def test_chunk_settings_follow_simworld():
    s = _settings(3, 6e-3)
    dt = s["time_step"]
    n = 1000
    per = n // 3
    ends = []
    for c in range(3):
        d, first, steps = time_chunk(s, n, c)
        t = s["start_time"]
        for _ in range(per * c):
            t += dt
        assert d["start_time"] == t and first == per * c
        for _ in range(per):
            t += dt
        assert d["stop_boundary_acc_time"] == t
        assert steps == int(per + 6e-3 / dt)
        ends.append(first + per)
    assert ends[-1] == 3 * per <= n
    lib = capi.load_library()
    bad = capi.SceneBuffers(dict(s, time_parallel=0)).scene.settings
    out = capi.Settings()
    assert lib.fdtd_time_chunk(C.byref(bad), n, 0, C.byref(out), None, None) == 1
    ok = capi.SceneBuffers(s).scene.settings
    assert lib.fdtd_time_chunk(C.byref(ok), n, 3, C.byref(out), None, None) == 1


def _scene(settings):
    rng = np.random.default_rng(71)
    lo, hi = scenes.interior_bounds(settings, 2)
    src = scenes.gaussian_pressure_source(rng.uniform(lo, hi), 5e-5, offset=2e-4)
    listeners = rng.uniform(lo, hi, size=(5, 3))
    return SceneBuffers(settings, pressure_sources=[src], listeners=listeners)


def test_chunks_superpose_to_the_whole_run():
    n, k = 180, 3
    s = _settings(k, 1.0)              # the overlap reaches past the end of the run
    whole = Oracle(_scene(dict(s, time_parallel=0)), reclassify_each_step=False).step(n)
    total = np.zeros_like(whole)
    for c in range(k):
        d, first, steps = time_chunk(s, n, c)
        part = Oracle(_scene(d), reclassify_each_step=False).step(n - first)
        total[first:] += part
        if c == 0:
            # before the gate a chunk is the whole run; after it, the source is gone
            assert np.array_equal(part[:n // k], whole[:n // k])
    err = np.linalg.norm(total - whole) / np.linalg.norm(whole)
    print("superposition rel L2", err)
    assert err <= 1e-12


def test_gate_is_exact_to_the_step():
    """A chunk whose stop time sits one step later than another's differs from
    it by exactly one step's source term, first seen in the row of that step."""
    n = 60
    s = _settings(2, 1.0)
    d, first, _ = time_chunk(s, n, 0)
    a = Oracle(_scene(d), reclassify_each_step=False).step(n)
    d2 = dict(d, stop_boundary_acc_time=d["stop_boundary_acc_time"] + s["time_step"])
    b = Oracle(_scene(d2), reclassify_each_step=False).step(n)
    assert np.array_equal(a[:n // 2], b[:n // 2])
    assert not np.array_equal(a[n // 2], b[n // 2])
