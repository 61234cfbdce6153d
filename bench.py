#!/usr/bin/env python3
"""Benchmark of the FDTD time-stepping hot path on one MI355X.

Workload (one "step" = one collocated time step of the whole grid):
256^3 MAC grid, h = 2.5 mm, 16-cell PML, five rigid ellipsoids with seeded
Gaussian-impulse boundary accelerations (rasterized ghost cells, FP64),
128 listeners (FP64 MLS), FP32 fields -- the 256^3 scene of the engine's
headline metric.  Inputs are resident in HBM before timing starts.

Prints ONE JSON line:
  value        whole-job Gcell-updates/s = N * steps / wall time
  roofline     the step kernel: algorithmic bytes (12 N + 44 N_pml per step)
               / device time per step from two HIP events recorded on the
               kernel's stream around the timed region (the step kernels run
               back to back, with one block advance per 64 steps: an upper
               bound on the kernel's mean duration); traffic from the
               committed counter summary (profiles/step_kernel_256cube_traffic.json);
               state_bytes / state_frac: the same for the bytes the engine's
               merged PML state moves (12 N + 44 N_pml - 24 N_merged)
  staggered    the reference's staggered (MAC) scheme on a 256^3 free field
  engquist     a 256^3 free field without PML (the Engquist boundary)
  moving       a 256^3 scene with a moving, spinning ellipsoid, and the same
               ellipsoid held still
  batch        64 uneven scenes advanced together vs one at a time
  coupled      the same scene with fully-coupled ghost cells
  serving      a Poisson request stream (200 requests): continuous batching vs
               the plain one-request-at-a-time path (all steps in one call)
  cpu_baseline FP64 CPU oracle (OpenMP, all host threads granted; plus one
               thread) on bounded samples of the three single-scene workloads
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "FDTD Gcell-updates/s (%HBM peak); served solves/s + p99 latency vs plain"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernel_256cube_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=256, help="grid edge (256 is the headline scene)")
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=120)
    ap.add_argument("--serve", type=int, default=200, help="requests in the serving comparison (0: skip)")
    ap.add_argument("--serve-per-step", action="store_true",
                    help="also time the plain path in the reference's per-step driver pattern")
    ap.add_argument("--block", type=int, default=64, help="server steps per block")
    ap.add_argument("--staggered-steps", type=int, default=200,
                    help="steps of the staggered-scheme measurement (0: skip)")
    ap.add_argument("--batch", type=int, default=64, help="scenes in the batched measurement (0: skip)")
    ap.add_argument("--other-steps", type=int, default=200,
                    help="steps of the Engquist-boundary and moving-object measurements (0: skip)")
    return ap.parse_args()


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads(n):
    """Sets the OpenMP thread count of the calling thread (the oracle's loops)."""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass


def cpu_baseline(steps3):
    """FP64 oracle (the CPU restatement of the reference path, OpenMP) timed
    over its stepping loop on bounded samples of the three single-scene
    workloads: 64^3 free field (500 steps, all), 128^3 sphere (200 of 2000
    steps), 256^3 five ellipsoids (steps3 of 4000), each with every host
    thread this process may use, plus a 1-thread run of the 128^3 sample.
    Per-step reclassification as in the reference; the value is the 256^3
    rate."""
    from acoustics_amd import scenes
    from oracle.oracle import Oracle, build
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    nproc = os.cpu_count() or 1
    if threads <= 0:
        threads = nproc
    build()

    def rate(sb, steps, nthreads):
        _omp_threads(nthreads)
        o = Oracle(sb, reclassify_each_step=True)
        o.step(1, record=False)
        t0 = time.perf_counter()
        o.step(steps, record=False)
        t = time.perf_counter() - t0
        o.close()
        return sb.n_cells * steps / t / 1e9, t

    out = {}
    samples = [("cfg1_64cube_free_field", scenes.config1(seed=1)[0], 500),
               ("cfg2_128cube_sphere", scenes.config2(seed=2)[0], 200),
               ("cfg3_256cube_ellipsoids", scenes.config3(seed=3)[0], steps3)]
    for name, sb, steps in samples:
        g, t = rate(sb, steps, threads)
        out[name] = {"Gcell_per_s": round(g, 4), "steps": steps, "seconds": round(t, 2)}
    g1, t1 = rate(samples[1][1], 20, 1)
    _omp_threads(threads)
    return {"value": out["cfg3_256cube_ellipsoids"]["Gcell_per_s"], "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"256^3 headline scene, {steps3} FP64 steps after 1 warmup step (of 4000), per-step "
                      f"reclassification as in the reference, {threads} OpenMP threads",
            "nproc": nproc, "cpu_model": _cpu_model(), "configs": out,
            "single_thread": {"workload": "cfg2_128cube_sphere", "steps": 20, "Gcell_per_s": round(g1, 5),
                              "seconds": round(t1, 2)}}


def staggered(steps, n=256):
    """The reference's staggered (MAC) scheme on an n^3 free-field scene (h =
    0.64/n m, 16-cell PML, one Gaussian pressure pulse, 8 listeners): whole-job
    Gcell-updates/s and the step kernel against HBM peak.  Algorithmic bytes
    per step for that scheme: 32 N + 24 N_pml."""
    from acoustics_amd import scenes
    from acoustics_amd.solver import Solver
    sb, _ = scenes.staggered_free_field(seed=7, div=(n, n, n), h=0.64 / n, pml=16, n_listeners=8)
    from acoustics_amd import capi
    with Solver(sb, series_capacity=max(1024, steps + 64)) as s:
        info = s.info()
        s.set_timing(capi.FDTD_TIMING_SPAN)
        s.step(20)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        _, step_ms = s.last_kernel_ms()
        kern_ms = step_ms
        finite = bool(np.isfinite(s.series()).all())
    ach = info["bytes_per_step"] / (kern_ms * 1e-3) / 1e9
    return {"workload": f"{n}^3 free field, staggered MAC scheme, 16-cell PML, 1 pressure source, 8 listeners",
            "value": round(info["n_cells"] * steps / wall / 1e9, 3), "unit": "Gcell-updates/s", "steps": steps,
            "bytes_per_step": info["bytes_per_step"], "kernel_ms": round(kern_ms, 5), "step_ms": round(step_ms, 5),
            "achieved_GBps": round(ach, 1), "frac": round(ach / HBM_PEAK_GBPS, 4), "listeners_finite": finite}


def _timed(sb, steps, warm=20):
    from acoustics_amd import capi
    from acoustics_amd.solver import Solver
    with Solver(sb, series_capacity=max(1024, steps + warm + 64)) as s:
        info = s.info()
        s.set_timing(capi.FDTD_TIMING_SPAN)
        s.step(warm)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        _, step_ms = s.last_kernel_ms()
        finite = bool(np.isfinite(s.series()).all())
    return info, wall, step_ms, finite


def coupled(steps, n=256):
    """The headline scene with fully-coupled ghost cells (MAC_Grid.cpp:
    1312-1425): per step, the coupled system's static solution operator
    (one FP64 pass over the ghost cells) runs before the step kernel."""
    from acoustics_amd import scenes
    from acoustics_amd.capi import FDTD_BOUNDARY_FULLY_COUPLED
    sb, _ = scenes.config3(seed=3, n=n, h=0.64 / n, boundary=FDTD_BOUNDARY_FULLY_COUPLED)
    info, wall, step_ms, finite = _timed(sb, steps)
    return {"workload": f"{n}^3 headline scene with fully-coupled ghost cells ({info['n_ghost_cells']} ghost cells)",
            "value": round(info["n_cells"] * steps / wall / 1e9, 3), "unit": "Gcell-updates/s", "steps": steps,
            "step_ms": round(step_ms, 5), "listeners_finite": finite}


def other_paths(steps, n=256):
    """Two more paths at n^3 (h = 0.64/n): the Engquist boundary on a
    free field without PML (one pressure pulse, 8 listeners), and a moving
    ellipsoid (10 m/s, spinning at 200 rad/s, 16-cell PML) beside the same
    ellipsoid held still.  Whole-job Gcell-updates/s and device time per step
    (HIP events around the timed steps; for the moving scene this includes
    the per-step motion kernel and the moving-object tiles)."""
    from acoustics_amd import scenes
    from acoustics_amd.capi import SceneBuffers
    warm = 20
    out = {}
    sb, _ = scenes.engquist_free_field(seed=8, div=(n, n, n), h=0.64 / n, n_listeners=8)
    info, wall, step_ms, finite = _timed(sb, steps, warm)
    out["engquist"] = {"workload": f"{n}^3 free field, no PML (Engquist boundary on layer 1), 1 pressure source, 8 listeners",
                       "value": round(info["n_cells"] * steps / wall / 1e9, 3), "unit": "Gcell-updates/s",
                       "steps": steps, "step_ms": round(step_ms, 5), "listeners_finite": finite}
    mv, _ = scenes.moving_object(seed=13, n=n, h=0.64 / n, pml=16, steps=steps + warm + 1, kind="ellipsoid",
                                 speed=10.0, spin=200.0)
    info, wall, step_ms, finite = _timed(mv, steps, warm)
    still = SceneBuffers(mv.settings_dict, [dict(o, velocity=(0.0, 0.0, 0.0), angular_velocity=(0.0, 0.0, 0.0))
                                            for o in mv.objects_list], mv.impulse_list, mv.pressure_list, mv.listeners_np)
    _, wall0, step0_ms, _ = _timed(still, steps, warm)
    out["moving"] = {"workload": f"{n}^3 MAC grid, 16-cell PML, one ellipsoid moving at 10 m/s and spinning at "
                                 f"200 rad/s (per-step reclassification and fresh cells), 8 listeners",
                     "value": round(info["n_cells"] * steps / wall / 1e9, 3), "unit": "Gcell-updates/s",
                     "steps": steps, "step_ms": round(step_ms, 5), "step_ms_static": round(step0_ms, 5),
                     "value_static": round(info["n_cells"] * steps / wall0 / 1e9, 3), "listeners_finite": finite}
    return out


def batched(n_scenes):
    """Uneven batch (edges U{32..128} per axis, 8-cell PML, one sphere or
    ellipsoid with an impulse, 16 listeners, 1000-3000 steps each): all scenes
    advanced together (one launch per step, 32-step graphs, autotuned tile
    plan) against the same scenes solved one after another with the
    single-scene API (all steps in one call each)."""
    from acoustics_amd import scenes
    from acoustics_amd.solver import Batch, Solver
    sbs, steps = scenes.config4(seed=4, n_scenes=n_scenes)
    cells = sum(sb.n_cells * st for sb, st in zip(sbs, steps))

    def step_bytes(sb):   # 12 N + 44 N_pml with the 8-cell PML on every face
        d = np.asarray(sb.divisions)
        return 12 * int(np.prod(d)) + 44 * (int(np.prod(d)) - int(np.prod(d - 16)))
    alg_bytes = sum(step_bytes(sb) * st for sb, st in zip(sbs, steps))
    t0 = time.perf_counter()
    with Batch(sbs, steps) as b:
        t1 = time.perf_counter()
        b.run()
        ok = all(bool(np.isfinite(b.series(i)).all()) for i in range(len(sbs)))
    t_batch = time.perf_counter() - t1
    t_batch_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    for sb, st in zip(sbs, steps):
        with Solver(sb) as s:
            s.step(st)
            s.series()
    t_plain = time.perf_counter() - t0
    return {"scenes": len(sbs), "cell_updates": cells,
            "batched_Gcell_per_s": round(cells / t_batch / 1e9, 3),
            "batched_wall_s": round(t_batch, 3), "batched_wall_with_setup_s": round(t_batch_all, 3),
            "one_at_a_time_Gcell_per_s": round(cells / t_plain / 1e9, 3), "one_at_a_time_wall_s": round(t_plain, 3),
            "speedup": round(t_plain / t_batch_all, 3), "listeners_finite": ok,
            "roofline": {"bound": "hbm", "algorithmic_bytes": alg_bytes,
                         "achieved": round(alg_bytes / t_batch / 1e9, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": round(alg_bytes / t_batch / 1e9 / HBM_PEAK_GBPS, 4),
                         "source": "sum over scenes of (12 N + 44 N_pml) x steps / batched wall time"}}


def _percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def serving(n_requests, block, per_step=False):
    """Poisson stream of impact-sound requests (rate 200/s, edges 48-192,
    1-3 ellipsoids, 32 listeners, 0.01-0.05 s of sound), served three ways:
      served      the continuous-batching server (all admitted requests share
                  each step's launch, graph-captured blocks of `block` steps);
      plain       one request at a time through the single-scene API in the
                  reference's own driver pattern: advance one step, then fetch
                  the listener pressures (SimWorld::StepWorld calls
                  stepSystem() and FetchPressureData() every step);
      plain_bulk  one request at a time, all steps in one call, series read
                  once at the end.
    Latency = finish - arrival."""
    import threading
    from acoustics_amd import scenes
    from acoustics_amd.solver import Server, Solver
    stream = scenes.config5_stream(seed=5, n_requests=n_requests)
    cells_steps = sum(sb.n_cells * st for _, sb, st in stream)

    # served
    lat = {}
    with Server(hbm_budget_bytes=160 << 30, steps_per_block=block) as srv:
        ids = [None] * len(stream)
        t0 = time.perf_counter()

        def submitter():
            for q, (ta, sb, st) in enumerate(stream):
                d = t0 + ta - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[q] = (srv.submit(sb, st), time.perf_counter())
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for q in range(len(stream)):
            rid, ts = ids[q]
            srv.wait(rid, 600000)
            status, series, el = srv.take(rid)
            if status != 0:
                raise RuntimeError(f"served request {q} failed with status {status}")
            lat[q] = el
        t_served = time.perf_counter() - t0
    served_lat = [lat[q] for q in range(len(stream))]

    def one_at_a_time(per_step):
        lats = []
        t0 = time.perf_counter()
        for ta, sb, st in stream:
            d = t0 + ta - time.perf_counter()
            if d > 0:
                time.sleep(d)
            with Solver(sb) as s:
                if per_step:
                    for _ in range(st):
                        s.step(1)
                        s.listen()
                else:
                    s.step(st)
                    s.series()
            lats.append((time.perf_counter() - (t0 + ta)) * 1e3)
        return lats, time.perf_counter() - t0

    def summary(lats, wall):
        return {"solves_per_s": round(len(stream) / wall, 3), "p50_ms": round(_percentile(lats, 50), 2),
                "p99_ms": round(_percentile(lats, 99), 2), "wall_s": round(wall, 3)}

    bulk_lat, t_bulk = one_at_a_time(False)
    out = {
        "requests": len(stream), "cell_updates": cells_steps, "steps_per_block": block,
        "served": summary(served_lat, t_served),
        "plain_bulk": summary(bulk_lat, t_bulk),
        "speedup_vs_plain_bulk": round(t_bulk / t_served, 3),
        "served_Gcell_per_s": round(cells_steps / t_served / 1e9, 2),
    }
    if per_step:
        plain_lat, t_plain = one_at_a_time(True)
        out["plain_per_step"] = summary(plain_lat, t_plain)
        out["speedup_vs_plain_per_step"] = round(t_plain / t_served, 3)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)
    from acoustics_amd import capi, scenes
    from acoustics_amd.solver import Solver

    n = a.n
    sb, _ = scenes.config3(seed=3, n=n, h=0.64 / n)
    s = Solver(sb, tile_config=a.tile, series_capacity=max(4096, a.steps + a.warmup + 8))
    info = s.info()
    N, Npml = info["n_cells"], info["n_pml_cells"]
    bytes_step = info["bytes_per_step"]
    state_bytes = info["state_bytes_per_step"]
    # timed region: K steps back to back on the solver's stream, bracketed by
    # two HIP events on that stream (events between the kernels would add a
    # few microseconds to every step).  The device time per step is the step
    # kernel's duration plus one block advance per 64 steps and the launch
    # gaps, an upper bound on the kernel's average duration.
    s.set_timing(capi.FDTD_TIMING_SPAN)
    s.step(a.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(a.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    _, kern_ms = s.last_kernel_ms()
    series = s.series()
    finite = bool((series == series).all())
    s.close()

    value = N * a.steps / wall / 1e9
    achieved = bytes_step / (kern_ms * 1e-3) / 1e9 if kern_ms > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            tf = json.load(open(TRAFFIC_FILE))
            if tf.get("n") == n:
                traffic = tf.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields; f64 ghost cells and listeners",
        "data": "synthetic (seeded scene)",
        "config": {"workload": f"{n}^3 MAC grid, h={0.64 / n * 1e3:.2f} mm, 16-cell PML, 5 rigid ellipsoids "
                               f"(rasterized ghost cells), 128 listeners, collocated scheme",
                   "cells": N, "pml_cells": Npml, "bytes_per_step": bytes_step, "listeners_finite": finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "traffic_source": "not measured in this run: HBM bytes per launch from the committed counter "
                                       "summary profiles/step_kernel_256cube_traffic.json (rocprofv3 --pmc "
                                       "FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel of known size)",
                     "kernel": "step_kernel", "kernel_ms": round(kern_ms, 5),
                     "kernel_ms_source": "HIP events around the timed steps on the solver stream, per step",
                     "state_bytes": state_bytes,
                     "state_frac": round(state_bytes / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4) if kern_ms > 0 else 0.0},
    }
    if a.staggered_steps > 0:
        out["staggered"] = staggered(a.staggered_steps, n)
    if a.other_steps > 0:
        out["coupled"] = coupled(a.other_steps, n)
        out["coupled"]["ghost_ms_per_step"] = round(out["coupled"]["step_ms"] - kern_ms, 5)
        out.update(other_paths(a.other_steps, n))
    if a.batch > 0:
        out["batch"] = batched(a.batch)
    if a.serve > 0:
        out["serving"] = serving(a.serve, a.block, a.serve_per_step)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
