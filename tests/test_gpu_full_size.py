"""The workloads at their full sizes on the device, against the FP64 oracle's
committed series (tests/golden/full, written by make_golden_full.py):

* the 128^3 sphere (2000 steps, 64 listeners on a 0.2 m shell) and the 256^3
  scene of five ellipsoids (4000 steps, 128 listeners), each with rasterized
  and with fully-coupled ghost cells: listener relative L2 error <= 1e-4 over
  the whole series;
* the 64-scene batch (edges 32-128, 1000-3000 steps): every scene equals the
  single-scene path bit for bit, and four scenes match the oracle to 1e-4;
* a Poisson stream of requests (edges 48-192, 0.01-0.05 s of sound, 1-3
  ellipsoids, 32 listeners) through the server: every result equals the same
  request run alone, bit for bit.
The scene of each fixture is rebuilt from its builder arguments; its
listener positions and object centres must equal the stored ones."""
import json
import os
import time

import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.solver import Batch, Server, Solver, rel_l2

pytestmark = pytest.mark.gpu

FULL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full")
TOL = 1e-4


def load(name):
    z = np.load(os.path.join(FULL, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    built = getattr(scenes, meta["builder"])(**meta["kwargs"])
    if meta["index"] is None:
        sb, steps = built
    else:
        sb, steps = built[0][meta["index"]], built[1][meta["index"]]
    assert steps == meta["steps"]
    assert np.array_equal(np.asarray(sb.listeners_np, dtype=np.float64).reshape(-1, 3), z["listeners"])
    ctr = np.asarray([o["center"] for o in sb.objects_list], dtype=np.float64).reshape(-1, 3)
    assert np.array_equal(ctr, z["centers"])
    return sb, steps, z["series"].astype(np.float64)


@pytest.mark.parametrize("name", ["sphere128_raster", "sphere128_coupled", "ellipsoids256_raster",
                                  "ellipsoids256_coupled"])
def test_full_size_scene_matches_oracle(hip_lib, name):
    sb, steps, ref = load(name)
    print(f"full-size case {name}: divisions={sb.divisions} steps={steps} listeners={sb.n_listeners}", flush=True)
    t = time.time()
    with Solver(sb, steps_per_graph=32) as s:
        s.step(steps)
        got = s.series()
    err = rel_l2(got, ref)
    print(f"{name}: listener rel L2 {err:.3e} ({time.time() - t:.1f} s)", flush=True)
    assert np.all(np.isfinite(got))
    assert err <= TOL


def test_full_batch_of_64_scenes(hip_lib):
    sbs, steps = scenes.config4(seed=4)
    print(f"64-scene batch: {sum(int(np.prod(sb.divisions)) * n for sb, n in zip(sbs, steps)) / 1e9:.1f} G cell-updates",
          flush=True)
    with Batch(sbs, steps, steps_per_graph=32) as b:
        b.run()
        got = [b.series(i) for i in range(len(sbs))]
    for i in range(4):
        sb, n, ref = load(f"batch64_scene{i}")
        err = rel_l2(got[i], ref)
        print(f"batch scene {i} {sb.divisions} {n} steps: rel L2 vs oracle {err:.3e}", flush=True)
        assert err <= TOL
    for i, sb in enumerate(sbs):
        with Solver(sb) as s:
            s.step(steps[i])
            alone = s.series()
        assert np.array_equal(got[i], alone), i


def test_full_serving_stream(hip_lib):
    stream = scenes.config5_stream(seed=5, n_requests=24)
    print(f"stream: {len(stream)} requests, "
          f"{sum(int(np.prod(sc.divisions)) * n for _, sc, n in stream) / 1e9:.0f} G cell-updates", flush=True)
    with Server(steps_per_block=64) as srv:
        t0 = time.time()
        ids = []
        for t_arr, sc, n in stream:
            while time.time() - t0 < t_arr:
                time.sleep(0.0005)
            ids.append(srv.submit(sc, n))
        served = []
        for rid in ids:
            assert srv.wait(rid, 600000) == 0
            served.append(srv.take(rid)[1])
    for (t_arr, sc, n), got in zip(stream, served):
        with Solver(sc) as s:
            s.step(n)
            alone = s.series()
        assert got.shape == alone.shape
        assert np.array_equal(got, alone)
