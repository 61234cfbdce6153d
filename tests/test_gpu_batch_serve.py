"""Batched and served execution: every scene's listener series must equal
the single-scene path bit for bit (the per-scene arithmetic does not depend
on what else shares the launches), finished scenes stop at their own step
counts, and the server admits by HBM budget, cancels, and rejects requests
larger than the budget while the others complete unaffected."""
import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.solver import Batch, Server, Solver, rel_l2, scene_device_bytes
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _small_batch(seed=41, n=6):
    sbs, steps = scenes.config4(seed=seed, n_scenes=n, edge_lo=24, edge_hi=40, steps_lo=60, steps_hi=150)
    return sbs, steps


def _alone(sb, steps):
    with Solver(sb) as s:
        s.step(steps)
        return s.series()


def test_batch_matches_single_scene_path(hip_lib):
    sbs, steps = _small_batch()
    print("batch:", [sb.divisions for sb in sbs], steps, flush=True)
    with Batch(sbs, steps, steps_per_graph=16) as b:
        b.run()
        got = [b.series(i) for i in range(len(sbs))]
    for i, sb in enumerate(sbs):
        ref = _alone(sb, steps[i])
        assert got[i].shape == ref.shape
        assert np.array_equal(got[i], ref), i


def test_batch_matches_oracle(hip_lib):
    sbs, steps = _small_batch(seed=42, n=3)
    with Batch(sbs, steps, steps_per_graph=8) as b:
        b.run()
        for i, sb in enumerate(sbs):
            ref = Oracle(sb, reclassify_each_step=False).step(steps[i])
            assert rel_l2(b.series(i), ref) <= 1e-4


def test_server_alone_and_in_mix_identical(hip_lib):
    sbs, steps = _small_batch(seed=43, n=5)
    with Server(steps_per_block=16) as srv:
        ids = [srv.submit(sb, st) for sb, st in zip(sbs, steps)]
        for rid in ids:
            assert srv.wait(rid, 120000) == 0
        mixed = [srv.take(rid)[1] for rid in ids]
    for i, sb in enumerate(sbs):
        with Server(steps_per_block=16) as srv:
            rid = srv.submit(sb, steps[i])
            assert srv.wait(rid, 120000) == 0
            alone = srv.take(rid)[1]
        assert np.array_equal(alone, mixed[i]), i
        assert np.array_equal(alone, _alone(sb, steps[i])), i


def test_server_budget_cancel_and_reject(hip_lib):
    sbs, steps = _small_batch(seed=44, n=4)
    need = [scene_device_bytes(sb) for sb in sbs]
    budget = int(max(need) * 1.5) + (8 << 20)
    with Server(hbm_budget_bytes=budget, steps_per_block=16) as srv:
        ids = [srv.submit(sb, st) for sb, st in zip(sbs, steps)]
        srv.cancel(ids[1])
        big, _ = scenes.config1(seed=3, n=96)
        with pytest.raises(RuntimeError):
            srv.submit(big, 100)  # over the whole budget: rejected at submit
        for i, rid in enumerate(ids):
            st = srv.wait(rid, 120000)
            if i == 1:
                assert st in (0, 9)  # finished before the cancel landed, or cancelled
            else:
                assert st == 0
        stats = srv.stats()
        assert stats["peak_resident_bytes"] <= budget
        for i, rid in enumerate(ids):
            if i == 1:
                continue
            st, ser, _ = srv.take(rid)
            assert np.array_equal(ser, _alone(sbs[i], steps[i]))


def test_engquist_scenes_in_batch_and_server(hip_lib):
    """Scenes with the Engquist boundary (no PML) share launches with PML
    scenes in a batch and in the server, and give the single-scene series."""
    sbs, steps = _small_batch(seed=44, n=3)
    sbs = sbs + [scenes.engquist_free_field(seed=61, div=(30, 34, 38))[0], scenes.engquist_sphere(seed=62, n=40)[0]]
    steps = steps + [120, 90]
    print("batch:", [sb.divisions for sb in sbs], steps, flush=True)
    with Batch(sbs, steps, steps_per_graph=16) as b:
        b.run()
        got = [b.series(i) for i in range(len(sbs))]
    with Server(steps_per_block=16) as srv:
        ids = [srv.submit(sb, st) for sb, st in zip(sbs, steps)]
        for rid in ids:
            assert srv.wait(rid, 120000) == 0
        served = [srv.take(rid)[1] for rid in ids]
    for i, sb in enumerate(sbs):
        ref = _alone(sb, steps[i])
        assert np.array_equal(got[i], ref), i
        assert np.array_equal(served[i], ref), i


def test_staggered_and_speaker_scenes_in_batch_and_server(hip_lib):
    """Staggered-scheme scenes (their float4 rows in the fused launch) and a
    speaker scene share launches with collocated scenes in a batch and in the
    server, and give the single-scene series bit for bit."""
    sbs, steps = _small_batch(seed=45, n=2)
    stag_a, _ = scenes.staggered_free_field(seed=63, div=(36, 40, 44), h=0.005, pml=6)
    stag_b, _ = scenes.staggered_free_field(seed=64, div=(48, 32, 40), h=0.005, pml=8)
    spk, _ = scenes.speaker_scene(seed=65, n=40)
    sbs = sbs + [stag_a, stag_b, spk]
    steps = steps + [110, 140, 100]
    print("batch:", [sb.divisions for sb in sbs], steps, flush=True)
    with Batch(sbs, steps, steps_per_graph=16) as b:
        b.run()
        got = [b.series(i) for i in range(len(sbs))]
    with Server(steps_per_block=16) as srv:
        ids = [srv.submit(sb, st) for sb, st in zip(sbs, steps)]
        for rid in ids:
            assert srv.wait(rid, 120000) == 0
        served = [srv.take(rid)[1] for rid in ids]
    for i, sb in enumerate(sbs):
        ref = _alone(sb, steps[i])
        assert np.abs(ref).max() > 0, i
        assert np.array_equal(got[i], ref), i
        assert np.array_equal(served[i], ref), i


def _blow_up_scene(seed=9, n=32):
    """A scene whose time step is twice the FDTD stability limit
    (c dt / h = 2/sqrt(3)): its highest spatial mode grows about 14-fold per
    step, so the field overflows to infinity and NaN within ~100 steps."""
    rng = np.random.default_rng(seed)
    s = scenes.base_settings((n, n, n), 0.005, 5, dt=scenes.cfl_dt(0.005, cfl=2.0))
    lo, hi = scenes.interior_bounds(s, 2)
    src = scenes.gaussian_pressure_source(rng.uniform(lo, hi), 2e-4)
    return scenes.SceneBuffers(s, pressure_sources=[src], listeners=rng.uniform(lo, hi, size=(4, 3)))


def test_server_evicts_a_blown_up_request(hip_lib):
    """A request that blows up is evicted with FDTD_ERR_NONFINITE at the next
    periodic check, long before its last step, while the requests sharing its
    launches finish with the series they give alone."""
    sbs, steps = _small_batch(seed=45, n=4)
    bad_steps, block = 20000, 16
    with Server(steps_per_block=block) as srv:
        bad = srv.submit(_blow_up_scene(), bad_steps)
        ids = [srv.submit(sb, st) for sb, st in zip(sbs, steps)]
        assert srv.wait(bad, 120000) == 7           # FDTD_ERR_NONFINITE
        for rid in ids:
            assert srv.wait(rid, 120000) == 0
        stats = srv.stats()
        served = [srv.take(rid)[1] for rid in ids]
        st, ser, _ = srv.take(bad)
    assert st == 7 and ser is None
    assert stats["failed"] == 1 and stats["completed"] == len(ids)
    # evicted within a few checks (16 blocks apart), not after 20000 steps
    assert stats["blocks_run"] < 100, stats
    for i, sb in enumerate(sbs):
        assert np.array_equal(served[i], _alone(sb, steps[i])), i


def test_server_take_while_admitting(hip_lib):
    """Clients that take each result the moment it is ready, while another
    thread keeps submitting (so the worker admits, finishes and rebuilds
    concurrently), get every request's own series."""
    import threading
    import time
    sbs, steps = _small_batch(seed=46, n=12)
    got, ids = {}, {}
    with Server(steps_per_block=8) as srv:
        def submitter():
            for i, (sb, st) in enumerate(zip(sbs, steps)):
                ids[i] = srv.submit(sb, st)
        t = threading.Thread(target=submitter)
        t.start()
        pending = set(range(len(sbs)))
        deadline = time.monotonic() + 120
        while pending:
            assert time.monotonic() < deadline, pending
            time.sleep(0.0005)
            for i in list(pending):
                if i in ids and srv.poll(ids[i]) != 8:    # not FDTD_ERR_BUSY
                    st, ser, _ = srv.take(ids[i])
                    assert st == 0, (i, st)
                    got[i] = ser
                    pending.discard(i)
        t.join()
    for i, sb in enumerate(sbs):
        assert np.array_equal(got[i], _alone(sb, steps[i])), i


def test_batch_create_failure_frees_uploaded_scenes(hip_lib):
    """A batch whose later scene is refused frees what it already built (the
    first scene's arena, the program buffers, its stream): five refused
    creations would otherwise keep five 128^3 arenas (5 x 54.8 MB); the
    runtime itself keeps about 12 MB per stream created and destroyed."""
    import ctypes as C
    from acoustics_amd import capi
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    sb, _ = scenes.config2(seed=3)
    need = scene_device_bytes(sb)
    print("case: refused batch, arena bytes", need, flush=True)
    with Batch([sb], [8], tile_config=0) as b:   # warm the allocator and the module
        b.run()
    before = free_bytes()
    for _ in range(5):
        with pytest.raises(capi.FDTDError):
            Batch([sb, sb], [100, -1], tile_config=0)
    after = free_bytes()
    print("free before", before, "after", after, flush=True)
    assert before - after < 2 * need
