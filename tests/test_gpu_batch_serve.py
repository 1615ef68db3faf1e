"""Batched and served forms: every job's image is bit-identical to the same
job rendered alone (its RNG streams depend only on its own iteration, pixel,
depth and seed), and failure handling forced from the host (budget
exhaustion, cancellation, deadlines) leaves the other jobs untouched."""
import math
import time

import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import scenes
from luxion_amd import _ffi as F

pytestmark = pytest.mark.gpu


def orbit_camera(k, w, h, target=(0.0, 1.0, 0.0), radius=3.4, height=1.0, fovy=40.0):
    a = 0.35 * math.sin(1.7 * k)
    eye = (target[0] + radius * math.sin(a), height + 0.1 * math.cos(2.3 * k), target[2] + radius * math.cos(a))
    return scenes.look_at_camera(eye, target, (0, 1, 0), fovy, w, h)


@pytest.fixture(scope="module")
def scenes_pair(ctx):
    corn = lx.Scene(ctx, scenes.cornell(64, 64))
    ico_data = scenes.icosphere_hdri(64, 36, level=4, pieces=8, env_size=(128, 64), trace_depth=6)
    ico = lx.Scene(ctx, ico_data)
    return corn, ico


def _jobs(corn, ico):
    jobs = []
    spec = [(corn, 32, 32, 3, True, False, 0, 1), (corn, 48, 40, 1, True, False, 7, 5), (ico, 40, 24, 2, False, True, 0, 1),
            (corn, 64, 64, 4, False, False, 3, 2), (ico, 56, 32, 3, False, True, 11, 9), (corn, 16, 24, 5, True, False, 0, 1)]
    for k, (sc, w, h, spp, dmis, emis, seed, first) in enumerate(spec):
        cam = orbit_camera(k, w, h) if sc is corn else scenes.look_at_camera((0, 0.3, 3.0), (0, 0, 0), (0, 1, 0), 40, w, h)
        jobs.append((sc, cam, dict(spp=spp, dmis=dmis, emis=emis, seed=seed, first_iter=first)))
    return jobs


@pytest.fixture(params=[1, 0], ids=["overlapped", "sequential"])
def server_mode(ctx, request):
    """Servers started with overlapped rounds (the default) and without."""
    saved = ctx.tuning()
    ctx.set_tuning(overlap_rounds=request.param)
    yield request.param
    ctx.set_tuning(**saved)


def _alone(sc, cam, spp, dmis, emis, seed, first_iter):
    r = lx.Render(sc, dmis=dmis, emis=emis, seed=seed, camera=cam)
    r.iterate(first_iter, spp)
    return r.accum()[0]


def test_batch_equals_individual_renders(ctx, scenes_pair):
    corn, ico = scenes_pair
    spec = _jobs(corn, ico)
    jobs = [lx.make_job(sc, cam, **kw) for sc, cam, kw in spec]
    res = lx.batch_run(ctx, jobs)
    for (sc, cam, kw), r in zip(spec, res):
        assert r["status"] == 0 and r["spp"] == kw["spp"]
        want = _alone(sc, cam, kw["spp"], kw["dmis"], kw["emis"], kw["seed"], kw["first_iter"])
        assert np.array_equal(r["image"], want)


def test_batch_with_small_path_capacity_schedules_subsets(ctx, scenes_pair):
    corn, ico = scenes_pair
    spec = _jobs(corn, ico)
    jobs = [lx.make_job(sc, cam, **kw) for sc, cam, kw in spec]
    full = lx.batch_run(ctx, jobs)
    ctx.set_tuning(max_paths_per_round=4096)  # the 64x64 job alone fills a round
    try:
        part = lx.batch_run(ctx, jobs)
    finally:
        ctx.set_tuning(max_paths_per_round=0)
    for a, b in zip(full, part):
        assert np.array_equal(a["image"], b["image"])
    # jobs finishing in later rounds finish later
    assert max(r["finish_ms"] for r in part) >= min(r["finish_ms"] for r in part)


def test_server_result_same_alone_and_in_mix(ctx, scenes_pair, server_mode):
    corn, ico = scenes_pair
    spec = _jobs(corn, ico)
    srv = lx.Server(ctx, max_jobs=8, max_depth=8)
    try:
        a = spec[4]
        ja = srv.submit(lx.make_job(a[0], a[1], **a[2]))
        assert srv.poll(ja, 30000)
        alone = srv.result(ja)["image"].copy()
        ids = [srv.submit(lx.make_job(sc, cam, **kw)) for sc, cam, kw in spec]
        for i in ids:
            assert srv.poll(i, 60000)
        for (sc, cam, kw), i in zip(spec, ids):
            r = srv.result(i)
            assert r["status"] == 0 and r["spp"] == kw["spp"]
            want = _alone(sc, cam, kw["spp"], kw["dmis"], kw["emis"], kw["seed"], kw["first_iter"])
            assert np.array_equal(r["image"], want)
        assert np.array_equal(srv.result(ids[4])["image"], alone)
        st = srv.stats()
        assert st["jobs_done"] == len(spec) + 1
        assert st["peak_jobs_in_flight"] > 1  # jobs overlapped on the device
        assert st["p99_ms"] >= st["p50_ms"] > 0
    finally:
        srv.stop()


def test_server_failure_handling_isolated(ctx, scenes_pair, server_mode):
    corn, ico = scenes_pair
    srv = lx.Server(ctx, max_jobs=8, max_depth=8)
    try:
        cam = orbit_camera(3, 32, 32)
        srv.inject_budget_failure(1)
        rejected = srv.submit(lx.make_job(corn, cam, spp=2, dmis=True))
        long_job = srv.submit(lx.make_job(corn, orbit_camera(1, 64, 64), spp=100000, dmis=True))
        deadline = srv.submit(lx.make_job(corn, orbit_camera(2, 64, 64), spp=100000, dmis=True, deadline_ms=50.0))
        ok = srv.submit(lx.make_job(corn, cam, spp=3, dmis=True, seed=5))
        assert srv.poll(rejected, 30000) and srv.result(rejected)["status"] == F.LX_OK + 6  # LX_ERR_BUDGET
        assert srv.poll(ok, 30000) and srv.result(ok)["status"] == 0
        assert np.array_equal(srv.result(ok)["image"], _alone(corn, cam, 3, True, False, 5, 1))
        assert srv.poll(deadline, 60000) and srv.result(deadline)["status"] == 8  # LX_ERR_TIMEOUT
        srv.cancel(long_job)
        assert srv.poll(long_job, 30000) and srv.result(long_job)["status"] == 7  # LX_ERR_CANCELLED
        st = srv.stats()
        assert st["jobs_rejected"] == 1 and st["jobs_failed"] == 2 and st["jobs_done"] == 1
    finally:
        srv.stop()


def test_workload_runners_small(ctx, scenes_pair):
    from luxion_amd import workloads
    corn, ico = scenes_pair
    specs = [(ico, dict(emis=True), (0.0, 0.0, 0.0), 3.0), (corn, dict(dmis=True), (0.0, 1.0, 0.0), 3.4)]
    spec = workloads.mixed_jobs(specs, 6, seed=5, side=(16, 48), spp=(1, 4))
    b, res = workloads.run_batched(ctx, spec)
    assert b["images_bitexact"] and b["speedup_vs_per_job"] > 0
    s = workloads.run_served(ctx, workloads.mixed_jobs(specs, 6, seed=6, side=(16, 48), spp=(1, 4)), 500.0,
                             max_depth=8)
    assert s["p99_ms"] >= s["p50_ms"] > 0 and s["jobs"] == 6


def test_server_many_small_jobs_overlapped_equal_alone(ctx, scenes_pair, server_mode):
    """More jobs than slots, with pool ranges freed and reused while rounds
    are in flight: every image still equals the job rendered alone."""
    corn, ico = scenes_pair
    srv = lx.Server(ctx, max_jobs=3, max_depth=8)
    try:
        spec = []
        for k in range(10):
            w, h = 16 + 8 * (k % 4), 16 + 4 * (k % 3)
            spec.append((corn, orbit_camera(k, w, h), dict(spp=1 + k % 4, dmis=True, seed=k, first_iter=1 + k)))
        ids = [srv.submit(lx.make_job(sc, cam, **kw)) for sc, cam, kw in spec]
        for i in ids:
            assert srv.poll(i, 60000)
        for (sc, cam, kw), i in zip(spec, ids):
            r = srv.result(i)
            assert r["status"] == 0 and r["spp"] == kw["spp"]
            assert np.array_equal(r["image"], _alone(sc, cam, kw["spp"], True, False, kw["seed"], kw["first_iter"]))
        assert srv.stats()["jobs_done"] == 10
    finally:
        srv.stop()


def test_server_jobs_beyond_path_capacity_take_turns(ctx, scenes_pair, server_mode):
    """Path capacity for one 64x64 job per round, several jobs in flight: the
    jobs take turns, and each image still equals the job rendered alone."""
    corn, ico = scenes_pair
    srv = lx.Server(ctx, max_jobs=4, max_depth=8, max_paths=4096)
    try:
        spec = [(corn, orbit_camera(k, 64 if k % 2 == 0 else 32, 64 if k % 2 == 0 else 32),
                 dict(spp=2 + k % 3, dmis=True, seed=3 * k)) for k in range(6)]
        ids = [srv.submit(lx.make_job(sc, cam, **kw)) for sc, cam, kw in spec]
        for i in ids:
            assert srv.poll(i, 60000)
        for (sc, cam, kw), i in zip(spec, ids):
            r = srv.result(i)
            assert r["status"] == 0 and r["spp"] == kw["spp"]
            assert np.array_equal(r["image"], _alone(sc, cam, kw["spp"], True, False, kw["seed"], 1))
    finally:
        srv.stop()
