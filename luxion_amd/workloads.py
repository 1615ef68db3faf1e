"""Job-mix workloads of the batched and served forms.

mixed_jobs: jobs on one or more resident scenes, square sides drawn from
U{side_lo..side_hi}, spp from U{spp_lo..spp_hi}, and a per-job camera on an
orbit whose angle and height come from the job's own seed.

run_batched: one lx_batch_run over the mix versus the same jobs rendered one
after another through their own lx_render (graph-replayed rounds), with the
images of both compared bit for bit (so the two reach the same RMSE against
any reference).

run_per_job: the same jobs rendered one after another, each through its
own lx_render (the plain per-job launches the batched form is measured
against).

run_served: Poisson arrivals at rate lam into an lx_server; latency is
submit -> result ready, per job.
"""
import math
import time

import numpy as np

from . import Render, Server, batch_run, make_job, scenes


def orbit_camera(seed, width, height, target=(0.0, 0.0, 0.0), radius=3.0, fovy=40.0):
    r = np.random.default_rng(0x5EED0000 + seed)
    ang = float(r.uniform(0.0, 2.0 * math.pi))
    h = float(r.uniform(0.1, 1.2))
    eye = (target[0] + radius * math.sin(ang), target[1] + h, target[2] + radius * math.cos(ang))
    return scenes.look_at_camera(eye, target, (0, 1, 0), fovy, width, height)


def mixed_jobs(scene_specs, n_jobs, seed=3, side=(64, 512), spp=(1, 64)):
    """scene_specs: list of (Scene, dict of job flags, camera target, orbit
    radius).  Returns [(scene, camera, kwargs)] with kwargs for make_job."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_jobs):
        sc, flags, target, radius = scene_specs[k % len(scene_specs)]
        s = int(rng.integers(side[0], side[1] + 1))
        n = int(rng.integers(spp[0], spp[1] + 1))
        cam = orbit_camera(seed * 100003 + k, s, s, target, radius)
        out.append((sc, cam, dict(spp=n, seed=k + 1, first_iter=1, **flags)))
    return out


def path_samples(spec):
    return sum(cam["width"] * cam["height"] * kw["spp"] for _, cam, kw in spec)


def run_batched(ctx, spec, per_job=True):
    """Returns a dict: batched wall time and spp/s, per-job wall time and
    spp/s, their ratio, and whether every image matched bit for bit."""
    jobs = [make_job(sc, cam, **kw) for sc, cam, kw in spec]
    t0 = time.perf_counter()
    res = batch_run(ctx, jobs)
    t_batch = time.perf_counter() - t0
    dev_ms = max(r["finish_ms"] for r in res) - min(r["start_ms"] for r in res)
    assert all(r["status"] == 0 for r in res)
    ps = path_samples(spec)
    rays = sum(r["rays"] for r in res)
    out = {"jobs": len(spec), "path_samples": ps, "batched_s": round(t_batch, 4),
           "batched_device_s": round(dev_ms / 1e3, 4), "batched_spp_per_s": round(ps / t_batch, 1),
           "batched_mrays_per_s": round(rays / t_batch / 1e6, 1), "rays_per_path_sample": round(rays / ps, 3)}
    if not per_job:
        return out, res
    # per-job launches: each job's own render, buffers and graph prepared
    # beforehand (a warm-up sample, then reset), timed from first launch to
    # its last sample being ready
    renders = []
    for sc, cam, kw in spec:
        r = Render(sc, camera=cam, seed=kw["seed"], dmis=kw.get("dmis", False), emis=kw.get("emis", False))
        r.iterate(1, 1)
        r.sync()
        r.reset()
        renders.append(r)
    same = True
    t_jobs = 0.0
    for r, (sc, cam, kw), rb in zip(renders, spec, res):
        t0 = time.perf_counter()
        r.iterate(kw["first_iter"], kw["spp"])
        r.sync()
        t_jobs += time.perf_counter() - t0
        same &= bool(np.array_equal(r.accum()[0], rb["image"]))
        r.close()
    out.update({"per_job_s": round(t_jobs, 4), "per_job_spp_per_s": round(ps / t_jobs, 1),
                "speedup_vs_per_job": round(t_jobs / t_batch, 3), "images_bitexact": same})
    return out, res


def run_per_job(spec):
    """Plain per-job launches: every job through its own lx_render, under
    the context's current tuning, buffers and graph prepared beforehand (a
    warm-up sample, then reset), timed from each job's first launch to its
    last sample being ready.  Returns the summed time and the images."""
    t_jobs = 0.0
    images = []
    for sc, cam, kw in spec:
        r = Render(sc, camera=cam, seed=kw["seed"], dmis=kw.get("dmis", False), emis=kw.get("emis", False))
        r.iterate(1, 1)
        r.sync()
        r.reset()
        t0 = time.perf_counter()
        r.iterate(kw["first_iter"], kw["spp"])
        r.sync()
        t_jobs += time.perf_counter() - t0
        images.append(r.accum()[0])
        r.close()
    ps = path_samples(spec)
    return {"per_job_s": round(t_jobs, 4), "per_job_spp_per_s": round(ps / t_jobs, 1), "images": images}


def run_served(ctx, spec, lam, seed=11, max_jobs=64, max_depth=12, srv=None):
    """Submit spec's jobs with exponential inter-arrival gaps of mean 1/lam
    seconds; wait for all; return latency percentiles and throughput."""
    own = srv is None
    if own:
        srv = Server(ctx, max_jobs=max_jobs, max_depth=max_depth)
    try:
        st0 = srv.stats()
        gaps = np.random.default_rng(seed).exponential(1.0 / lam, len(spec))
        ids = []
        t_start = time.perf_counter()
        t_next = t_start
        for (sc, cam, kw), g in zip(spec, gaps):
            t_next += g
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(make_job(sc, cam, **kw)))
        for i in ids:
            assert srv.poll(i, 600000.0)
        t_end = time.perf_counter()
        lat = np.array([srv.result(i)["finish_ms"] - srv.result(i)["submit_ms"] for i in ids])
        assert all(srv.result(i)["status"] == 0 for i in ids)
        st1 = srv.stats()
        rays = st1["rays"] - st0["rays"]
        ps = path_samples(spec)
        wall = t_end - t_start
        return {"lambda_per_s": lam, "jobs": len(spec), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2), "mean_ms": round(float(lat.mean()), 2),
                "wall_s": round(wall, 3), "spp_per_s": round(ps / wall, 1), "mrays_per_s": round(rays / wall / 1e6, 1),
                "rounds": int(st1["rounds"] - st0["rounds"]),
                "peak_jobs_in_flight": int(st1["peak_jobs_in_flight"])}
    finally:
        if own:
            srv.stop()
