"""Benchmark: Mrays/s + spp/s of the wavefront path tracer at 1920x1080 on one
MI355X, on the single-mesh + HDRI workload (1,310,720-triangle displaced
icosphere in 64 primitives + ground plane, 2048x1024 HDRI with a 0.5-degree
sun of radiance 1e4, depth 12, environment importance sampling + MIS).

A step is one 1-spp frame: camera rays, then the bounces (by
default all in the persistent tail kernel: radiance + environment shadow
rays, uber-shader, MIS, Russian roulette, accumulation).  A round renders
twenty consecutive frames as twenty copies of the job (tuning round_frames),
so the last long paths of a frame share the GPU with the other frames' paths
and the launch's final drain is paid once per twenty frames; the
accumulated image is the same bits as one frame after another.  Inputs are resident in HBM before the timed
region.  Per-kernel times for the roofline come from a separate pass with
HIP events around every kernel.  Output: one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
NODE_BYTES, TRI_BYTES, INST_BYTES = 128, 48, 64
RAY_BYTES, HIT_BYTES, OCC_BYTES = 32, 16, 4
# per path-bounce shading: path state read + write, three hit records, three
# index triples, normals and UVs, the material, the next ray and two shadow
# rays, one queue word (texels not counted)
SHADE_BYTES = 2 * 48 + 3 * 16 + 12 + 36 + 24 + 64 + 32 + 2 * 32 + 4
# launch shapes for this workload, applied unless --autotune 1: the fastest
# measured for the 1080p frame (tail kernel for every bounce, 256-lane
# blocks, 16 waves per CU, no queue sort, 16x16 camera squares in bands of 8
# square rows, twenty frames per round, one round at a time); fixed so that
# runs, and the profiles of the same command, time the same kernels
WORKLOAD_TUNING = dict(wavefront_bounces=0, block_size=256, waves_per_cu=16, queue_chunk=32, tail_min_waves=4,
                       bvh_width=4, sort_mode=0, pixel_tile=16, tile_band=8, overlap_rounds=0, round_frames=20)
# the batched and served legs (and the per-job renders they are compared
# with) run one sample per round, overlapped, with the ray-cell sort: the
# library's round shape for plain per-job launches
FORMS_TUNING = dict(WORKLOAD_TUNING, sort_mode=1, overlap_rounds=1, round_frames=1)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (a line 'traffic_bytes_per_launch:
    N' in profiles/<kernel>_cfg1_1080p_emis.md), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cfg1_1080p_emis.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def measured_issue(kernel):
    """VALU busy and lane-utilisation fractions of `kernel` on this workload,
    from the same committed summary (lines 'VALU busy = ... = X %' and 'VALU
    lane utilisation = ... = Y %'), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cfg1_1080p_emis.md")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                s = line.strip()
                for key, tag in (("valu_busy", "VALU busy ="), ("lane_utilisation", "VALU lane utilisation =")):
                    if tag in s and s.endswith("%") and key not in out:
                        out[key] = round(float(s.rsplit("=", 1)[1].strip().rstrip("%").strip()) / 100.0, 3)
    except (OSError, ValueError):
        return None
    return out or None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--accuracy", type=int, default=1,
                    help="1: compare a 16-spp render of the workload with the fp32 oracle (bits) and the FP64 "
                         "oracle (RMSE) on 4,096 seeded pixels")
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--batched", type=int, default=256,
                    help="jobs in the batched-form mix on the same scene (0 = skip)")
    ap.add_argument("--served", choices=["off", "light", "full"], default="light",
                    help="served-form Poisson sweep: light = the HDRI scene + Cornell box resident; full = the "
                         "HDRI scene + instanced interior + 30M-triangle grid resident")
    ap.add_argument("--served-jobs", type=int, default=48, help="jobs per arrival rate")
    ap.add_argument("--wavefront-bounces", type=int, default=None,
                    help="bounces launched one kernel each before the tail kernel (default: the context's)")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: run the library's autotuner on this workload (8-spp job) before timing instead "
                         "of applying WORKLOAD_TUNING (its picks vary between shapes within the noise)")
    ap.add_argument("--set", action="append", default=[], metavar="KNOB=VALUE",
                    help="set a tuning knob (lx_tuning field) after autotuning, e.g. --set sort_mode=1")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run the plain 16-byte copy kernel over this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import ctypes
        import luxion_amd as lx
        ctx = lx.Context(0)
        g = ctypes.c_double()
        lx.check(ctx.lib.lx_debug_copy_bandwidth(ctx.h, args.copy_calibration << 20, 4, ctypes.byref(g)))
        print(json.dumps({"copy_mib": args.copy_calibration, "dispatches": 5, "gbps": round(g.value, 1)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported: each GPU "
                                   "would render its own share (spp or jobs) and reduce the accumulation"}))
        sys.exit(2)

    import numpy as np
    import luxion_amd as lx
    from luxion_amd import scenes

    W, H = 1920, 1080
    t0 = time.time()
    data = scenes.icosphere_hdri(W, H, level=args.level)
    log(f"scene generated: {data.num_tris} triangles in {time.time() - t0:.1f} s")
    ctx = lx.Context(0)
    t0 = time.time()
    sc = lx.Scene(ctx, data)
    info = sc.info()
    log(f"scene on device: {info['device_bytes'] / 2**20:.1f} MiB, BVH {info['bvh_nodes']} nodes, "
        f"build {info['build_ms']:.0f} ms")
    depth = data.trace_depth

    if args.autotune:
        t0 = time.time()
        best = ctx.autotune([lx.make_job(sc, data.camera, 8, emis=True)])
        log(f"autotuned in {time.time() - t0:.1f} s: {best}")
    else:
        ctx.set_tuning(**WORKLOAD_TUNING)
    if args.wavefront_bounces is not None:
        ctx.set_tuning(wavefront_bounces=args.wavefront_bounces)
    for kv in args.set:
        k, v = kv.split("=", 1)
        ctx.set_tuning(**{k: int(v)})
    tuning = ctx.tuning()
    r = lx.Render(sc, emis=True)
    r.iterate(1, args.warmup)
    r.sync()
    log(f"warmup done ({args.warmup} steps)")

    # timed region: the frames as a caller renders them (rounds of
    # round_frames frames, overlapped when overlap_rounds says so),
    # synchronised on both sides
    first = args.warmup + 1
    r.reset()
    t0 = time.perf_counter()
    r.iterate(first, args.steps)
    r.sync()
    t_timed = time.perf_counter() - t0
    st_main = r.stats()
    rays = st_main["radiance_rays"] + st_main["dl_shadow_rays"] + st_main["env_shadow_rays"]
    log(f"timed region: {args.steps} steps in {t_timed * 1e3:.1f} ms, {rays} rays "
        f"(overlap_rounds {tuning['overlap_rounds']}, round_frames {tuning['round_frames']})")

    # one round after another, replayed from a captured hipGraph
    ctx.set_tuning(overlap_rounds=0, round_frames=1)
    r.reset()
    r.iterate(first, 1)
    r.sync()
    r.reset()
    t0 = time.perf_counter()
    r.iterate(first, args.steps)
    r.sync()
    t_graph = time.perf_counter() - t0
    st_graph = r.stats()
    log(f"sequential rounds (graph replay): {args.steps} steps in {t_graph * 1e3:.1f} ms")

    # kernel timing pass: the timed region's rounds again, with HIP events
    # before and after each kernel on the stream it runs on (per-kernel
    # device time for the roofline)
    ctx.set_tuning(overlap_rounds=tuning["overlap_rounds"], round_frames=tuning["round_frames"])
    r.reset()
    r.set_timing(events=True)
    t0 = time.perf_counter()
    r.iterate(first, args.steps)
    r.sync()
    t_events = time.perf_counter() - t0
    st = r.stats()
    r.close()  # frees the wide rounds' path buffers before the other legs
    assert st["radiance_rays"] == st_main["radiance_rays"], "kernel timing pass traced a different ray batch"
    log(f"kernel timing pass: {args.steps} steps in {t_events * 1e3:.1f} ms")

    # counting replay of the same frames: node / triangle / instance visits
    rc = lx.Render(sc, emis=True)
    rc.set_timing(events=False, count_work=True)
    rc.iterate(first, args.steps)
    sc_ = rc.stats()
    assert sc_["radiance_rays"] == st["radiance_rays"], "counting replay traced a different ray batch"
    rc.close()

    def trace_bytes(n_closest, n_any, c):
        return (n_closest * (RAY_BYTES + HIT_BYTES) + n_any * (RAY_BYTES + OCC_BYTES) +
                INST_BYTES * (c["closest_insts"] + c["any_insts"]) +
                NODE_BYTES * (c["closest_nodes"] + c["any_nodes"]) + TRI_BYTES * (c["closest_tris"] + c["any_tris"]))

    n_closest = sc_["radiance_rays"] + sc_["dl_shadow_rays"]
    n_any = sc_["env_shadow_rays"]
    tail = {k[5:]: v for k, v in sc_.items() if k.startswith("tail_") and k[5:] in
            ("closest_nodes", "closest_tris", "closest_insts", "any_nodes", "any_tris", "any_insts")}
    wave = {k: sc_[k] - tail[k] for k in tail}
    tail_bytes = trace_bytes(sc_["tail_closest_rays"], sc_["tail_any_rays"], tail)
    wave_bytes = trace_bytes(n_closest - sc_["tail_closest_rays"], n_any - sc_["tail_any_rays"], wave)
    per_ray = {"nodes_closest": sc_["closest_nodes"] / max(1, n_closest),
               "tris_closest": sc_["closest_tris"] / max(1, n_closest),
               "nodes_any": sc_["any_nodes"] / max(1, n_any), "tris_any": sc_["any_tris"] / max(1, n_any)}
    # the dominant kernel: the tail kernel (trace + shade of every bounce it
    # runs) or, when every bounce is its own launch, the trace kernel
    if st["tail_ms"] >= st["trace_ms"]:
        kname, kms, klaunch = "k_finish", st["tail_ms"], st["tail_launches"]
        shaded = sc_["tail_bounces"]
        kbytes = tail_bytes + SHADE_BYTES * shaded
    else:
        kname, kms, klaunch = "k_trace", st["trace_ms"], st["trace_launches"]
        kbytes = wave_bytes
    launch_ms = kms / max(1, klaunch)
    bytes_per_launch = kbytes / max(1, klaunch)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = measured_traffic(kname)

    value = rays / t_timed / 1e6
    value_graph = (st_graph["radiance_rays"] + st_graph["dl_shadow_rays"] + st_graph["env_shadow_rays"]) \
        / t_graph / 1e6
    value_events = rays / t_events / 1e6
    spp_s = W * H * args.steps / t_timed

    cpu = None
    o = None
    if args.cpu_baseline:
        from oracle import lxo
        threads = min(16, os.cpu_count() or 1)
        o = lxo.OracleScene(data)
        log(f"oracle BVH built ({lxo.lib().lxo_scene_bvh_build_ms(o.h):.0f} ms); timing CPU baseline")
        rng = np.random.default_rng(0)
        n_pix = W * H
        tot_rays, tot_ms, iters = 0, 0.0, 0
        while tot_ms < args.cpu_seconds * 1e3 and iters < 8:
            _, cst = o.render(first + iters, 1, emis=True, threads=threads)
            tot_rays += cst["radiance_rays"] + cst["dl_shadow_rays"] + cst["env_shadow_rays"]
            tot_ms += cst["wall_ms"]
            iters += 1
            log(f"cpu baseline: {iters} frame(s), {tot_ms / 1e3:.1f} s")
        cpu = {"value": tot_rays / (tot_ms * 1e-3) / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{iters} full 1920x1080 frame(s) at 1 spp of the same workload (iterations "
                         f"{first}..{first + iters - 1}), oracle/lxo_oracle.cpp restatement with its "
                         f"reference-style binary SAH BVH, {threads} threads; BVH build excluded"}
        # one thread on 200,000 random pixels of the same frame
        pix = np.sort(rng.choice(n_pix, 200000, replace=False)).astype(np.uint32)
        _, c1 = o.render(first, 1, pixels=pix, emis=True, threads=1)
        cpu["single_thread_value"] = (c1["radiance_rays"] + c1["dl_shadow_rays"] + c1["env_shadow_rays"]) \
            / (c1["wall_ms"] * 1e-3) / 1e6
        cpu["single_thread_sample"] = "200,000 random pixels of one frame (iteration %d), 1 thread" % first
        # the CPU-runnable Cornell workload, GPU and CPU on the same job
        cd = scenes.cornell(256, 256)
        oc = lxo.OracleScene(cd)
        _, cc = oc.render(1, 16, dmis=True, threads=threads)
        csc = lx.Scene(ctx, cd)
        rg = lx.Render(csc, dmis=True)
        rg.iterate(1, 2)
        rg.sync()
        rg.reset()
        t0 = time.perf_counter()
        rg.iterate(1, 16)
        rg.sync()
        tg = time.perf_counter() - t0
        gs = rg.stats()
        rg.close()
        cpu["cornell_256_16spp_dmis"] = {
            "cpu_mrays_per_s": (cc["radiance_rays"] + cc["dl_shadow_rays"]) / (cc["wall_ms"] * 1e-3) / 1e6,
            "cpu_threads": threads, "cpu_ms": cc["wall_ms"],
            "gpu_mrays_per_s": (gs["radiance_rays"] + gs["dl_shadow_rays"]) / tg / 1e6, "gpu_ms": tg * 1e3}
        log(f"cpu baseline: one thread {cpu['single_thread_value']:.3f} Mrays/s; cornell {cpu['cornell_256_16spp_dmis']}")

    accuracy = None
    if args.accuracy:
        # Same seeds: the GPU's 16-spp sums on 4,096 seeded pixels against the
        # fp32 oracle (bit for bit) and the FP64 oracle (same random streams,
        # double arithmetic); converged: both 16-spp means against an FP64
        # estimate from 256 later, independent samples per pixel.
        from oracle import lxo
        threads = min(16, os.cpu_count() or 1)
        if o is None:
            o = lxo.OracleScene(data)
        o64 = lxo.OracleScene(data, fp64=True)
        spp, pix_n = 16, 4096
        pix = np.sort(np.random.default_rng(1).choice(W * H, pix_n, replace=False)).astype(np.uint32)
        ra = lx.Render(sc, emis=True)
        ra.iterate(1, spp)
        g = ra.accum()[0].reshape(-1, 3)[pix]
        ra.close()
        f32, _ = o.render(1, spp, pixels=pix, emis=True, threads=threads)
        f64, _ = o64.render(1, spp, pixels=pix, emis=True, threads=threads)
        conv, _ = o64.render(1 + spp, 16 * spp, pixels=pix, emis=True, threads=threads)
        g64, f64d, convd = g.astype(np.float64) / spp, f64.astype(np.float64) / spp, conv.astype(np.float64) / (16 * spp)
        mean = float(convd.mean())
        rmse = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)))
        accuracy = {
            "pixels": pix_n, "spp": spp, "seeds": "reference streams (seed 0), iterations 1..16",
            "bitexact_vs_fp32_oracle": bool(np.array_equal(g.view(np.uint32), f32.view(np.uint32))),
            "rel_rmse_vs_fp64_oracle_same_seeds": rmse(g64, f64d) / mean,
            "rel_rmse_vs_converged_fp64": rmse(g64, convd) / mean,
            "fp64_oracle_rel_rmse_vs_converged_fp64": rmse(f64d, convd) / mean,
            "converged_reference": "FP64 oracle, 256 spp per pixel (iterations 17..272)"}
        log(f"accuracy: {accuracy}")

    batched = served = None
    if not args.autotune:
        ctx.set_tuning(**FORMS_TUNING)
    if args.batched:
        from luxion_amd import workloads
        spec = workloads.mixed_jobs([(sc, dict(emis=True), (0.0, 0.0, 0.0), 3.0)], args.batched, seed=3)
        log(f"batched: {len(spec)} jobs, {workloads.path_samples(spec) / 1e6:.1f} M path-samples")
        batched, _ = workloads.run_batched(ctx, spec)
        log(f"batched: {batched}")
        if not args.autotune:
            # both forms with wide rounds: per-job renders take their samples
            # twenty to a launch, the batch up to 2048 / jobs samples of every
            # job per round
            ctx.set_tuning(round_frames=WORKLOAD_TUNING["round_frames"])
            wide, _ = workloads.run_batched(ctx, spec)
            ctx.set_tuning(**FORMS_TUNING)
            batched["wide_rounds"] = {
                "round_frames": WORKLOAD_TUNING["round_frames"], "batched_s": wide["batched_s"],
                "batched_spp_per_s": wide["batched_spp_per_s"], "per_job_s": wide["per_job_s"],
                "speedup_vs_per_job": wide["speedup_vs_per_job"], "images_bitexact": wide["images_bitexact"]}
            log(f"batched with wide rounds: {batched['wide_rounds']}")
    if args.served != "off":
        from luxion_amd import workloads
        specs = [(sc, dict(emis=True), (0.0, 0.0, 0.0), 3.0)]
        resident = [data.name]
        if args.served == "light":
            cb = lx.Scene(ctx, scenes.cornell(256, 256))
            specs.append((cb, dict(dmis=True), (0.0, 1.0, 0.0), 3.4))
            resident.append("cornell")
        else:
            t0 = time.time()
            ii = scenes.instanced_interior()
            sc2 = lx.Scene(ctx, ii)
            gr = scenes.displaced_grid()
            sc3 = lx.Scene(ctx, gr)
            log(f"served scenes built in {time.time() - t0:.1f} s")
            specs += [(sc2, dict(dmis=True), (0.0, 3.0, 0.0), 40.0), (sc3, dict(emis=True), (0.0, 0.0, 0.0), 14.0)]
            resident += [f"{ii.name} ({ii.num_tris} unique tris)", f"{gr.name} ({gr.num_tris} tris)"]
        srv = lx.Server(ctx, max_jobs=64, max_depth=16)
        try:
            sweep = []
            for k, lam in enumerate((50, 100, 200, 500, 1000, 2000)):
                spec = workloads.mixed_jobs(specs, args.served_jobs, seed=100 + k)
                res_ = workloads.run_served(ctx, spec, lam, seed=200 + k, srv=srv)
                log(f"served lambda={lam}: {res_}")
                sweep.append(res_)
        finally:
            srv.stop()
        served = {"resident_scenes": resident, "job_mix": "side U{64..512}, spp U{1..64}, seeded orbit cameras",
                  "sweep": sweep}

    out = {
        "metric": "Mrays/s + spp/s at 1920x1080 (1 GPU); p50/p99 job latency under mixed load",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_timed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": "single mesh + HDRI: displaced icosphere level %d (%d triangles, 64 primitives) + "
                               "ground plane, 2048x1024 RGBA32F HDRI with 0.5-degree sun of radiance 1e4, "
                               "1920x1080, 1 spp per step, depth 12, -emis (environment IS + MIS)"
                               % (args.level, data.num_tris),
                   "width": W, "height": H, "triangles": data.num_tris, "spp_per_step": 1, "max_depth": depth,
                   "flags": "-emis"},
        "spp_per_s": round(spp_s, 1),
        "rays_per_step": {"radiance": st["radiance_rays"] // args.steps,
                          "env_shadow": st["env_shadow_rays"] // args.steps,
                          "dl_shadow": st["dl_shadow_rays"] // args.steps},
        "sequential_rounds": {"value": round(value_graph, 2), "ms_per_step": round(t_graph * 1e3 / args.steps, 3),
                              "note": "one frame per round, rounds one after another (round_frames 1, "
                                      "overlap_rounds 0), graph replay"},
        "kernel_timing_pass": {"value": round(value_events, 2), "ms_per_step": round(t_events * 1e3 / args.steps, 3),
                               "note": "the timed region's rounds with HIP events around every kernel "
                                       "(source of kernel_ms_per_step and roofline)"},
        "kernel_ms_per_step": {"trace": round(st["trace_ms"] / args.steps, 3),
                               "shade": round(st["shade_ms"] / args.steps, 3),
                               "tail": round(st["tail_ms"] / args.steps, 3),
                               "other": round(st["other_ms"] / args.steps, 3)},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(launch_ms, 4),
                     "frames_per_launch": round(args.steps / max(1, klaunch), 2),
                     "issue": measured_issue(kname),
                     "limiter": "VALU issue with lanes idle behind each wave's longest ray (traversal "
                                "divergence), not HBM bandwidth: HBM traffic is under half the algorithmic bytes",
                     "per_ray": {k: round(v, 2) for k, v in per_ray.items()},
                     "timed_region_frame": {
                         "achieved": round(kbytes / args.steps / (t_timed / args.steps) / 1e9, 1),
                         "frac": round(kbytes / args.steps / (t_timed / args.steps) / 1e9 / HBM_PEAK_GBPS, 4),
                         "note": "the same algorithmic bytes per frame over the timed region's frame time "
                                 "(its rounds shaped by round_frames / overlap_rounds)"}},
        "tuning": tuning,
        "forms_tuning": None if args.autotune else {k: FORMS_TUNING[k] for k in ("sort_mode", "overlap_rounds",
                                                                            "round_frames")},
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
