"""Benchmark: Mrays/s + spp/s of the wavefront path tracer at 1920x1080 on one
MI355X.

Workloads (--workload):
  hdri      (default) single mesh + HDRI: 1,310,720-triangle displaced
            icosphere in 64 primitives + ground plane, 2048x1024 HDRI with a
            0.5-degree sun of radiance 1e4, depth 12, environment importance
            sampling + MIS (-emis).
  interior  instanced interior: 64 procedural meshes of 10k-500k triangles
            under 10,008 instances, 8 emissive meshes, glass, 1024^2 checker
            albedo and alpha-mask textures, depth 16, emitter NEE + MIS
            (-dmis).

A step is one 1-spp frame: camera rays, then every bounce in the persistent
tail kernel (radiance + shadow rays, uber-shader, MIS, Russian roulette,
accumulation).  The headline `value` renders one frame per launch (one
sample per round), with the next frame's round in flight on a second stream
while the current one drains; the same frames twenty to a launch
(`wide_rounds`) and one round after another (`sequential_rounds`) are
reported beside it.  Inputs are resident in HBM before every timed region.
Per-kernel times for the roofline come from a separate pass with HIP events
around every kernel of rounds run one after another.  Output: one JSON line.
"""
import argparse
import hashlib
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
# launch shapes applied unless --autotune 1 (fixed, so that runs and the
# profiles of the same command time the same kernels): tail kernel for every
# bounce (making its camera paths itself), 128-lane (HDRI) / 256-lane
# (interior) blocks, 16 waves per CU, BVH4, no queue sort, refills once 48
# (HDRI) / 8 (interior) lanes of a wave are idle; the headline renders one
# frame per round with one more round in flight (HDRI) or two (interior)
WORKLOAD_TUNING = {
    "hdri": dict(wavefront_bounces=0, block_size=128, waves_per_cu=16, queue_chunk=48, tail_min_waves=4, bvh_width=4,
                 sort_mode=0, pixel_tile=16, tile_band=8, overlap_rounds=1, round_frames=1),
    "interior": dict(wavefront_bounces=0, block_size=256, waves_per_cu=16, queue_chunk=8, tail_min_waves=4,
                     bvh_width=4, sort_mode=0, pixel_tile=8, tile_band=0, overlap_rounds=2, round_frames=1),
}
WIDE_FRAMES = 20
# the batched and served legs' plain per-job launches: one sample per round,
# overlapped, no queue sort
FORMS_TUNING = dict(WORKLOAD_TUNING["hdri"], block_size=256, sort_mode=0, overlap_rounds=1, round_frames=1)
# device sources of the kernels the profiles measure: a profile's figures are
# this build's only if its recorded digest matches
DEVICE_SOURCES = ["luxion_amd/csrc/lx_kernels.hip", "luxion_amd/csrc/lx_kernels.h", "luxion_amd/csrc/lx_path_dev.h",
                  "luxion_amd/csrc/lx_trace_dev.h", "luxion_amd/csrc/lx_shade_dev.h", "luxion_amd/csrc/lx_types.h",
                  "include/luxion_detmath.h", "include/luxion_mi355x.h"]


def device_source_digest():
    h = hashlib.sha256()
    for rel in DEVICE_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profile_figures(kernel, workload):
    """HBM traffic per launch and VALU issue fractions of `kernel` from the
    committed rocprofv3 summary profiles/<kernel>_<workload>.md (lines
    'traffic_bytes_per_launch: N', 'valu_busy: X', 'lane_utilisation: Y',
    'device_source_digest: D'), stamped with whether the summary was made
    from these device sources; None if there is no summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                k, _, v = line.strip().partition(":")
                if k in ("traffic_bytes_per_launch", "valu_busy", "lane_utilisation", "device_source_digest",
                         "fetch_bytes_per_launch", "write_bytes_per_launch", "valu_insts_per_ray"):
                    out[k] = v.strip()
    except OSError:
        return None
    if not out:
        return None
    digest = out.pop("device_source_digest", None)
    fig = {k: float(v) for k, v in out.items()}
    fig["source"] = f"profiles/{kernel}_{workload}.md"
    fig["from_this_build"] = digest == device_source_digest()
    fig["note"] = ("rocprofv3 counters of a separate profiling run of this command, not of this run; "
                   "from_this_build: the summary's device-source digest matches these sources")
    return fig


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity": affinity, "model": model}


def rays_of(st):
    return st["radiance_rays"] + st["dl_shadow_rays"] + st["env_shadow_rays"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["hdri", "interior"], default="hdri")
    ap.add_argument("--legs", default="timed,wide,sequential,events,count",
                    help="render legs to run (comma list of timed, wide, sequential, events, count); a profiling "
                         "run names only the legs it profiles")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (the GPU box's CPU share for one GPU is 16)")
    ap.add_argument("--accuracy", type=int, default=1,
                    help="1: compare a 16-spp render of the workload with the fp32 oracle (bits) and the FP64 "
                         "oracle (RMSE) on 4,096 seeded pixels")
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--batched", type=int, default=256,
                    help="jobs in the batched-form mix on the HDRI scene (0 = skip)")
    ap.add_argument("--served", choices=["off", "light", "full"], default="full",
                    help="served-form Poisson sweep: full = the HDRI scene + instanced interior + 30M-triangle "
                         "grid resident; light = the HDRI scene + a Cornell box")
    ap.add_argument("--served-jobs", type=int, default=200, help="jobs per arrival rate")
    ap.add_argument("--wavefront-bounces", type=int, default=None,
                    help="bounces launched one kernel each before the tail kernel (default: the workload's)")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: run the library's autotuner on this workload (8-spp job) before timing instead "
                         "of applying WORKLOAD_TUNING")
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
    legs = set(args.legs.split(","))

    import numpy as np
    import luxion_amd as lx
    from luxion_amd import scenes

    W, H = 1920, 1080
    t0 = time.time()
    if args.workload == "hdri":
        data = scenes.icosphere_hdri(W, H, level=args.level)
        flags, flag_name = dict(emis=True), "-emis"
    else:
        data = scenes.instanced_interior(W, H)
        flags, flag_name = dict(dmis=True), "-dmis"
    log(f"scene generated: {data.num_tris} unique triangles in {time.time() - t0:.1f} s")
    ctx = lx.Context(0)
    t0 = time.time()
    sc = lx.Scene(ctx, data)
    info = sc.info()
    log(f"scene on device: {info['device_bytes'] / 2**20:.1f} MiB, BVH {info['bvh_nodes']} nodes, "
        f"build {info['build_ms']:.0f} ms")
    depth = data.trace_depth
    tune = WORKLOAD_TUNING[args.workload]

    if args.autotune:
        t0 = time.time()
        best = ctx.autotune([lx.make_job(sc, data.camera, 8, **flags)])
        log(f"autotuned in {time.time() - t0:.1f} s: {best}")
    else:
        ctx.set_tuning(**tune)
    if args.wavefront_bounces is not None:
        ctx.set_tuning(wavefront_bounces=args.wavefront_bounces)
    for kv in args.set:
        k, v = kv.split("=", 1)
        ctx.set_tuning(**{k: int(v)})
    tuning = ctx.tuning()
    r = lx.Render(sc, **flags)
    r.iterate(1, args.warmup)
    r.sync()
    log(f"warmup done ({args.warmup} steps)")
    first = args.warmup + 1
    steps = args.steps

    def timed(n=steps):
        r.reset()
        t = time.perf_counter()
        r.iterate(first, n)
        r.sync()
        return time.perf_counter() - t, r.stats()

    # the headline: one frame per launch as a caller renders them (rounds
    # overlapped as overlap_rounds says), synchronised on both sides
    t_timed = st_main = None
    if "timed" in legs:
        t_timed, st_main = timed()
        log(f"timed region: {steps} steps in {t_timed * 1e3:.1f} ms, {rays_of(st_main)} rays "
            f"(overlap_rounds {tuning['overlap_rounds']}, round_frames {tuning['round_frames']})")

    # the same frames one round after another, replayed from a captured hipGraph
    t_graph = st_graph = None
    if "sequential" in legs:
        ctx.set_tuning(overlap_rounds=0, round_frames=1)
        r.reset()
        r.iterate(first, 1)
        r.sync()
        t_graph, st_graph = timed()
        log(f"sequential rounds (graph replay): {steps} steps in {t_graph * 1e3:.1f} ms")

    # kernel timing pass: one frame per launch, rounds one after another,
    # with HIP events before and after each kernel on the stream it runs on
    # (so each launch has the GPU to itself: per-kernel device time for the
    # roofline)
    st = None
    if "events" in legs:
        ctx.set_tuning(overlap_rounds=0, round_frames=1)
        r.set_timing(events=True)
        t_events, st = timed()
        r.set_timing(events=False)
        if st_main is not None:
            assert st["radiance_rays"] == st_main["radiance_rays"], "kernel timing pass traced a different ray batch"
        log(f"kernel timing pass: {steps} steps in {t_events * 1e3:.1f} ms")

    # the same frames twenty to a launch (one round of twenty job copies)
    t_wide = st_wide = None
    if "wide" in legs:
        ctx.set_tuning(overlap_rounds=0, round_frames=WIDE_FRAMES)
        r.reset()
        r.iterate(first, WIDE_FRAMES)
        r.sync()
        t_wide, st_wide = timed()
        log(f"wide rounds ({WIDE_FRAMES} frames per launch): {steps} steps in {t_wide * 1e3:.1f} ms")
    ctx.set_tuning(**{k: tuning[k] for k in ("overlap_rounds", "round_frames")})
    r.close()  # frees the wide rounds' path buffers before the other legs

    # counting replay of the same frames: node / triangle / instance visits
    sc_ = None
    if "count" in legs:
        rc = lx.Render(sc, **flags)
        rc.set_timing(events=False, count_work=True)
        rc.iterate(first, steps)
        sc_ = rc.stats()
        rc.close()
        if st is not None:
            assert sc_["radiance_rays"] == st["radiance_rays"], "counting replay traced a different ray batch"

    out = {
        "metric": "Mrays/s + spp/s at 1920x1080 (1 GPU); p50/p99 job latency under mixed load",
        "value": None,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural scene, seeded)",
    }
    if args.workload == "hdri":
        wl = ("single mesh + HDRI: displaced icosphere level %d (%d triangles, 64 primitives) + ground plane, "
              "2048x1024 RGBA32F HDRI with 0.5-degree sun of radiance 1e4, 1920x1080, 1 spp per launch, depth 12, "
              "-emis (environment IS + MIS)" % (args.level, data.num_tris))
    else:
        wl = ("instanced interior: 64 procedural meshes (%d unique triangles, %d instanced) under %d instances, "
              "8 emissive meshes, glass (T 1, IOR 1.5), 1024^2 RGBA8 checker albedo + alpha-mask textures (MASK), "
              "1920x1080, 1 spp per launch, depth 16, -dmis (emitter NEE + MIS)"
              % (data.num_tris, data.num_instanced_tris, len(data.instances)))
    out["config"] = {"workload": wl, "width": W, "height": H, "triangles": data.num_tris, "spp_per_step": 1,
                     "max_depth": depth, "flags": flag_name,
                     "launch_shape": "one frame (1 spp) per tail-kernel launch; %d round(s) in flight"
                                     % (1 + tuning["overlap_rounds"])}
    out["accel_build"] = {"ms": round(info["build_ms"], 1), "bvh_nodes": info["bvh_nodes"],
                          "sah_cost": round(info["sah_cost"], 4),
                          "builder": "auto: BLASes of 2^14 triangles or more on the GPU, smaller ones on the host"}
    cam_rays = W * H
    if st_main is not None:
        rays = rays_of(st_main)
        out["value"] = round(rays / t_timed / 1e6, 2)
        out["ms_per_step"] = round(t_timed * 1e3 / steps, 3)
        out["spp_per_s"] = round(W * H * steps / t_timed, 1)
        sec = st_main["radiance_rays"] - cam_rays * steps
        out["rays_per_step"] = {"camera": cam_rays, "secondary_radiance": sec // steps,
                                "env_shadow": st_main["env_shadow_rays"] // steps,
                                "dl_shadow": st_main["dl_shadow_rays"] // steps}
        out["secondary_plus_shadow_mrays_per_s"] = round(
            (sec + st_main["env_shadow_rays"] + st_main["dl_shadow_rays"]) / t_timed / 1e6, 2)
    if st_graph is not None:
        out["sequential_rounds"] = {"value": round(rays_of(st_graph) / t_graph / 1e6, 2),
                                    "ms_per_step": round(t_graph * 1e3 / steps, 3),
                                    "note": "one frame per launch, rounds one after another (overlap_rounds 0), "
                                            "graph replay"}
    if st_wide is not None:
        out["wide_rounds"] = {"value": round(rays_of(st_wide) / t_wide / 1e6, 2),
                              "ms_per_step": round(t_wide * 1e3 / steps, 3), "frames_per_launch": WIDE_FRAMES,
                              "note": "the same frames, twenty per tail-kernel launch (round_frames 20), one round "
                                      "at a time; not the headline"}
    if st is not None:
        out["kernel_ms_per_step"] = {"trace": round(st["trace_ms"] / steps, 3),
                                     "shade": round(st["shade_ms"] / steps, 3),
                                     "tail": round(st["tail_ms"] / steps, 3),
                                     "other": round(st["other_ms"] / steps, 3)}
    if st is not None and sc_ is not None:
        def trace_bytes(n_closest, n_any, c):
            return (n_closest * (RAY_BYTES + HIT_BYTES) + n_any * (RAY_BYTES + OCC_BYTES) +
                    INST_BYTES * (c["closest_insts"] + c["any_insts"]) +
                    NODE_BYTES * (c["closest_nodes"] + c["any_nodes"]) +
                    TRI_BYTES * (c["closest_tris"] + c["any_tris"]))

        n_closest = sc_["radiance_rays"] + sc_["dl_shadow_rays"]
        n_any = sc_["env_shadow_rays"]
        tail = {k[5:]: v for k, v in sc_.items() if k.startswith("tail_") and k[5:] in
                ("closest_nodes", "closest_tris", "closest_insts", "any_nodes", "any_tris", "any_insts")}
        wave = {k: sc_[k] - tail[k] for k in tail}
        tail_bytes = trace_bytes(sc_["tail_closest_rays"], sc_["tail_any_rays"], tail)
        wave_bytes = trace_bytes(n_closest - sc_["tail_closest_rays"], n_any - sc_["tail_any_rays"], wave)
        per_ray = {"nodes_closest": sc_["closest_nodes"] / max(1, n_closest),
                   "tris_closest": sc_["closest_tris"] / max(1, n_closest),
                   "insts_closest": sc_["closest_insts"] / max(1, n_closest),
                   "nodes_any": sc_["any_nodes"] / max(1, n_any), "tris_any": sc_["any_tris"] / max(1, n_any)}
        if st["tail_ms"] >= st["trace_ms"]:
            kname, kms, klaunch = "k_finish", st["tail_ms"], st["tail_launches"]
            kbytes = tail_bytes + SHADE_BYTES * sc_["tail_bounces"]
        else:
            kname, kms, klaunch = "k_trace", st["trace_ms"], st["trace_launches"]
            kbytes = wave_bytes
        launch_ms = kms / max(1, klaunch)
        bytes_per_launch = kbytes / max(1, klaunch)
        achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
        prof = profile_figures(kname, args.workload)
        roof = {"bound": "valu", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": int(prof["traffic_bytes_per_launch"]) if prof and "traffic_bytes_per_launch" in prof
                else None,
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(launch_ms, 4),
                "rays_per_frame": int(rays_of(st) // steps),
                "frames_per_launch": round(steps / max(1, klaunch), 2),
                "frac_basis": "algorithmic bytes (counting replay: ray, hit, node, triangle, instance bytes + 380 B "
                              "per shaded path-bounce) / event-timed launch / 8 TB/s",
                "limiter": "VALU issue: lanes idle behind each wave's longest ray (traversal divergence); HBM "
                           "traffic is below the algorithmic bytes (BVH and triangles largely cached), so the "
                           "HBM roof is not the binding one",
                "per_ray": {k: round(v, 2) for k, v in per_ray.items()},
                "profile": prof}
        if prof and "traffic_bytes_per_launch" in prof:
            roof["measured_traffic_frac"] = round(prof["traffic_bytes_per_launch"] / (launch_ms * 1e-3) / 1e9
                                                  / HBM_PEAK_GBPS, 4)
        if t_timed is not None:
            roof["timed_region_frame"] = {
                "achieved": round(kbytes / steps / (t_timed / steps) / 1e9, 1),
                "frac": round(kbytes / steps / (t_timed / steps) / 1e9 / HBM_PEAK_GBPS, 4),
                "note": "the same algorithmic bytes per frame over the headline's frame time (overlapped rounds)"}
        out["roofline"] = roof
    out["tuning"] = tuning

    ci = cpu_info()
    cpu = None
    o = None
    if args.cpu_baseline and args.workload == "hdri":
        from oracle import lxo
        threads = max(1, min(args.cpu_threads, ci["affinity"]))
        o = lxo.OracleScene(data)
        log(f"oracle BVH built ({lxo.lib().lxo_scene_bvh_build_ms(o.h):.0f} ms); timing CPU baseline")
        rng = np.random.default_rng(0)
        tot_rays, tot_ms, iters = 0, 0.0, 0
        while tot_ms < args.cpu_seconds * 1e3 and iters < 8:
            _, cst = o.render(first + iters, 1, emis=True, threads=threads)
            tot_rays += rays_of(cst)
            tot_ms += cst["wall_ms"]
            iters += 1
            log(f"cpu baseline: {iters} frame(s), {tot_ms / 1e3:.1f} s")
        cpu = {"value": tot_rays / (tot_ms * 1e-3) / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{iters} full 1920x1080 frame(s) at 1 spp of the same workload (iterations "
                         f"{first}..{first + iters - 1}), oracle/lxo_oracle.cpp restatement with its "
                         f"reference-style binary SAH BVH, {threads} threads; BVH build excluded",
               "host": dict(ci, threads_used=threads,
                            note="threads capped at the GPU box's CPU share for one GPU (16)")}
        pix = np.sort(rng.choice(W * H, 200000, replace=False)).astype(np.uint32)
        _, c1 = o.render(first, 1, pixels=pix, emis=True, threads=1)
        cpu["single_thread_value"] = rays_of(c1) / (c1["wall_ms"] * 1e-3) / 1e6
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
        log(f"cpu baseline: one thread {cpu['single_thread_value']:.3f} Mrays/s; cornell "
            f"{cpu['cornell_256_16spp_dmis']}")
    out["cpu_baseline"] = cpu

    accuracy = None
    if args.accuracy:
        # Same seeds: the GPU's 16-spp sums on 4,096 seeded pixels against the
        # fp32 oracle (bit for bit) and the FP64 oracle (same random streams,
        # double arithmetic); converged: both 16-spp means against an FP64
        # estimate from 256 later, independent samples per pixel.
        from oracle import lxo
        threads = max(1, min(args.cpu_threads, ci["affinity"]))
        if o is None:
            o = lxo.OracleScene(data)
        o64 = lxo.OracleScene(data, fp64=True)
        spp, pix_n = 16, 4096
        conv_spp = 16 * spp if args.workload == "hdri" else 4 * spp
        pix = np.sort(np.random.default_rng(1).choice(W * H, pix_n, replace=False)).astype(np.uint32)
        ra = lx.Render(sc, **flags)
        ra.iterate(1, spp)
        g = ra.accum()[0].reshape(-1, 3)[pix]
        ra.close()
        f32, _ = o.render(1, spp, pixels=pix, threads=threads, **flags)
        f64, _ = o64.render(1, spp, pixels=pix, threads=threads, **flags)
        conv, _ = o64.render(1 + spp, conv_spp, pixels=pix, threads=threads, **flags)
        g64, f64d, convd = g.astype(np.float64) / spp, f64.astype(np.float64) / spp, conv.astype(np.float64) / conv_spp
        mean = float(convd.mean())
        rmse = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)))
        accuracy = {
            "pixels": pix_n, "spp": spp, "seeds": "reference streams (seed 0), iterations 1..16",
            "bitexact_vs_fp32_oracle": bool(np.array_equal(g.view(np.uint32), f32.view(np.uint32))),
            "rel_rmse_vs_fp64_oracle_same_seeds": rmse(g64, f64d) / mean,
            "rel_rmse_vs_converged_fp64": rmse(g64, convd) / mean,
            "fp64_oracle_rel_rmse_vs_converged_fp64": rmse(f64d, convd) / mean,
            "converged_reference": "FP64 oracle, %d spp per pixel (iterations 17..%d)" % (conv_spp, 16 + conv_spp)}
        log(f"accuracy: {accuracy}")
    out["accuracy"] = accuracy

    batched = served = None
    hdri_sc = sc if args.workload == "hdri" else None
    if args.batched and hdri_sc is not None:
        from luxion_amd import workloads
        spec = workloads.mixed_jobs([(hdri_sc, dict(emis=True), (0.0, 0.0, 0.0), 3.0)], args.batched, seed=3)
        log(f"batched: {len(spec)} jobs, {workloads.path_samples(spec) / 1e6:.1f} M path-samples")
        # plain per-job launches, each job through its own render, at the two
        # per-job shapes: one sample per launch (overlapped) and twenty
        ctx.set_tuning(**FORMS_TUNING)
        per1 = workloads.run_per_job(spec)
        ctx.set_tuning(**dict(FORMS_TUNING, overlap_rounds=0, round_frames=WIDE_FRAMES))
        per20 = workloads.run_per_job(spec)
        best_per = min(per1["per_job_s"], per20["per_job_s"])
        # the batched form as one launch per bounce for all jobs, rounds
        # captured in a hipGraph and replayed, launch knobs from the autotuner
        # (timed on a 32-job, 4-spp slice of the same mix), several samples of
        # every job per round
        ctx.set_tuning(**dict(FORMS_TUNING, overlap_rounds=0, round_frames=WIDE_FRAMES))
        t0 = time.time()
        tune_jobs = [lx.make_job(s_, cam, **dict(kw, spp=min(4, kw["spp"]))) for s_, cam, kw in spec[:32]]
        best = ctx.autotune(tune_jobs)
        ctx.set_tuning(overlap_rounds=0, round_frames=WIDE_FRAMES, use_graph=1)
        log(f"batched: autotuned in {time.time() - t0:.1f} s: {best}")
        bres, res = workloads.run_batched(ctx, spec, per_job=False)
        same = all(np.array_equal(a, b["image"]) for a, b in zip(per20["images"], res))
        batched = dict(bres)
        batched.update({
            "form": "one launch per bounce for all jobs; rounds captured in a hipGraph and replayed; up to %d "
                    "samples of every job per round (at most 2048 job copies per round); autotuned launch "
                    "knobs" % WIDE_FRAMES,
            "autotuned": {k: best[k] for k in ("block_size", "bvh_width", "queue_chunk", "waves_per_cu",
                                               "sort_mode", "pixel_tile")},
            "per_job_one_sample_per_launch_s": per1["per_job_s"],
            "per_job_twenty_samples_per_launch_s": per20["per_job_s"],
            "speedup_vs_per_job_one_sample_per_launch": round(per1["per_job_s"] / bres["batched_s"], 3),
            "speedup_vs_strongest_per_job": round(best_per / bres["batched_s"], 3),
            "baseline_note": "speedup_vs_strongest_per_job compares with the faster of the two per-job shapes "
                             "(each job rendered alone through its own render)",
            "images_bitexact": bool(same)})
        log(f"batched: {batched}")
    out["batched"] = batched
    if args.served != "off" and hdri_sc is not None:
        from luxion_amd import workloads
        ctx.set_tuning(**FORMS_TUNING)
        specs = [(hdri_sc, dict(emis=True), (0.0, 0.0, 0.0), 3.0)]
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
            resident += [f"{ii.name} ({ii.num_tris} unique tris, {len(ii.instances)} instances)",
                         f"{gr.name} ({gr.num_tris} tris)"]
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
        served = {"resident_scenes": resident, "jobs_per_rate": args.served_jobs,
                  "job_mix": "side U{64..512}, spp U{1..64}, seeded orbit cameras, scenes in turn", "sweep": sweep}
    out["served"] = served
    out["forms_tuning"] = {k: FORMS_TUNING[k] for k in ("sort_mode", "overlap_rounds", "round_frames")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
