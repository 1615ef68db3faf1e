/*
 * lx_batch.cpp -- batched form: many render jobs (any scenes, resolutions,
 * spp, depths and flags) share one launch per bounce.
 *
 * The reference renders one job per process, serially (run_best_scenes.ps1:22-50).
 * Here every round takes one sample of every job that still needs samples
 * (up to a path capacity); k_begin_round does the scheduling on the device,
 * so the round is captured once into a hipGraph and replayed.  Each job's
 * pixels keep their own RNG streams (iteration, pixel, depth, job seed), so a
 * job's image is bit-identical whether it runs alone or in a batch.
 */
#include "lx_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void lx_validate_views(int32_t debug_mode, int32_t shading_model) {
    LX_REQUIRE(debug_mode >= LX_DEBUG_OFF && debug_mode <= LX_DEBUG_ALPHA, LX_ERR_INVALID,
               "material_debug_mode must be one of LX_DEBUG_*");
    LX_REQUIRE(shading_model == LX_SHADING_UBER || shading_model == LX_SHADING_LEGACY, LX_ERR_INVALID,
               "shading_model must be LX_SHADING_UBER or LX_SHADING_LEGACY");
}

int lx_job_depth(int32_t max_depth, int trace_depth, int32_t debug_mode) {
    /* a material debug view traces one bounce (src/pathtrace.cu:107) */
    if (debug_mode != LX_DEBUG_OFF) return 1;
    return max_depth > 0 ? max_depth : trace_depth;
}

uint32_t lx_job_flags(int32_t debug_mode, int32_t shading_model) {
    return ((uint32_t)debug_mode & LX_JF_DEBUG_MASK) | (shading_model == LX_SHADING_LEGACY ? LX_JF_LEGACY : 0u);
}

void lx_validate_job(const lx_job& j) {
    LX_REQUIRE(j.scene != nullptr, LX_ERR_INVALID, "job without scene");
    lx_scene* s = j.scene;
    LX_REQUIRE(s->accel_built, LX_ERR_STATE, "job scene: call lx_scene_build_accel first");
    LX_REQUIRE(j.camera.width > 0 && j.camera.height > 0, LX_ERR_INVALID, "job camera resolution must be positive");
    LX_REQUIRE(j.spp >= 1, LX_ERR_INVALID, "job spp must be >= 1");
    LX_REQUIRE(j.first_iter >= 1, LX_ERR_INVALID, "iterations are 1-based");
    int depth = j.max_depth > 0 ? j.max_depth : s->trace_depth;
    LX_REQUIRE(depth <= LX_MAX_DEPTH, LX_ERR_INVALID, "max_depth above the supported maximum");
    lx_validate_views(j.material_debug_mode, j.shading_model);
    if (j.dmis)
        LX_REQUIRE(!s->emissive.empty() && !(s->total_area < 1e-5f), LX_ERR_INVALID,
                   "-dmis requires the scene to contain emissive geometry");
    if (j.emis) LX_REQUIRE(s->has_env, LX_ERR_INVALID, "-emis requires an environment map, pass one with -e");
}

LxJob lx_make_dev_job(const lx_job& j, uint32_t accum_off) {
    LxJob J;
    std::memset(&J, 0, sizeof(J));
    J.scene = j.scene->d_scene.as<LxDevScene>();
    J.cam = j.camera;
    J.num_pix = (uint32_t)j.camera.width * (uint32_t)j.camera.height;
    J.accum_off = accum_off;
    J.depth = lx_job_depth(j.max_depth, j.scene->trace_depth, j.material_debug_mode);
    J.flags = lx_job_flags(j.material_debug_mode, j.shading_model);
    J.dmis = j.dmis;
    J.emis = j.emis;
    J.intensity = j.envmap_intensity;
    J.seed = j.seed;
    J.iter = (int32_t)j.first_iter - 1;
    J.spp_done = 0;
    return J;
}

/* Host replica of k_begin_round's greedy schedule: which round finishes each
 * job, and how many rounds the batch needs. */
int lx_simulate_rounds(const std::vector<uint32_t>& npix, const std::vector<int32_t>& spp, uint32_t cap,
                       std::vector<int>* finish_round) {
    size_t n = npix.size();
    std::vector<int32_t> done(n, 0);
    if (finish_round) finish_round->assign(n, -1);
    int rounds = 0;
    while (true) {
        uint32_t off = 0, na = 0;
        bool any = false;
        for (size_t j = 0; j < n; j++) {
            if (done[j] < spp[j] && na < LX_SCHED_MAX && off + npix[j] <= cap) {
                off += npix[j];
                na++;
                done[j]++;
                any = true;
                if (done[j] == spp[j] && finish_round) (*finish_round)[j] = rounds;
            }
        }
        if (!any) break;
        rounds++;
    }
    return rounds;
}

uint32_t lx_batch_path_cap(lx_context* ctx, const std::vector<uint32_t>& npix, bool dmis, uint64_t extra_bytes,
                           bool paths_in_memory) {
    uint64_t total = 0, mx = 0;
    for (uint32_t p : npix) {
        total += p;
        mx = std::max<uint64_t>(mx, p);
    }
    /* bytes per path slot: two path buffers + records (+ emitter records) + spill share */
    uint64_t per = 2 * 15 * 4 + (11 + 3 + (dmis ? 10 : 0)) * 4;
    uint64_t free_b = ctx->budget > ctx->in_use.load() + extra_bytes ? ctx->budget - ctx->in_use.load() - extra_bytes : 0;
    uint64_t cap = paths_in_memory ? std::min<uint64_t>(total, free_b / per / 2) : total;
    if (ctx->tuning.max_paths_per_round) cap = std::min<uint64_t>(cap, ctx->tuning.max_paths_per_round);
    cap = std::min<uint64_t>(cap, (1ull << 31) - 1);
    LX_REQUIRE(cap >= mx, LX_ERR_BUDGET, "HBM budget too small for the largest job's paths");
    return (uint32_t)cap;
}

extern "C" lx_status lx_batch_run(lx_context* ctx, const lx_job* jobs, uint32_t n, lx_job_result* results) {
    try {
        LX_REQUIRE(ctx && jobs && results && n > 0, LX_ERR_INVALID, "null argument");
        double t_submit = now_ms();
        for (uint32_t i = 0; i < n; i++) lx_validate_job(jobs[i]);
        int width = jobs[0].scene->dev.bvh_width;
        const LxDevScene* uscene = jobs[0].scene->d_scene.as<LxDevScene>();
        for (uint32_t i = 1; i < n; i++)
            if (jobs[i].scene != jobs[0].scene) uscene = nullptr;
        for (uint32_t i = 0; i < n; i++)
            LX_REQUIRE(jobs[i].scene->dev.bvh_width == width, LX_ERR_INVALID,
                       "every job of a batch must reference scenes built at one BVH width");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->set_device();
        std::vector<LxJob> hj(n);
        std::vector<int32_t> hspp(n);
        std::vector<uint32_t> npix(n);
        uint64_t accum_pix = 0;
        bool any_dmis = false, any_tan = false;
        int max_depth = 1;
        for (uint32_t i = 0; i < n; i++) {
            hj[i] = lx_make_dev_job(jobs[i], (uint32_t)accum_pix);
            npix[i] = hj[i].num_pix;
            accum_pix += npix[i];
            hspp[i] = jobs[i].spp;
            any_dmis |= jobs[i].dmis != 0;
            for (auto& m : jobs[i].scene->mats) any_tan |= m.normal_tex >= 0;
            max_depth = std::max(max_depth, hj[i].depth);
        }
        LX_REQUIRE(accum_pix < (1ull << 32), LX_ERR_INVALID, "batch too large");
        DevBuf d_jobs, d_spp, d_cnt, d_sched, d_spill, d_accum;
        d_accum.alloc(ctx, accum_pix * 12);
        /* a round whose tail kernel makes the camera paths keeps no path in
         * memory: no path buffers, and the budget goes to the colour pool */
        const bool paths_mem = !lx_tail_makes_camera(ctx, max_depth);
        uint32_t cap = lx_batch_path_cap(ctx, npix, any_dmis, (uint64_t)lx_spill_words(ctx->num_cus * 16u) * 4 + (1u << 20),
                                         paths_mem);
        /* wide rounds (tuning.round_frames = R > 1): each round takes up to R
         * samples of every job as R copies of it (copy k: samples k, k + R,
         * ... into region k of a colour pool added in sample order), when the
         * budget holds R copies of every job's paths at once */
        uint32_t R = (uint32_t)std::max(1, ctx->tuning.round_frames);
        int32_t max_spp = 1;
        for (int32_t v : hspp) max_spp = std::max(max_spp, v);
        R = std::min<uint32_t>(R, (uint32_t)max_spp);
        R = std::min<uint32_t>(R, LX_SCHED_MAX / std::max<uint32_t>(1u, n)); /* copies the schedule holds */
        bool wide = R > 1 && (uint64_t)n * R <= LX_SCHED_MAX && accum_pix * R < (1ull << 31) &&
                    (uint64_t)cap >= accum_pix;
        DevBuf d_colw;
        if (wide) {
            std::vector<uint32_t> npw;
            for (uint32_t k = 0; k < R; k++) npw.insert(npw.end(), npix.begin(), npix.end());
            try {
                uint32_t capw = lx_batch_path_cap(ctx, npw, any_dmis,
                                                  (uint64_t)lx_spill_words(ctx->num_cus * 16u) * 4 + (1u << 20) +
                                                      accum_pix * R * 12,
                                                  paths_mem);
                wide = (uint64_t)capw >= accum_pix * R;
                if (wide) {
                    d_colw.alloc(ctx, accum_pix * R * 12);
                    cap = capw;
                }
            } catch (const LxError& e) {
                if (e.status != LX_ERR_BUDGET) throw;
                wide = false;
            }
        }
        WaveBufs wb;
        if (paths_mem) wb.alloc(ctx, cap, any_dmis, true, any_tan);
        uint32_t waves = (uint32_t)std::max(1, ctx->tuning.waves_per_cu);
        uint32_t tg = std::min((uint32_t)ctx->num_cus * std::max(1u, waves * 64u / LX_TRACE_BLOCK),
                               std::max(1u, (cap + LX_TRACE_BLOCK - 1) / LX_TRACE_BLOCK));
        uint32_t sg = std::min((uint32_t)ctx->num_cus * 8u, std::max(1u, (cap + LX_SHADE_BLOCK - 1) / LX_SHADE_BLOCK));
        uint32_t cg = std::min((uint32_t)ctx->num_cus * 8u, std::max(1u, (cap + 255u) / 256u));
        d_spill.alloc(ctx, (size_t)lx_spill_words(tg) * 4);
        d_jobs.upload(ctx, hj.data(), n * sizeof(LxJob));
        d_spp.upload(ctx, hspp.data(), n * 4);
        d_cnt.alloc(ctx, sizeof(LxCounters));
        d_sched.alloc(ctx, sizeof(LxSched));
        hipStream_t s = ctx->stream;
        LX_HIP(lx_launch_fill_u32(d_accum.as<uint32_t>(), accum_pix * 3, 0u, s));
        LX_HIP(lx_launch_fill_u32(d_cnt.as<uint32_t>(), sizeof(LxCounters) / 4, 0u, s));
        std::vector<int> fin;
        int rounds = lx_simulate_rounds(npix, hspp, cap, &fin);
        DevBuf d_jobsw, d_sppw;
        std::vector<LxJob> hjw;
        std::vector<int32_t> hsw;
        if (wide) {
            /* copy k of job j (index k n + j): first sample k, step R, region k */
            hjw.resize((size_t)n * R);
            hsw.resize((size_t)n * R);
            rounds = 0;
            for (uint32_t k = 0; k < R; k++)
                for (uint32_t i = 0; i < n; i++) {
                    LxJob J = hj[i];
                    J.iter = hj[i].iter + 1 + (int32_t)k - (int32_t)R;
                    J.flags = (J.flags & ~LX_JF_STEP_MASK) | (R << LX_JF_STEP_SHIFT);
                    J.accum_off = (uint32_t)(k * accum_pix) + hj[i].accum_off;
                    hjw[(size_t)k * n + i] = J;
                    hsw[(size_t)k * n + i] = (int32_t)k < hspp[i] ? (hspp[i] - (int32_t)k + (int32_t)R - 1) / (int32_t)R : 0;
                }
            for (uint32_t i = 0; i < n; i++) {
                fin[i] = (hspp[i] + (int32_t)R - 1) / (int32_t)R - 1;
                rounds = std::max(rounds, fin[i] + 1);
            }
            d_jobsw.upload(ctx, hjw.data(), hjw.size() * sizeof(LxJob));
            d_sppw.upload(ctx, hsw.data(), hsw.size() * 4);
        }
        /* overlapped rounds (tuning.overlap_rounds) need every job in every
         * round until it is done, so that round k is sample k of each job in
         * both forms and the per-pixel sums keep their order */
        bool overlap = !wide && ctx->tuning.overlap_rounds != 0 && rounds >= 2 && n <= LX_SCHED_MAX &&
                       (uint64_t)cap >= accum_pix;
        WaveBufs wb1;
        DevBuf d_jobs1, d_spp1, d_cnt1, d_sched1, d_spill1, d_col[2];
        std::vector<LxJob> hjp[2];
        std::vector<int32_t> hsp[2];
        if (overlap) {
            try {
                if (paths_mem) wb1.alloc(ctx, cap, any_dmis, true, any_tan);
                d_spill1.alloc(ctx, d_spill.bytes);
                d_cnt1.alloc(ctx, sizeof(LxCounters));
                d_sched1.alloc(ctx, sizeof(LxSched));
                d_jobs1.alloc(ctx, n * sizeof(LxJob));
                d_spp1.alloc(ctx, n * 4);
                for (DevBuf& b : d_col) b.alloc(ctx, accum_pix * 12);
            } catch (const LxError& e) {
                if (e.status != LX_ERR_BUDGET) throw;
                overlap = false; /* one round after another in what the budget leaves */
            }
        }
        if (overlap) {
            /* parity p's copy of job j: iterations first + p, first + p + 2, ... */
            for (int p = 0; p < 2; p++) {
                hjp[p] = hj;
                hsp[p].resize(n);
                for (uint32_t i = 0; i < n; i++) {
                    hjp[p][i].iter = hj[i].iter - 1 + p;
                    hjp[p][i].flags |= 2u << LX_JF_STEP_SHIFT;
                    hsp[p][i] = (hspp[i] + 1 - p) / 2;
                }
            }
            LX_HIP(hipMemcpyAsync(d_jobs.ptr, hjp[0].data(), n * sizeof(LxJob), hipMemcpyHostToDevice, s));
            LX_HIP(hipMemcpyAsync(d_spp.ptr, hsp[0].data(), n * 4, hipMemcpyHostToDevice, s));
            LX_HIP(hipMemcpyAsync(d_jobs1.ptr, hjp[1].data(), n * sizeof(LxJob), hipMemcpyHostToDevice, s));
            LX_HIP(hipMemcpyAsync(d_spp1.ptr, hsp[1].data(), n * 4, hipMemcpyHostToDevice, s));
            LX_HIP(lx_launch_fill_u32(d_cnt1.as<uint32_t>(), sizeof(LxCounters) / 4, 0u, s));
        }
        /* a round on the context stream: with wide rounds, the job copies'
         * colour regions filled with -0.0, the round, then the regions added
         * into the accumulation in sample order */
        auto plain_round = [&]() {
            if (wide) {
                LX_HIP(lx_launch_fill_u32(d_colw.as<uint32_t>(), accum_pix * R * 3, 0x80000000u, s));
                lx_enqueue_round(ctx, d_jobsw.as<LxJob>(), d_sppw.as<int32_t>(), n * R, d_cnt.as<LxCounters>(),
                                 d_sched.as<LxSched>(), cap, wb, d_colw.as<float>(), d_spill.as<uint32_t>(), max_depth,
                                 tg, sg, cg, width, uscene, nullptr, false, nullptr, any_dmis, true);
                LX_HIP(lx_launch_add_regions(d_accum.as<float>(), d_colw.as<float>(), accum_pix * 3, R, s));
            } else {
                lx_enqueue_round(ctx, d_jobs.as<LxJob>(), d_spp.as<int32_t>(), n, d_cnt.as<LxCounters>(),
                                 d_sched.as<LxSched>(), cap, wb, d_accum.as<float>(), d_spill.as<uint32_t>(), max_depth,
                                 tg, sg, cg, width, uscene, nullptr, false, nullptr, any_dmis);
            }
        };
        /* one captured round, replayed (every round is the same launches over
         * device-side job state; no memset nodes); an event after every round
         * dates job completion */
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        bool use_graph = ctx->tuning.use_graph != 0 && !overlap;
        if (use_graph) {
            LX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                plain_round();
            } catch (...) {
                hipGraph_t tmp;
                (void)hipStreamEndCapture(s, &tmp);
                throw;
            }
            LX_HIP(hipStreamEndCapture(s, &g));
            LX_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        }
        std::vector<hipEvent_t> ev(rounds + 1);
        for (auto& e : ev) LX_HIP(hipEventCreate(&e));
        /* overlapped form: round k on stream k % 2 into colour pool k % 2
         * (filled with -0.0: the one add per pixel copies the colour's bits),
         * added into the accumulation on the context's stream in round order */
        hipStream_t ps[2] = {nullptr, nullptr};
        hipEvent_t e_up = nullptr, e_done[2] = {nullptr, nullptr}, e_free[2] = {nullptr, nullptr};
        struct Cleanup {
            hipStream_t* ps;
            hipEvent_t *up, *done, *free;
            ~Cleanup() {
                for (int p = 0; p < 2; p++) {
                    if (ps[p]) (void)hipStreamDestroy(ps[p]);
                    if (done[p]) (void)hipEventDestroy(done[p]);
                    if (free[p]) (void)hipEventDestroy(free[p]);
                }
                if (*up) (void)hipEventDestroy(*up);
            }
        } cleanup{ps, &e_up, e_done, e_free};
        if (overlap) {
            for (int p = 0; p < 2; p++) {
                LX_HIP(hipStreamCreateWithFlags(&ps[p], hipStreamNonBlocking));
                LX_HIP(hipEventCreateWithFlags(&e_done[p], hipEventDisableTiming));
                LX_HIP(hipEventCreateWithFlags(&e_free[p], hipEventDisableTiming));
            }
            LX_HIP(hipEventCreateWithFlags(&e_up, hipEventDisableTiming));
            LX_HIP(hipEventRecord(e_up, s));
            for (int p = 0; p < 2; p++) LX_HIP(hipStreamWaitEvent(ps[p], e_up, 0));
        }
        LxJob* pj[2] = {d_jobs.as<LxJob>(), d_jobs1.as<LxJob>()};
        int32_t* pspp[2] = {d_spp.as<int32_t>(), d_spp1.as<int32_t>()};
        LxCounters* pcnt[2] = {d_cnt.as<LxCounters>(), d_cnt1.as<LxCounters>()};
        LxSched* psch[2] = {d_sched.as<LxSched>(), d_sched1.as<LxSched>()};
        uint32_t* pspill[2] = {d_spill.as<uint32_t>(), d_spill1.as<uint32_t>()};
        WaveBufs* pwb[2] = {&wb, &wb1};
        LX_HIP(hipEventRecord(ev[0], s));
        double t_start = now_ms();
        for (int k = 0; k < rounds; k++) {
            if (overlap) {
                int p = k & 1;
                if (k >= 2) LX_HIP(hipStreamWaitEvent(ps[p], e_free[p], 0));
                LX_HIP(lx_launch_fill_u32(d_col[p].as<uint32_t>(), accum_pix * 3, 0x80000000u, ps[p]));
                lx_enqueue_round(ctx, pj[p], pspp[p], n, pcnt[p], psch[p], cap, *pwb[p], d_col[p].as<float>(),
                                 pspill[p], max_depth, tg, sg, cg, width, uscene, nullptr, false, ps[p], any_dmis, true);
                LX_HIP(hipEventRecord(e_done[p], ps[p]));
                LX_HIP(hipStreamWaitEvent(s, e_done[p], 0));
                LX_HIP(lx_launch_add_f32(d_accum.as<float>(), d_col[p].as<float>(), accum_pix * 3, s));
                LX_HIP(hipEventRecord(e_free[p], s));
            } else if (use_graph) {
                LX_HIP(hipGraphLaunch(ge, s));
            } else {
                plain_round();
            }
            LX_HIP(hipEventRecord(ev[k + 1], s));
        }
        LX_HIP(hipStreamSynchronize(s));
        double t_end = now_ms();
        LxCounters c;
        LX_HIP(hipMemcpy(&c, d_cnt.ptr, sizeof(c), hipMemcpyDeviceToHost));
        if (overlap) {
            LxCounters c1;
            LX_HIP(hipMemcpy(&c1, d_cnt1.ptr, sizeof(c1), hipMemcpyDeviceToHost));
            lx_counters_add(c, c1);
        }
        std::vector<float> round_ms(rounds + 1, 0.0f);
        for (int k = 1; k <= rounds; k++) LX_HIP(hipEventElapsedTime(&round_ms[k], ev[0], ev[k]));
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        LX_REQUIRE(c.stack_overflow == 0, LX_ERR_STATE, "traversal stack overflow");
        std::vector<LxJob> done(n);
        if (wide) {
            LX_HIP(hipMemcpy(hjw.data(), d_jobsw.ptr, hjw.size() * sizeof(LxJob), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; i++) {
                done[i] = hjw[i];
                for (uint32_t k = 1; k < R; k++) done[i].spp_done += hjw[(size_t)k * n + i].spp_done;
            }
        } else {
            LX_HIP(hipMemcpy(done.data(), d_jobs.ptr, n * sizeof(LxJob), hipMemcpyDeviceToHost));
        }
        if (overlap) {
            std::vector<LxJob> done1(n);
            LX_HIP(hipMemcpy(done1.data(), d_jobs1.ptr, n * sizeof(LxJob), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; i++) done[i].spp_done += done1[i].spp_done;
        }
        uint64_t tot_rays = c.rad_rays + c.dl_rays + c.env_rays;
        for (uint32_t i = 0; i < n; i++) {
            lx_job_result& r = results[i];
            if (r.rgb_sum)
                LX_HIP(hipMemcpy(r.rgb_sum, (uint8_t*)d_accum.ptr + (size_t)hj[i].accum_off * 12, (size_t)npix[i] * 12,
                                 hipMemcpyDeviceToHost));
            r.status = done[i].spp_done == hspp[i] ? LX_OK : LX_ERR_STATE;
            r.spp_done = (uint32_t)done[i].spp_done;
            /* rays are counted for the batch; attributed to jobs by path-samples */
            r.rays = (uint64_t)((double)tot_rays * ((double)npix[i] * hspp[i]) /
                                std::max(1.0, [&]() {
                                    double t = 0;
                                    for (uint32_t k = 0; k < n; k++) t += (double)npix[k] * hspp[k];
                                    return t;
                                }()));
            r.submit_ms = t_submit;
            r.start_ms = t_start;
            r.finish_ms = fin[i] >= 0 ? t_start + round_ms[fin[i] + 1] : t_end;
        }
        lx_set_error("");
        return LX_OK;
    } catch (const LxError& e) {
        lx_set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc&) {
        lx_set_error("host allocation failed");
        return LX_ERR_OOM;
    }
}
