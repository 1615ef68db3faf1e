/*
 * lx_render.cpp -- bounce scheduler for one render job.
 *
 * Replaces pathtraceInit / pathtrace / pathtraceFree / copyImageToHost
 * (src/pathtrace/pathtrace_buffers.cu:21-190, src/pathtrace.cu:97-367).
 * The reference synchronises with the host every bounce (a blocking read of
 * the live-path count, pathtrace.cu:317, plus a device-wide sync in
 * checkCUDAError); here each round is a fixed sequence of launches whose
 * sizes come from device counters, so a round is captured once into a
 * hipGraph and replayed per sample with no host round trip.
 */
#include "lx_internal.h"

#include <algorithm>
#include <chrono>


uint32_t lx_spill_words(uint32_t trace_grid) { return (uint32_t)LX_SPILL * trace_grid * LX_TRACE_BLOCK; }

static void carve(uint8_t*& p, float** f, size_t n) {
    *f = (float*)p;
    p += ((n * 4 + 255) / 256) * 256;
}

void WaveBufs::alloc(lx_context* ctx, uint32_t c, bool dmis, bool emis, bool tangents) {
    (void)emis;
    cap = c;
    size_t arr = ((size_t)c * 4 + 255) / 256 * 256;
    pathA.alloc(ctx, 15 * arr);
    pathB.alloc(ctx, 15 * arr);
    uint8_t* pa = (uint8_t*)pathA.ptr;
    uint8_t* pb = (uint8_t*)pathB.ptr;
    LxPathBuf* bufs[2] = {&A, &B};
    uint8_t* bases[2] = {pa, pb};
    for (int k = 0; k < 2; k++) {
        uint8_t* p = bases[k];
        LxPathBuf& P = *bufs[k];
        carve(p, &P.ox, c);
        carve(p, &P.oy, c);
        carve(p, &P.oz, c);
        carve(p, &P.dx, c);
        carve(p, &P.dy, c);
        carve(p, &P.dz, c);
        carve(p, &P.tx, c);
        carve(p, &P.ty, c);
        carve(p, &P.tz, c);
        carve(p, &P.cx, c);
        carve(p, &P.cy, c);
        carve(p, &P.cz, c);
        carve(p, &P.pdf, c);
        carve(p, (float**)&P.pix, c);
        carve(p, (float**)&P.job, c);
    }
    /* sort stage: keys and permutation per path, one histogram */
    sort.alloc(ctx, 2 * arr + ((size_t)LX_SORT_HIST * 4 + 255) / 256 * 256);
    keys = (uint32_t*)sort.ptr;
    order = (uint32_t*)((uint8_t*)sort.ptr + arr);
    hist = (uint32_t*)((uint8_t*)sort.ptr + 2 * arr);
    /* records: 7 always, 3 tangent, 4 environment, 10 direct light */
    size_t nrec = 7 + 4 + (tangents ? 3 : 0) + (dmis ? 10 : 0);
    rec.alloc(ctx, nrec * arr);
    uint8_t* p = (uint8_t*)rec.ptr;
    std::memset(&R, 0, sizeof(R));
    carve(p, &R.t, c);
    carve(p, &R.nx, c);
    carve(p, &R.ny, c);
    carve(p, &R.nz, c);
    carve(p, &R.u, c);
    carve(p, &R.v, c);
    carve(p, (float**)&R.mat, c);
    carve(p, &R.env_t, c);
    carve(p, &R.edx, c);
    carve(p, &R.edy, c);
    carve(p, &R.edz, c);
    if (tangents) {
        carve(p, &R.tgx, c);
        carve(p, &R.tgy, c);
        carve(p, &R.tgz, c);
    }
    if (dmis) {
        carve(p, &R.dl_t, c);
        carve(p, &R.dl_nx, c);
        carve(p, &R.dl_ny, c);
        carve(p, &R.dl_nz, c);
        carve(p, &R.dl_u, c);
        carve(p, &R.dl_v, c);
        carve(p, (float**)&R.dl_mat, c);
        carve(p, &R.lpx, c);
        carve(p, &R.lpy, c);
        carve(p, &R.lpz, c);
    }
}

void lx_enqueue_round(lx_context* ctx, LxJob* d_jobs, const int32_t* d_spp, uint32_t njobs, LxCounters* d_cnt,
                      LxSched* d_sched, uint32_t path_cap, WaveBufs& wb, float* d_accum, uint32_t* d_spill,
                      int max_depth, uint32_t trace_grid, uint32_t shade_grid, uint32_t cam_grid, int width,
                      const LxDevScene* uscene, hipEvent_t* ev, bool count, hipStream_t stream, bool dl_query,
                      bool colour_fresh) {
    hipStream_t s = stream ? stream : ctx->stream;
    int wf = lx_wavefront_bounces(ctx, max_depth);
    const int sm = ctx->tuning.sort_mode;
    const uint32_t tile = (uint32_t)ctx->tuning.pixel_tile, band = (uint32_t)ctx->tuning.tile_band;
    /* a tail kernel that starts at bounce 0 with the queue in slot order makes
     * the camera paths itself (the same rays, with no path-state round trip
     * through memory); otherwise the camera kernel writes them first */
    const bool tail_camera = lx_tail_makes_camera(ctx, max_depth);
    LX_HIP(lx_launch_begin_round(d_jobs, d_spp, njobs, d_cnt, d_sched, path_cap, s));
    if (!tail_camera) LX_HIP(lx_launch_camera(d_jobs, d_sched, d_cnt, wb.A, cam_grid, s, tile, band));
    LxPathBuf* cur = &wb.A;
    LxPathBuf* nxt = &wb.B;
    /* the sort stages write wb.order and the next launch reads it; a
     * per-bounce stage sits inside the timed bracket of the launch it serves */
    const uint32_t sort_grid = std::min(shade_grid, 2u * (uint32_t)std::max(1, ctx->num_cus));
    auto sort = [&](int bounce, int mode) -> const uint32_t* {
        LX_HIP(lx_launch_sort(d_jobs, *cur, wb.R, d_cnt, bounce, mode, wb.keys, wb.hist, wb.order, sort_grid, s));
        return wb.order;
    };
    for (int b = 0; b < wf; b++) {
        if (ev) LX_HIP(hipEventRecord(ev[2 * b], s));
        const uint32_t* ot = (sm & LX_SORT_RAYS) ? sort(b, LX_SORT_RAYS) : nullptr;
        LX_HIP(lx_launch_trace(d_jobs, *cur, wb.R, d_cnt, b, d_spill, trace_grid, width, s, count, ot));
        if (ev) LX_HIP(hipEventRecord(ev[2 * b + 1], s));
        const uint32_t* os = (sm & LX_SORT_MATERIALS) ? sort(b, LX_SORT_MATERIALS) : nullptr;
        LX_HIP(lx_launch_shade(d_jobs, *cur, *nxt, wb.R, d_cnt, b, d_accum, shade_grid, s, os));
        std::swap(cur, nxt);
    }
    if (wf < max_depth) {
        /* ev[2wf] -> ev[2wf+1] brackets the tail kernel alone (its queue sort
         * runs before the bracket) */
        const uint32_t* of = (sm & LX_SORT_RAYS) ? sort(wf, LX_SORT_RAYS) : nullptr;
        if (ev) LX_HIP(hipEventRecord(ev[2 * wf], s));
        LX_HIP(lx_launch_finish(d_jobs, *cur, d_cnt, wf, d_spill, d_accum, trace_grid * LX_TRACE_BLOCK,
                                ctx->tuning.block_size, ctx->tuning.tail_min_waves, width, ctx->tuning.queue_chunk,
                                uscene, njobs == 1, s, count, of, tail_camera ? d_sched : nullptr, tile, band,
                                dl_query, colour_fresh));
        if (ev && 2 * wf + 1 < 2 * max_depth) LX_HIP(hipEventRecord(ev[2 * wf + 1], s));
    }
    if (ev) LX_HIP(hipEventRecord(ev[2 * max_depth], s));
    LX_HIP(lx_launch_end_round(d_jobs, njobs, s));
}

void lx_counters_add(LxCounters& a, const LxCounters& b) {
    a.rad_rays += b.rad_rays;
    a.dl_rays += b.dl_rays;
    a.env_rays += b.env_rays;
    a.bounces += b.bounces;
    a.tail_bounces += b.tail_bounces;
    for (int w = 0; w < 2; w++) {
        a.work[w].closest_rays += b.work[w].closest_rays;
        a.work[w].any_rays += b.work[w].any_rays;
        a.work[w].closest_nodes += b.work[w].closest_nodes;
        a.work[w].closest_tris += b.work[w].closest_tris;
        a.work[w].closest_insts += b.work[w].closest_insts;
        a.work[w].any_nodes += b.work[w].any_nodes;
        a.work[w].any_tris += b.work[w].any_tris;
        a.work[w].any_insts += b.work[w].any_insts;
    }
    a.stack_overflow |= b.stack_overflow;
}

int lx_wavefront_bounces(const lx_context* ctx, int max_depth) {
    int wf = ctx->tuning.wavefront_bounces;
    if (wf < 0 || wf > max_depth) wf = max_depth;
    return wf;
}

bool lx_tail_makes_camera(const lx_context* ctx, int max_depth) {
    const int wf = lx_wavefront_bounces(ctx, max_depth);
    return wf == 0 && wf < max_depth && !(ctx->tuning.sort_mode & LX_SORT_RAYS);
}
