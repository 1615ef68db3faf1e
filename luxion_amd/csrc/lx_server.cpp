/*
 * lx_server.cpp -- served form: continuous batching of render jobs.
 *
 * A request thread owns a fixed table of job slots, a path capacity and an
 * accumulation pool sized from the HBM budget.  Jobs submitted from any
 * thread are admitted FIFO whenever a slot and enough accumulation space are
 * free; every round takes one sample of every admitted job in one shared
 * launch per bounce.  With overlapped rounds (the default) up to four rounds
 * are in flight on streams of their own, each adding its colours into the
 * jobs' pool ranges in round order; otherwise the round is a captured hipGraph
 * over the whole slot table (empty slots have a zero sample target and are
 * never scheduled).  Finished jobs leave between rounds and their results are
 * copied out, so jobs that arrive while others run overlap with them on the
 * device.
 * A job's pixels use only its own RNG streams, so its image does not depend
 * on which other jobs share its rounds.
 *
 * The reference has no batch or serve mode (run_best_scenes.ps1 runs one
 * process per job, serially).
 */
#include "lx_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {

struct SrvJob {
    uint64_t id = 0;
    lx_job job{};
    lx_job_result* res = nullptr;
    int state = 0; /* 0 queued, 1 active, 2 done */
    int slot = -1;
    uint32_t npix = 0, accum_off = 0;
    bool cancel = false;
    lx_status status = LX_OK;
    /* overlapped rounds */
    uint32_t launched = 0, merged = 0; /* samples launched / added into the pool */
    int64_t last_round = -1;           /* round of the last launched sample */
    bool stopped = false;              /* finished (or failed): no more samples */
};

/* First-fit allocator over pixel ranges of the accumulation pool. */
struct RangeAlloc {
    std::map<uint32_t, uint32_t> free_; /* offset -> length */
    void init(uint32_t n) {
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(uint32_t n, uint32_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                uint32_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint32_t off, uint32_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

} // namespace

struct lx_server {
    lx_context* ctx = nullptr;
    lx_server_opts opts{};
    uint32_t slots = 0, cap = 0, pool_pix = 0;
    int max_depth = LX_MAX_DEPTH;
    int width = 4; /* BVH width every job's scene must have */
    std::thread th;
    std::mutex m;
    std::condition_variable cv, done_cv;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<SrvJob*> queue;
    std::vector<SrvJob*> slot_job;
    std::map<uint64_t, SrvJob*> jobs;
    RangeAlloc ranges;
    uint32_t inject_fail = 0;
    /* device */
    DevBuf d_jobs, d_spp, d_cnt, d_sched, d_spill, d_accum;
    WaveBufs wb;
    std::vector<LxJob> h_jobs;    /* host mirror (async-copy source) */
    std::vector<int32_t> h_spp;
    uint32_t tg = 0, sg = 0, cg = 0;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    /* overlapped rounds (tuning.overlap_rounds at start): round r runs on
     * par[r % kDepth]'s stream with its own job table, counters, paths and
     * colour pool; only the sampled jobs' pixel ranges of the pool are filled
     * with -0.0 and then added into d_accum, round after round, on the
     * context's stream.  Up to kDepth rounds are in flight: a served round is
     * small and latency-bound (a few jobs' pixels), so more than two rounds
     * share the GPU at low load */
    static constexpr int kDepth = 4;
    bool overlap = false;
    struct Par {
        DevBuf jobs, spp, cnt, sched, spill, col, rng;
        WaveBufs wb;
        std::vector<LxJob> hj;      /* async-copy sources: untouched until this parity's next round */
        std::vector<int32_t> hs;
        std::vector<uint32_t> hr;
        std::vector<SrvJob*> in_round;
        uint32_t nr = 0;
        int64_t round = -1;
        double t_launch = 0.0;
        LxCounters last{};          /* counters as of this set's last processed round */
        hipStream_t s = nullptr;
        hipEvent_t done = nullptr, merged = nullptr;
    } par[kDepth];
    /* stats */
    std::vector<double> lat;
    lx_server_stats st{};
    uint64_t rays_prev = 0;

    void loop();
    void loop_overlapped();
    void release_slot(SrvJob* j) {
        ranges.release(j->accum_off, j->npix);
        slot_job[j->slot] = nullptr;
    }
};

static void finish_job(lx_server* S, SrvJob* j, lx_status status, double t) {
    j->state = 2;
    j->status = status;
    if (j->res) {
        j->res->status = status;
        j->res->finish_ms = t;
    }
    if (status == LX_OK) {
        S->lat.push_back(t - j->res->submit_ms);
        S->st.jobs_done++;
    } else if (status == LX_ERR_BUDGET) {
        S->st.jobs_rejected++;
    } else {
        S->st.jobs_failed++;
    }
}

void lx_server::loop() {
    while (true) {
        std::vector<int> to_upload;
        {
            std::unique_lock<std::mutex> lk(m);
            double t = now_ms();
            /* cancellations and deadlines of active jobs */
            for (uint32_t k = 0; k < slots; k++) {
                SrvJob* j = slot_job[k];
                if (!j) continue;
                bool expired = j->job.deadline_ms > 0 && t - j->res->submit_ms > j->job.deadline_ms;
                if (j->cancel || expired) {
                    finish_job(this, j, j->cancel ? LX_ERR_CANCELLED : LX_ERR_TIMEOUT, t);
                    ranges.release(j->accum_off, j->npix);
                    slot_job[k] = nullptr;
                    h_spp[k] = 0;
                    to_upload.push_back((int)k);
                }
            }
            /* admissions, FIFO */
            while (!queue.empty()) {
                SrvJob* j = queue.front();
                if (j->cancel) {
                    queue.pop_front();
                    finish_job(this, j, LX_ERR_CANCELLED, t);
                    continue;
                }
                if (inject_fail > 0 || j->npix > pool_pix || j->npix > cap) {
                    if (inject_fail > 0) inject_fail--;
                    queue.pop_front();
                    finish_job(this, j, LX_ERR_BUDGET, t);
                    continue;
                }
                int free_slot = -1;
                for (uint32_t k = 0; k < slots; k++)
                    if (!slot_job[k]) {
                        free_slot = (int)k;
                        break;
                    }
                uint32_t off = 0;
                if (free_slot < 0 || !ranges.alloc(j->npix, off)) break; /* wait for space */
                queue.pop_front();
                j->slot = free_slot;
                j->accum_off = off;
                j->state = 1;
                j->res->start_ms = t;
                slot_job[free_slot] = j;
                h_jobs[free_slot] = lx_make_dev_job(j->job, off);
                h_spp[free_slot] = j->job.spp;
                to_upload.push_back(free_slot);
            }
            bool any = false;
            uint32_t nact = 0;
            for (uint32_t k = 0; k < slots; k++)
                if (slot_job[k]) {
                    any = true;
                    nact++;
                }
            st.peak_jobs_in_flight = std::max(st.peak_jobs_in_flight, nact);
            if (!any && to_upload.empty()) {
                if (stop && queue.empty()) break;
                cv.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
        }
        /* device work for one round (the context mutex serialises the stream) */
        try {
            std::lock_guard<std::mutex> dl(ctx->mu);
            ctx->set_device();
            hipStream_t s = ctx->stream;
            for (int k : to_upload) {
                LX_HIP(hipMemcpyAsync(d_jobs.as<LxJob>() + k, &h_jobs[k], sizeof(LxJob), hipMemcpyHostToDevice, s));
                LX_HIP(hipMemcpyAsync(d_spp.as<int32_t>() + k, &h_spp[k], 4, hipMemcpyHostToDevice, s));
                SrvJob* j = slot_job[k];
                if (j)
                    LX_HIP(lx_launch_fill_u32(d_accum.as<uint32_t>() + 3 * (size_t)j->accum_off, 3 * (size_t)j->npix,
                                              0u, s));
            }
            bool any = false;
            for (uint32_t k = 0; k < slots; k++) any |= slot_job[k] != nullptr;
            if (any) {
                double t0 = now_ms();
                LX_HIP(hipGraphLaunch(ge, s));
                LX_HIP(hipStreamSynchronize(s));
                st.busy_ms += now_ms() - t0;
                st.rounds++;
                /* completions */
                std::vector<LxJob> cur(slots);
                LX_HIP(hipMemcpy(cur.data(), d_jobs.ptr, slots * sizeof(LxJob), hipMemcpyDeviceToHost));
                LxCounters c;
                LX_HIP(hipMemcpy(&c, d_cnt.ptr, sizeof(c), hipMemcpyDeviceToHost));
                LX_REQUIRE(c.stack_overflow == 0, LX_ERR_STATE, "traversal stack overflow");
                uint64_t rays = c.rad_rays + c.dl_rays + c.env_rays;
                std::lock_guard<std::mutex> lk(m);
                st.rays = rays;
                double t = now_ms();
                for (uint32_t k = 0; k < slots; k++) {
                    SrvJob* j = slot_job[k];
                    if (!j) continue;
                    j->res->spp_done = (uint32_t)cur[k].spp_done;
                    if (cur[k].spp_done >= j->job.spp) {
                        if (j->res->rgb_sum)
                            LX_HIP(hipMemcpy(j->res->rgb_sum, (uint8_t*)d_accum.ptr + (size_t)j->accum_off * 12,
                                             (size_t)j->npix * 12, hipMemcpyDeviceToHost));
                        j->res->rays = 0;
                        finish_job(this, j, LX_OK, t);
                        ranges.release(j->accum_off, j->npix);
                        slot_job[k] = nullptr;
                        h_spp[k] = 0;
                        LX_HIP(hipMemcpyAsync(d_spp.as<int32_t>() + k, &h_spp[k], 4, hipMemcpyHostToDevice, s));
                    }
                }
                LX_HIP(hipStreamSynchronize(s));
            }
        } catch (const LxError& e) {
            /* a device failure ends every in-flight job with its status */
            std::lock_guard<std::mutex> lk(m);
            double t = now_ms();
            for (uint32_t k = 0; k < slots; k++)
                if (slot_job[k]) {
                    finish_job(this, slot_job[k], e.status, t);
                    slot_job[k] = nullptr;
                }
            stop = true;
        }
        done_cv.notify_all();
    }
    done_cv.notify_all();
}

extern "C" {

/* The request loop with overlapped rounds: round r is launched while up to
 * kDepth - 1 earlier rounds (on the other streams) still run; once kDepth
 * rounds are in flight, or nothing is left to launch, the oldest is waited
 * for and its jobs' samples counted, results copied and slots released, so
 * the host work and the earlier rounds' last long paths overlap the newer
 * rounds.  A job's samples are added into its pool range in round order (one
 * sample per round), so its image is the bits of the loop without overlap. */
void lx_server::loop_overlapped() {
    int64_t r = 0;                /* next round */
    std::deque<int64_t> inflight; /* launched rounds not yet processed, oldest first */
    double t_done = 0.0;          /* when the last processed round was found complete */
    while (true) {
        std::vector<int> fresh; /* slots admitted now: their pool range starts at zero */
        bool work = false;
        {
            std::unique_lock<std::mutex> lk(m);
            double t = now_ms();
            const int64_t oldest = inflight.empty() ? r : inflight.front(); /* first round not yet processed */
            /* cancellations and deadlines of active jobs: no more samples; the slot
             * is free once no launched round holds the job */
            for (uint32_t k = 0; k < slots; k++) {
                SrvJob* j = slot_job[k];
                if (!j || j->stopped) continue;
                bool expired = j->job.deadline_ms > 0 && t - j->res->submit_ms > j->job.deadline_ms;
                if (j->cancel || expired) {
                    finish_job(this, j, j->cancel ? LX_ERR_CANCELLED : LX_ERR_TIMEOUT, t);
                    j->stopped = true;
                    if (j->last_round < oldest) release_slot(j);
                }
            }
            /* admissions, FIFO */
            while (!queue.empty()) {
                SrvJob* j = queue.front();
                if (j->cancel) {
                    queue.pop_front();
                    finish_job(this, j, LX_ERR_CANCELLED, t);
                    continue;
                }
                if (inject_fail > 0 || j->npix > pool_pix || j->npix > cap) {
                    if (inject_fail > 0) inject_fail--;
                    queue.pop_front();
                    finish_job(this, j, LX_ERR_BUDGET, t);
                    continue;
                }
                int free_slot = -1;
                for (uint32_t k = 0; k < slots; k++)
                    if (!slot_job[k]) {
                        free_slot = (int)k;
                        break;
                    }
                uint32_t off = 0;
                if (free_slot < 0 || !ranges.alloc(j->npix, off)) break; /* wait for space */
                queue.pop_front();
                j->slot = free_slot;
                j->accum_off = off;
                j->state = 1;
                j->res->start_ms = t;
                j->launched = j->merged = 0;
                j->last_round = -1;
                slot_job[free_slot] = j;
                fresh.push_back(free_slot);
            }
            uint32_t nact = 0;
            for (uint32_t k = 0; k < slots; k++) {
                SrvJob* j = slot_job[k];
                if (!j) continue;
                nact++;
                if (!j->stopped && j->launched < (uint32_t)j->job.spp) work = true;
            }
            st.peak_jobs_in_flight = std::max(st.peak_jobs_in_flight, nact);
            if (!work && inflight.empty()) {
                if (stop && queue.empty()) break;
                cv.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
        }
        try {
            std::lock_guard<std::mutex> dl(ctx->mu);
            ctx->set_device();
            hipStream_t ms = ctx->stream;
            bool launched = false;
            if (work && inflight.size() < (size_t)kDepth) {
                /* par[r % kDepth] is free: its previous round (r - kDepth) was processed */
                Par& P = par[r % kDepth];
                P.t_launch = now_ms();
                for (int k : fresh) {
                    SrvJob* j = slot_job[k];
                    LX_HIP(lx_launch_fill_u32(d_accum.as<uint32_t>() + 3 * (size_t)j->accum_off, 3 * (size_t)j->npix,
                                              0u, ms));
                }
                P.in_round.clear();
                P.nr = 0;
                /* the round's jobs, chosen here so that they fit the path capacity
                 * together (the round then schedules every one of them); the first
                 * slot rotates so that no job waits behind the same others */
                uint64_t used = 0;
                for (uint32_t q = 0; q < slots; q++) {
                    uint32_t k = (uint32_t)(((uint64_t)r + q) % slots);
                    SrvJob* j = slot_job[k];
                    if (j && !j->stopped && j->launched < (uint32_t)j->job.spp && used + j->npix <= cap) {
                        used += j->npix;
                        P.hj[k] = lx_make_dev_job(j->job, j->accum_off);
                        P.hj[k].iter = (int32_t)j->job.first_iter - 1 + (int32_t)j->launched; /* +1 at the round's start */
                        P.hs[k] = 1;
                        P.hr[2 * P.nr] = j->accum_off;
                        P.hr[2 * P.nr + 1] = j->npix;
                        P.nr++;
                        P.in_round.push_back(j);
                        j->launched++;
                        j->last_round = r;
                    } else {
                        P.hs[k] = 0; /* not scheduled this round */
                    }
                }
                LX_HIP(hipMemcpyAsync(P.jobs.ptr, P.hj.data(), slots * sizeof(LxJob), hipMemcpyHostToDevice, P.s));
                LX_HIP(hipMemcpyAsync(P.spp.ptr, P.hs.data(), slots * 4, hipMemcpyHostToDevice, P.s));
                LX_HIP(hipMemcpyAsync(P.rng.ptr, P.hr.data(), (size_t)P.nr * 8, hipMemcpyHostToDevice, P.s));
                LX_HIP(lx_launch_fill_ranges(P.col.as<uint32_t>(), P.rng.as<uint32_t>(), P.nr, 0x80000000u, P.s));
                lx_enqueue_round(ctx, P.jobs.as<LxJob>(), P.spp.as<int32_t>(), slots, P.cnt.as<LxCounters>(),
                                 P.sched.as<LxSched>(), cap, P.wb, P.col.as<float>(), P.spill.as<uint32_t>(),
                                 max_depth, tg, sg, cg, width, nullptr, nullptr, false, P.s, true, true);
                LX_HIP(hipEventRecord(P.done, P.s));
                LX_HIP(hipStreamWaitEvent(ms, P.done, 0));
                LX_HIP(lx_launch_add_ranges(d_accum.as<float>(), P.col.as<float>(), P.rng.as<uint32_t>(), P.nr, ms));
                LX_HIP(hipEventRecord(P.merged, ms));
                P.round = r;
                inflight.push_back(r);
                r++;
                launched = true;
            }
            if (!inflight.empty() && (inflight.size() >= (size_t)kDepth || !launched)) {
                Par& Q = par[inflight.front() % kDepth];
                LX_HIP(hipEventSynchronize(Q.merged));
                LX_HIP(hipMemcpy(&Q.last, Q.cnt.ptr, sizeof(LxCounters), hipMemcpyDeviceToHost));
                LX_REQUIRE(Q.last.stack_overflow == 0, LX_ERR_STATE, "traversal stack overflow");
                std::lock_guard<std::mutex> lk(m);
                st.rays = 0;
                for (const Par& P : par) st.rays += P.last.rad_rays + P.last.dl_rays + P.last.env_rays;
                st.rounds++;
                double t = now_ms();
                st.busy_ms += t - std::max(Q.t_launch, t_done); /* device time not counted twice */
                t_done = t;
                for (SrvJob* j : Q.in_round) {
                    j->merged++;
                    if (!j->stopped) j->res->spp_done = j->merged;
                    if (!j->stopped && j->merged == (uint32_t)j->job.spp) {
                        /* every sample of the job is in its range (later rounds do not touch it) */
                        if (j->res->rgb_sum)
                            LX_HIP(hipMemcpy(j->res->rgb_sum, (uint8_t*)d_accum.ptr + (size_t)j->accum_off * 12,
                                             (size_t)j->npix * 12, hipMemcpyDeviceToHost));
                        j->res->rays = 0;
                        finish_job(this, j, LX_OK, t);
                        j->stopped = true;
                        release_slot(j);
                    } else if (j->stopped && j->last_round == Q.round) {
                        release_slot(j); /* cancelled or expired while this round ran */
                    }
                }
                Q.in_round.clear();
                inflight.pop_front();
            }
        } catch (const LxError& e) {
            /* a device failure ends every in-flight job with its status */
            (void)hipDeviceSynchronize();
            std::lock_guard<std::mutex> lk(m);
            double t = now_ms();
            for (uint32_t k = 0; k < slots; k++)
                if (SrvJob* j = slot_job[k]) {
                    if (!j->stopped) finish_job(this, j, e.status, t);
                    slot_job[k] = nullptr;
                }
            inflight.clear();
            stop = true;
        }
        done_cv.notify_all();
    }
    done_cv.notify_all();
}

lx_status lx_server_start(lx_context* ctx, const lx_server_opts* o, lx_server** out) {
    try {
        LX_REQUIRE(ctx && out, LX_ERR_INVALID, "null argument");
        *out = nullptr;
        lx_server_opts opts{0, 64, 1, 0, 16};
        if (o) opts = *o;
        if (opts.max_depth <= 0) opts.max_depth = 16;
        LX_REQUIRE(opts.max_depth <= LX_MAX_DEPTH, LX_ERR_INVALID, "max_depth above the supported maximum");
        LX_REQUIRE(opts.max_jobs_in_flight >= 1 && opts.max_jobs_in_flight <= LX_SCHED_MAX, LX_ERR_INVALID,
                   "max_jobs_in_flight out of range");
        LX_REQUIRE(opts.spp_per_round <= 1, LX_ERR_UNSUPPORTED,
                   "one sample per job per round (per-pixel accumulation is race-free that way)");
        std::unique_ptr<lx_server> S(new lx_server());
        S->ctx = ctx;
        S->opts = opts;
        S->slots = (uint32_t)opts.max_jobs_in_flight;
        S->max_depth = opts.max_depth;
        std::lock_guard<std::mutex> dl(ctx->mu);
        S->width = ctx->tuning.bvh_width == 8 ? 8 : 4;
        S->overlap = ctx->tuning.overlap_rounds != 0;
        ctx->set_device();
        uint64_t avail = ctx->budget > ctx->in_use.load() ? ctx->budget - ctx->in_use.load() : 0;
        uint64_t budget = opts.hbm_budget_bytes ? std::min<uint64_t>(opts.hbm_budget_bytes, avail) : avail;
        /* split the budget: path capacity and accumulation pool (overlapped
         * rounds: a set of path buffers and a colour pool per round in flight,
         * besides the accumulation pool) */
        const uint64_t sets = S->overlap ? lx_server::kDepth : 1, pools = S->overlap ? lx_server::kDepth + 1 : 1;
        uint64_t per_path = 2 * 15 * 4 + 24 * 4;
        uint64_t cap = opts.max_paths_per_round ? opts.max_paths_per_round
                                                : std::min<uint64_t>(budget / 2 / per_path / sets, 16u << 20);
        cap = std::max<uint64_t>(cap, 4096);
        uint64_t pool = (budget - std::min(budget, cap * per_path * sets)) / 12 / 2 / pools;
        pool = std::min<uint64_t>(pool, (1ull << 31) - 1);
        LX_REQUIRE(pool >= 4096, LX_ERR_BUDGET, "HBM budget too small for the served form");
        S->cap = (uint32_t)cap;
        S->pool_pix = (uint32_t)pool;
        S->ranges.init(S->pool_pix);
        S->slot_job.assign(S->slots, nullptr);
        S->h_jobs.assign(S->slots, LxJob{});
        S->h_spp.assign(S->slots, 0);
        S->wb.alloc(ctx, S->cap, true, true, true);
        S->d_accum.alloc(ctx, (size_t)S->pool_pix * 12);
        S->d_jobs.upload(ctx, S->h_jobs.data(), S->slots * sizeof(LxJob));
        S->d_spp.upload(ctx, S->h_spp.data(), S->slots * 4);
        S->d_cnt.alloc(ctx, sizeof(LxCounters));
        S->d_sched.alloc(ctx, sizeof(LxSched));
        uint32_t waves = (uint32_t)std::max(1, ctx->tuning.waves_per_cu);
        S->tg = std::min((uint32_t)ctx->num_cus * std::max(1u, waves * 64u / LX_TRACE_BLOCK),
                         std::max(1u, (S->cap + LX_TRACE_BLOCK - 1) / LX_TRACE_BLOCK));
        S->sg = std::min((uint32_t)ctx->num_cus * 8u, std::max(1u, (S->cap + LX_SHADE_BLOCK - 1) / LX_SHADE_BLOCK));
        S->cg = std::min((uint32_t)ctx->num_cus * 8u, std::max(1u, (S->cap + 255u) / 256u));
        S->d_spill.alloc(ctx, (size_t)lx_spill_words(S->tg) * 4);
        hipStream_t s = ctx->stream;
        LX_HIP(lx_launch_fill_u32(S->d_cnt.as<uint32_t>(), sizeof(LxCounters) / 4, 0u, s));
        if (S->overlap) {
            for (auto& P : S->par) {
                P.wb.alloc(ctx, S->cap, true, true, true);
                P.jobs.alloc(ctx, S->slots * sizeof(LxJob));
                P.spp.alloc(ctx, S->slots * 4);
                P.cnt.alloc(ctx, sizeof(LxCounters));
                P.sched.alloc(ctx, sizeof(LxSched));
                P.spill.alloc(ctx, S->d_spill.bytes);
                P.col.alloc(ctx, (size_t)S->pool_pix * 12);
                P.rng.alloc(ctx, (size_t)S->slots * 8);
                P.hj.assign(S->slots, LxJob{});
                P.hs.assign(S->slots, 0);
                P.hr.assign(2 * (size_t)S->slots, 0u);
                LX_HIP(lx_launch_fill_u32(P.cnt.as<uint32_t>(), sizeof(LxCounters) / 4, 0u, s));
                LX_HIP(hipStreamCreateWithFlags(&P.s, hipStreamNonBlocking));
                LX_HIP(hipEventCreateWithFlags(&P.done, hipEventDisableTiming));
                LX_HIP(hipEventCreateWithFlags(&P.merged, hipEventDisableTiming));
            }
            LX_HIP(hipStreamSynchronize(s));
            ctx->refs++;
            lx_server* raw = S.release();
            raw->th = std::thread([raw]() { raw->loop_overlapped(); });
            *out = raw;
            lx_set_error("");
            return LX_OK;
        }
        LX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            lx_enqueue_round(ctx, S->d_jobs.as<LxJob>(), S->d_spp.as<int32_t>(), S->slots, S->d_cnt.as<LxCounters>(),
                             S->d_sched.as<LxSched>(), S->cap, S->wb, S->d_accum.as<float>(),
                             S->d_spill.as<uint32_t>(), S->max_depth, S->tg, S->sg, S->cg, S->width, nullptr,
                             nullptr);
        } catch (...) {
            hipGraph_t tmp;
            (void)hipStreamEndCapture(s, &tmp);
            throw;
        }
        LX_HIP(hipStreamEndCapture(s, &S->g));
        LX_HIP(hipGraphInstantiate(&S->ge, S->g, nullptr, nullptr, 0));
        LX_HIP(hipStreamSynchronize(s));
        ctx->refs++;
        lx_server* raw = S.release();
        raw->th = std::thread([raw]() { raw->loop(); });
        *out = raw;
        lx_set_error("");
        return LX_OK;
    } catch (const LxError& e) {
        lx_set_error(e.msg);
        return e.status;
    }
}

lx_status lx_server_submit(lx_server* S, const lx_job* job, lx_job_result* res, uint64_t* id) {
    try {
        LX_REQUIRE(S && job && res && id, LX_ERR_INVALID, "null argument");
        lx_validate_job(*job);
        int depth = lx_job_depth(job->max_depth, job->scene->trace_depth, job->material_debug_mode);
        LX_REQUIRE(depth <= S->max_depth, LX_ERR_INVALID, "job depth above the server's maximum");
        LX_REQUIRE(job->scene->dev.bvh_width == S->width, LX_ERR_INVALID,
                   "job scene BVH width differs from the server's (the context tuning's bvh_width at start)");
        SrvJob* j = new SrvJob();
        j->job = *job;
        j->res = res;
        j->npix = (uint32_t)job->camera.width * (uint32_t)job->camera.height;
        res->status = LX_ERR_STATE;
        res->spp_done = 0;
        res->rays = 0;
        res->submit_ms = now_ms();
        res->start_ms = res->finish_ms = 0;
        {
            std::lock_guard<std::mutex> lk(S->m);
            LX_REQUIRE(!S->stop, LX_ERR_STATE, "server is stopping");
            job->scene->refs++;
            j->id = S->next_id++;
            S->jobs[j->id] = j;
            S->queue.push_back(j);
            *id = j->id;
        }
        S->cv.notify_all();
        lx_set_error("");
        return LX_OK;
    } catch (const LxError& e) {
        lx_set_error(e.msg);
        return e.status;
    }
}

lx_status lx_server_poll(lx_server* S, uint64_t id, double timeout_ms, int32_t* done) {
    if (!S || !done) return LX_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) {
        lx_set_error("unknown job id");
        return LX_ERR_NOT_FOUND;
    }
    SrvJob* j = it->second;
    if (j->state != 2 && timeout_ms != 0) {
        auto pred = [&]() { return j->state == 2; };
        if (timeout_ms < 0) S->done_cv.wait(lk, pred);
        else S->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred);
    }
    *done = j->state == 2 ? 1 : 0;
    return LX_OK;
}

lx_status lx_server_cancel(lx_server* S, uint64_t id) {
    if (!S) return LX_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->m);
        auto it = S->jobs.find(id);
        if (it == S->jobs.end()) {
            lx_set_error("unknown job id");
            return LX_ERR_NOT_FOUND;
        }
        it->second->cancel = true;
    }
    S->cv.notify_all();
    return LX_OK;
}

lx_status lx_server_stats_get(lx_server* S, lx_server_stats* out) {
    if (!S || !out) return LX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->m);
    *out = S->st;
    if (!S->lat.empty()) {
        std::vector<double> v = S->lat;
        std::sort(v.begin(), v.end());
        auto q = [&](double p) {
            double x = p * (v.size() - 1);
            size_t a = (size_t)x;
            size_t b = std::min(a + 1, v.size() - 1);
            return v[a] + (x - a) * (v[b] - v[a]);
        };
        out->p50_ms = q(0.50);
        out->p99_ms = q(0.99);
    }
    out->peak_bytes = S->ctx->in_use.load();
    return LX_OK;
}

lx_status lx_server_inject_budget_failure(lx_server* S, uint32_t n) {
    if (!S) return LX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->m);
    S->inject_fail += n;
    return LX_OK;
}

lx_status lx_server_stop(lx_server* S) {
    if (!S) return LX_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->th.joinable()) S->th.join();
    lx_context* ctx = S->ctx;
    {
        std::lock_guard<std::mutex> dl(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (S->ge) (void)hipGraphExecDestroy(S->ge);
        if (S->g) (void)hipGraphDestroy(S->g);
        for (auto& P : S->par) {
            if (P.s) {
                (void)hipStreamSynchronize(P.s);
                (void)hipStreamDestroy(P.s);
            }
            if (P.done) (void)hipEventDestroy(P.done);
            if (P.merged) (void)hipEventDestroy(P.merged);
        }
    }
    for (auto& kv : S->jobs) {
        lx_scene_release(kv.second->job.scene);
        delete kv.second;
    }
    delete S;
    lx_context_release(ctx);
    return LX_OK;
}

} /* extern "C" */
