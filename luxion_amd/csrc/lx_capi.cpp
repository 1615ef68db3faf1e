/*
 * lx_capi.cpp -- extern "C" entry points declared in include/luxion_mi355x.h.
 * Every call converts exceptions into an lx_status and a per-thread message;
 * nothing terminates the process (the reference exits on every error,
 * include/common.h:16-36).
 */
#include "lx_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

static thread_local std::string g_err;

void lx_set_error(const std::string& msg) { g_err = msg; }

/* ------------------------------------------------------------- context */
void lx_context::set_device() const { LX_HIP(hipSetDevice(device)); }

void* lx_context::dalloc(size_t bytes) {
    uint64_t cur = in_use.load();
    if (budget && cur + bytes > budget)
        throw LxError{LX_ERR_BUDGET, "HBM budget exceeded: " + std::to_string(cur + bytes) + " > " +
                                         std::to_string(budget) + " bytes"};
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw LxError{LX_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e)};
    in_use += bytes;
    return p;
}

void lx_context::dfree(void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    in_use -= bytes;
}

extern "C" {

const char* lx_last_error(void) { return g_err.c_str(); }
int lx_abi_version(void) { return LX_ABI_VERSION; }

double lx_now_ms(void) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

lx_status lx_context_create(int device, uint64_t budget, lx_context** out) {
    LX_TRY
    LX_REQUIRE(out, LX_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int n = 0;
    LX_HIP(hipGetDeviceCount(&n));
    LX_REQUIRE(device >= 0 && device < n, LX_ERR_INVALID, "device ordinal out of range");
    lx_context* c = new lx_context();
    c->device = device;
    try {
        c->set_device();
        LX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        LX_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        LX_HIP(lx_noalloc_scratch_address(device, &c->noalloc_scratch));
        size_t fr = 0, tot = 0;
        LX_HIP(hipMemGetInfo(&fr, &tot));
        c->budget = budget ? budget : (uint64_t)fr;
        c->tuning.block_size = 256;
        c->tuning.bvh_width = 4;
        c->tuning.queue_chunk = 32;
        c->tuning.waves_per_cu = 16;
        c->tuning.use_graph = 1;
        c->tuning.wavefront_bounces = LX_DEFAULT_WAVEFRONT_BOUNCES;
        c->tuning.tail_min_waves = 4;
        c->tuning.sort_mode = LX_SORT_NONE;
        c->tuning.pixel_tile = 8;
        c->tuning.tile_band = 0;
        c->tuning.overlap_rounds = 1;
        c->tuning.round_frames = 1;
    } catch (...) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        throw;
    }
    *out = c;
    LX_CATCH
}

/* Objects are reference counted (render -> scene -> context) so handles can
 * be destroyed in any order; the last release frees. */
void lx_context_release(lx_context* c) {
    if (!c || --c->refs > 0) return;
    (void)hipSetDevice(c->device);
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

void lx_scene_release(lx_scene* s) {
    if (!s || --s->refs > 0) return;
    lx_context* c = s->ctx;
    if (c) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
    }
    delete s; /* device buffers return to the context here */
    lx_context_release(c);
}

void lx_context_destroy(lx_context* c) { lx_context_release(c); }

uint64_t lx_context_bytes_in_use(const lx_context* c) { return c ? c->in_use.load() : 0; }

/* The tuning is read under the context mutex by every launch sequence (a
 * server's request thread included), so reads and writes take it too. */
lx_status lx_tuning_get(const lx_context* c, lx_tuning* out) {
    LX_TRY
    LX_REQUIRE(c && out, LX_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(const_cast<lx_context*>(c)->mu);
    *out = c->tuning;
    LX_CATCH
}

lx_status lx_tuning_set(lx_context* c, const lx_tuning* t) {
    LX_TRY
    LX_REQUIRE(c && t, LX_ERR_INVALID, "null argument");
    lx_tuning_check(*t);
    std::lock_guard<std::mutex> lk(c->mu);
    c->tuning = *t;
    LX_CATCH
}

/* --------------------------------------------------------------- scene */
lx_status lx_scene_create(lx_context* ctx, const lx_scene_desc* desc, lx_scene** out) {
    LX_TRY
    LX_REQUIRE(ctx && desc && out, LX_ERR_INVALID, "null argument");
    *out = nullptr;
    lx_scene* s = new lx_scene();
    s->ctx = ctx;
    try {
        lx_scene_init(s, desc);
    } catch (...) {
        delete s;
        throw;
    }
    ctx->refs++;
    *out = s;
    LX_CATCH
}

lx_status lx_scene_build_accel(lx_scene* s, const lx_build_opts* o) {
    LX_TRY
    LX_REQUIRE(s, LX_ERR_INVALID, "null scene");
    /* a rebuild replaces the device buffers that renders and jobs point at */
    LX_REQUIRE(!s->accel_built || s->refs.load() == 1, LX_ERR_STATE,
               "scene is in use by renders or jobs: destroy them before rebuilding");
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    lx_scene_accel(s, o);
    LX_CATCH
}

lx_status lx_scene_get_camera(const lx_scene* s, lx_camera_desc* out, int32_t* depth) {
    LX_TRY
    LX_REQUIRE(s, LX_ERR_INVALID, "null scene");
    if (out) *out = s->cam;
    if (depth) *depth = s->trace_depth;
    LX_CATCH
}

lx_status lx_scene_get_info(const lx_scene* s, lx_scene_info* out) {
    LX_TRY
    LX_REQUIRE(s && out, LX_ERR_INVALID, "null argument");
    *out = s->info;
    LX_CATCH
}

lx_status lx_scene_read_env_cdfs(const lx_scene* s, float* marginal, float* conditional) {
    LX_TRY
    LX_REQUIRE(s && s->has_env, LX_ERR_STATE, "scene has no environment map");
    LX_REQUIRE(s->accel_built, LX_ERR_STATE, "call lx_scene_build_accel first");
    if (marginal) std::memcpy(marginal, s->marginal.data(), s->marginal.size() * 4);
    if (conditional) std::memcpy(conditional, s->conditional.data(), s->conditional.size() * 4);
    LX_CATCH
}

lx_status lx_scene_read_emitter_cdfs(const lx_scene* s, uint32_t* inst, float* prefix, uint32_t* num) {
    LX_TRY
    LX_REQUIRE(s && s->accel_built, LX_ERR_STATE, "call lx_scene_build_accel first");
    if (num) *num = (uint32_t)s->emissive.size();
    if (inst) std::memcpy(inst, s->emissive.data(), s->emissive.size() * 4);
    if (prefix) std::memcpy(prefix, s->emissive_prefix.data(), s->emissive_prefix.size() * 4);
    LX_CATCH
}

lx_status lx_scene_read_instance_tri_cdf(const lx_scene* s, uint32_t k, float* tri_prefix, uint32_t* n) {
    LX_TRY
    LX_REQUIRE(s && s->accel_built, LX_ERR_STATE, "call lx_scene_build_accel first");
    LX_REQUIRE(k < s->emissive.size(), LX_ERR_NOT_FOUND, "emissive index out of range");
    uint32_t a = s->tri_cdf_off[k], b = s->tri_cdf_off[k + 1];
    if (n) *n = b - a;
    if (tri_prefix) std::memcpy(tri_prefix, s->tri_cdf.data() + a, (size_t)(b - a) * 4);
    LX_CATCH
}

void lx_scene_destroy(lx_scene* s) { lx_scene_release(s); }

/* ------------------------------------------------------------ test hooks */
static lx_status trace_hook(lx_scene* s, const lx_ray* rays, uint32_t n, int any, int alpha, uint32_t iteration,
                            lx_hit* hits, uint8_t* occ) {
    LX_TRY
    LX_REQUIRE(s && s->accel_built, LX_ERR_STATE, "call lx_scene_build_accel first");
    LX_REQUIRE(rays && (any == 1 ? occ != nullptr : hits != nullptr), LX_ERR_INVALID, "null argument");
    if (n == 0) return LX_OK;
    lx_context* ctx = s->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    uint32_t grid = std::max(1u, std::min((n + LX_TRACE_BLOCK - 1) / LX_TRACE_BLOCK, (uint32_t)ctx->num_cus * 8u));
    DevBuf dr, dh, dsp, dov;
    dr.upload(ctx, rays, (size_t)n * sizeof(lx_ray));
    dh.alloc(ctx, any == 1 ? (size_t)n : (size_t)n * sizeof(lx_hit));
    dsp.alloc(ctx, (size_t)lx_spill_words(grid) * 4);
    dov.alloc(ctx, 256);
    LX_HIP(lx_launch_fill_u32(dov.as<uint32_t>(), 64, 0u, ctx->stream));
    LX_HIP(lx_launch_trace_rays(s->d_scene.as<LxDevScene>(), dr.as<lx_ray>(), n, any, alpha, iteration,
                                any == 1 ? nullptr : dh.as<lx_hit>(), any == 1 ? dh.as<uint8_t>() : nullptr,
                                dsp.as<uint32_t>(), dov.as<uint32_t>(), grid, s->dev.bvh_width, ctx->stream));
    uint32_t ov = 0;
    if (any == 1) LX_HIP(hipMemcpyAsync(occ, dh.ptr, n, hipMemcpyDeviceToHost, ctx->stream));
    else LX_HIP(hipMemcpyAsync(hits, dh.ptr, (size_t)n * sizeof(lx_hit), hipMemcpyDeviceToHost, ctx->stream));
    LX_HIP(hipMemcpyAsync(&ov, dov.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    LX_HIP(hipStreamSynchronize(ctx->stream));
    LX_REQUIRE(ov == 0, LX_ERR_STATE, "traversal stack overflow (BVH deeper than the supported stack)");
    LX_CATCH
}

lx_status lx_trace_closest(lx_scene* s, const lx_ray* rays, uint32_t n, int32_t alpha, uint32_t iteration,
                           lx_hit* hits) {
    return trace_hook(s, rays, n, 0, alpha, iteration, hits, nullptr);
}

lx_status lx_trace_occluded(lx_scene* s, const lx_ray* rays, uint32_t n, int32_t alpha, uint32_t iteration,
                            uint8_t* occ) {
    return trace_hook(s, rays, n, 1, alpha, iteration, nullptr, occ);
}

lx_status lx_trace_emitting(lx_scene* s, const lx_ray* rays, uint32_t n, uint32_t iteration, lx_hit* hits) {
    return trace_hook(s, rays, n, 2, 1, iteration, hits, nullptr);
}

/* The allocation-free variant of lx_trace_closest: device-resident rays and
 * hits, the spill columns and overflow flag in this device's scratch array
 * (one lane per ray, so n is capped by the array).  Calls that overlap on one
 * device race on the array; a context serialises its own calls. */
static void require_device_range(const lx_context* ctx, const void* p, size_t bytes, const char* what) {
    hipPointerAttribute_t at{};
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess || at.type != hipMemoryTypeDevice || at.device != ctx->device) {
        (void)hipGetLastError();
        throw LxError{LX_ERR_INVALID, std::string(what) + " is not a device allocation of this process on the "
                                                          "scene's device"};
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw LxError{LX_ERR_INVALID, std::string(what) + " is not a device allocation of this process on the "
                                                          "scene's device"};
    }
    uintptr_t b = (uintptr_t)base, q = (uintptr_t)p;
    LX_REQUIRE(q >= b && q - b <= size && size - (q - b) >= bytes, LX_ERR_INVALID,
               std::string(what) + " extends past the end of its device allocation");
}

lx_status lx_trace_closest_noalloc(lx_scene* s, const lx_ray* d_rays, uint32_t n, int32_t alpha, uint32_t iteration,
                                   lx_hit* d_hits) {
    LX_TRY
    LX_REQUIRE(s && s->accel_built, LX_ERR_STATE, "call lx_scene_build_accel first");
    LX_REQUIRE(d_rays && d_hits, LX_ERR_INVALID, "null argument");
    LX_REQUIRE(n <= LX_NOALLOC_MAX_RAYS, LX_ERR_INVALID,
               "lx_trace_closest_noalloc takes at most " + std::to_string(LX_NOALLOC_MAX_RAYS) +
                   " rays per call (got " + std::to_string(n) + "); split the batch");
    if (n == 0) return LX_OK;
    lx_context* ctx = s->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    LX_REQUIRE(ctx->noalloc_scratch, LX_ERR_STATE, "scratch array address unavailable on this device");
    require_device_range(ctx, d_rays, (size_t)n * sizeof(lx_ray), "d_rays");
    require_device_range(ctx, d_hits, (size_t)n * sizeof(lx_hit), "d_hits");
    uint32_t grid = (n + LX_TRACE_BLOCK - 1) / LX_TRACE_BLOCK; /* <= LX_NOALLOC_MAX_RAYS / LX_TRACE_BLOCK */
    uint32_t* spill = ctx->noalloc_scratch;
    uint32_t* overflow = ctx->noalloc_scratch + LX_NOALLOC_SPILL_WORDS;
    LX_HIP(lx_launch_fill_u32(overflow, 64, 0u, ctx->stream));
    LX_HIP(lx_launch_trace_rays(s->d_scene.as<LxDevScene>(), d_rays, n, 0, alpha, iteration, d_hits, nullptr, spill,
                                overflow, grid, s->dev.bvh_width, ctx->stream));
    uint32_t ov = 0;
    LX_HIP(hipMemcpyAsync(&ov, overflow, 4, hipMemcpyDeviceToHost, ctx->stream));
    LX_HIP(hipStreamSynchronize(ctx->stream));
    LX_REQUIRE(ov == 0, LX_ERR_STATE, "traversal stack overflow (BVH deeper than the supported stack)");
    LX_CATCH
}

lx_status lx_debug_math_eval(lx_context* ctx, int32_t fn, const float* x, const float* y, float* out, uint32_t n) {
    LX_TRY
    LX_REQUIRE(ctx && x && out, LX_ERR_INVALID, "null argument");
    if (n == 0) return LX_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    DevBuf dx, dy, dout;
    dx.upload(ctx, x, (size_t)n * 4);
    if (y) dy.upload(ctx, y, (size_t)n * 4);
    dout.alloc(ctx, (size_t)n * 4);
    LX_HIP(lx_launch_math_eval(fn, dx.as<float>(), y ? dy.as<float>() : nullptr, dout.as<float>(), n, ctx->stream));
    LX_HIP(hipMemcpyAsync(out, dout.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    LX_HIP(hipStreamSynchronize(ctx->stream));
    LX_CATCH
}

lx_status lx_debug_copy_bandwidth(lx_context* ctx, uint64_t bytes, uint32_t iters, double* gbps) {
    LX_TRY
    LX_REQUIRE(ctx && gbps && bytes >= 16 && iters > 0, LX_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    size_t n16 = (size_t)(bytes / 16);
    DevBuf a, b;
    a.alloc(ctx, n16 * 16);
    b.alloc(ctx, n16 * 16);
    LX_HIP(lx_launch_fill_u32(a.as<uint32_t>(), n16 * 4, 0x3f800000u, ctx->stream));
    uint32_t grid = (uint32_t)ctx->num_cus * 16u;
    LX_HIP(lx_launch_copy16(a.ptr, b.ptr, n16, grid, ctx->stream)); /* warm */
    hipEvent_t e0, e1;
    LX_HIP(hipEventCreate(&e0));
    LX_HIP(hipEventCreate(&e1));
    LX_HIP(hipEventRecord(e0, ctx->stream));
    for (uint32_t k = 0; k < iters; k++) LX_HIP(lx_launch_copy16(a.ptr, b.ptr, n16, grid, ctx->stream));
    LX_HIP(hipEventRecord(e1, ctx->stream));
    LX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbps = 2.0 * (double)n16 * 16.0 * iters / (ms * 1e-3) / 1e9;
    LX_CATCH
}

/* -------------------------------------------------------------- render */
static void render_grids(lx_context* ctx, uint32_t cap, uint32_t& tg, uint32_t& sg, uint32_t& cg) {
    uint32_t waves = (uint32_t)std::max(1, ctx->tuning.waves_per_cu);
    tg = (uint32_t)ctx->num_cus * std::max(1u, waves * 64u / LX_TRACE_BLOCK);
    tg = std::min(tg, std::max(1u, (cap + LX_TRACE_BLOCK - 1) / LX_TRACE_BLOCK));
    sg = std::min((uint32_t)ctx->num_cus * 8u, std::max(1u, (cap + LX_SHADE_BLOCK - 1) / LX_SHADE_BLOCK));
    cg = std::min((uint32_t)ctx->num_cus * 8u, std::max(1u, (cap + 255u) / 256u));
}

static void validate_flags(const lx_scene* s, int dmis, int emis) {
    if (dmis)
        LX_REQUIRE(!s->emissive.empty() && !(s->total_area < 1e-5f), LX_ERR_INVALID,
                   "-dmis requires the scene to contain emissive geometry");
    if (emis) LX_REQUIRE(s->has_env, LX_ERR_INVALID, "-emis requires an environment map, pass one with -e");
}

static bool scene_has_normal_maps(const lx_scene* s) {
    for (auto& m : s->mats)
        if (m.normal_tex >= 0) return true;
    return false;
}

lx_status lx_render_create(lx_scene* s, const lx_render_opts* o, lx_render** out) {
    LX_TRY
    LX_REQUIRE(s && o && out, LX_ERR_INVALID, "null argument");
    LX_REQUIRE(s->accel_built, LX_ERR_STATE, "call lx_scene_build_accel first");
    *out = nullptr;
    validate_flags(s, o->dmis, o->emis);
    lx_validate_views(o->material_debug_mode, o->shading_model);
    lx_context* ctx = s->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    lx_render* r = new lx_render();
    try {
        r->scene = s;
        r->ctx = ctx;
        r->opts = *o;
        r->cam = o->camera ? *o->camera : s->cam;
        LX_REQUIRE(r->cam.width > 0 && r->cam.height > 0, LX_ERR_INVALID, "camera resolution must be positive");
        r->depth = lx_job_depth(o->max_depth, s->trace_depth, o->material_debug_mode);
        LX_REQUIRE(r->depth <= LX_MAX_DEPTH, LX_ERR_INVALID, "max_depth above the supported maximum");
        r->npix = (uint32_t)r->cam.width * (uint32_t)r->cam.height;
        r->wb.alloc(ctx, r->npix, o->dmis != 0, o->emis != 0, scene_has_normal_maps(s));
        r->accum.alloc(ctx, (size_t)r->npix * 12);
        LX_HIP(lx_launch_fill_u32(r->accum.as<uint32_t>(), (size_t)r->npix * 3, 0u, ctx->stream));
        r->jobs.alloc(ctx, sizeof(LxJob));
        r->job_spp.alloc(ctx, 16);
        r->counters.alloc(ctx, sizeof(LxCounters));
        LX_HIP(lx_launch_fill_u32(r->counters.as<uint32_t>(), sizeof(LxCounters) / 4, 0u, ctx->stream));
        r->sched.alloc(ctx, sizeof(LxSched));
        render_grids(ctx, r->npix, r->trace_grid, r->shade_grid, r->cam_grid);
        r->spill.alloc(ctx, (size_t)lx_spill_words(r->trace_grid) * 4);
        LxJob& J = r->job;
        std::memset(&J, 0, sizeof(J));
        J.scene = s->d_scene.as<LxDevScene>();
        J.cam = r->cam;
        J.num_pix = r->npix;
        J.accum_off = 0;
        J.depth = r->depth;
        J.flags = lx_job_flags(o->material_debug_mode, o->shading_model);
        J.dmis = o->dmis;
        J.emis = o->emis;
        J.intensity = o->envmap_intensity;
        J.seed = o->seed;
        LX_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        delete r;
        throw;
    }
    s->refs++;
    *out = r;
    LX_CATCH
}

static void upload_job(lx_render* r, uint32_t first_iter, uint32_t n) {
    r->job.iter = (int32_t)first_iter - 1;
    r->job.spp_done = 0;
    r->job.active = 0;
    r->host_spp = (int32_t)n; /* source of an async copy: must outlive the call */
    LX_HIP(hipMemcpyAsync(r->jobs.ptr, &r->job, sizeof(LxJob), hipMemcpyHostToDevice, r->ctx->stream));
    LX_HIP(hipMemcpyAsync(r->job_spp.ptr, &r->host_spp, 4, hipMemcpyHostToDevice, r->ctx->stream));
}

static void enqueue(lx_render* r, hipEvent_t* ev) {
    lx_enqueue_round(r->ctx, r->jobs.as<LxJob>(), r->job_spp.as<int32_t>(), 1, r->counters.as<LxCounters>(),
                     r->sched.as<LxSched>(), r->npix, r->wb, r->accum.as<float>(), r->spill.as<uint32_t>(), r->depth,
                     r->trace_grid, r->shade_grid, r->cam_grid, r->scene->dev.bvh_width,
                     r->scene->d_scene.as<LxDevScene>(), ev, r->counting, nullptr, r->opts.dmis != 0);
}

/* Buffers and streams for `rounds` rounds in flight, made on first use.
 * Returns false (nothing new kept) when the HBM budget cannot hold them: the
 * render then runs its rounds one after another. */
static bool overlap_setup(lx_render* r, int rounds) {
    lx_context* ctx = r->ctx;
    int have = r->sets_ready;
    if (have < rounds) {
        try {
            for (int p = have; p < rounds; p++) {
                auto& O = r->extra[p - 1];
                O.wb.alloc(ctx, r->npix, r->opts.dmis != 0, r->opts.emis != 0, scene_has_normal_maps(r->scene));
                O.jobs.alloc(ctx, sizeof(LxJob));
                O.job_spp.alloc(ctx, 16);
                O.counters.alloc(ctx, sizeof(LxCounters));
                O.sched.alloc(ctx, sizeof(LxSched));
                O.spill.alloc(ctx, r->spill.bytes);
            }
            for (int p = 0; p < rounds; p++)
                if (!r->color[p].ptr) r->color[p].alloc(ctx, (size_t)r->npix * 12);
        } catch (const LxError& e) {
            if (e.status != LX_ERR_BUDGET) throw;
            for (int p = have; p < rounds; p++) {
                r->extra[p - 1] = lx_render::RoundSet();
                r->color[p].release();
            }
            if (have == 1) r->color[0].release();
            return false;
        }
        for (int p = have; p < rounds; p++)
            LX_HIP(lx_launch_fill_u32(r->extra[p - 1].counters.as<uint32_t>(), sizeof(LxCounters) / 4, 0u,
                                      ctx->stream));
        r->sets_ready = rounds;
    }
    for (int p = 0; p < rounds; p++) {
        if (!r->pstream[p]) LX_HIP(hipStreamCreateWithFlags(&r->pstream[p], hipStreamNonBlocking));
        if (!r->ev_done[p]) LX_HIP(hipEventCreateWithFlags(&r->ev_done[p], hipEventDisableTiming));
        if (!r->ev_free[p]) LX_HIP(hipEventCreateWithFlags(&r->ev_free[p], hipEventDisableTiming));
    }
    if (!r->ev_up) LX_HIP(hipEventCreateWithFlags(&r->ev_up, hipEventDisableTiming));
    return true;
}

/* n >= 2 rounds, D = min(rounds, n) in flight: round k on stream k % D with
 * set k % D; the job copy of set p starts at iteration first_iter + p and
 * steps by D.  Round k's colours are added to the accumulation on the
 * context's stream after round k - 1's, and round k + D waits for that add
 * before it reuses the colour buffer. */
static void iterate_overlapped(lx_render* r, uint32_t first_iter, uint32_t n, int rounds) {
    lx_context* ctx = r->ctx;
    hipStream_t ms = ctx->stream;
    const int D = (int)std::min<uint32_t>((uint32_t)rounds, n);
    auto set_of = [&](int p) -> lx_render::RoundSet* { return p == 0 ? nullptr : &r->extra[p - 1]; };
    for (int p = 0; p < D; p++) {
        lx_render::RoundSet* O = set_of(p);
        /* sources of async copies: members of the render, untouched until the next call */
        LxJob& J = r->ov_job[p];
        J = r->job;
        J.iter = (int32_t)first_iter - D + p;
        J.spp_done = 0;
        J.active = 0;
        J.flags = (J.flags & ~LX_JF_STEP_MASK) | ((uint32_t)D << LX_JF_STEP_SHIFT);
        r->ov_spp[p] = (int32_t)((n - (uint32_t)p + (uint32_t)D - 1) / (uint32_t)D); /* rounds of this set */
        LX_HIP(hipMemcpyAsync(O ? O->jobs.ptr : r->jobs.ptr, &J, sizeof(LxJob), hipMemcpyHostToDevice, ms));
        LX_HIP(hipMemcpyAsync(O ? O->job_spp.ptr : r->job_spp.ptr, &r->ov_spp[p], 4, hipMemcpyHostToDevice, ms));
    }
    LX_HIP(hipEventRecord(r->ev_up, ms));
    for (int p = 0; p < D; p++) LX_HIP(hipStreamWaitEvent(r->pstream[p], r->ev_up, 0));
    const size_t nf = (size_t)r->npix * 3;
    for (uint32_t k = 0; k < n; k++) {
        int p = (int)(k % (uint32_t)D);
        lx_render::RoundSet* O = set_of(p);
        hipStream_t s = r->pstream[p];
        if (k >= (uint32_t)D) LX_HIP(hipStreamWaitEvent(s, r->ev_free[p], 0));
        LX_HIP(lx_launch_fill_u32(r->color[p].as<uint32_t>(), nf, 0x80000000u, s));
        lx_enqueue_round(ctx, (O ? O->jobs : r->jobs).as<LxJob>(), (O ? O->job_spp : r->job_spp).as<int32_t>(), 1,
                         (O ? O->counters : r->counters).as<LxCounters>(), (O ? O->sched : r->sched).as<LxSched>(),
                         r->npix, O ? O->wb : r->wb, r->color[p].as<float>(), (O ? O->spill : r->spill).as<uint32_t>(),
                         r->depth, r->trace_grid, r->shade_grid, r->cam_grid, r->scene->dev.bvh_width,
                         r->scene->d_scene.as<LxDevScene>(), nullptr, false, s, r->opts.dmis != 0, true);
        LX_HIP(hipEventRecord(r->ev_done[p], s));
        LX_HIP(hipStreamWaitEvent(ms, r->ev_done[p], 0));
        LX_HIP(lx_launch_add_f32(r->accum.as<float>(), r->color[p].as<float>(), nf, ms));
        LX_HIP(hipEventRecord(r->ev_free[p], ms));
    }
}

/* Buffers for wide rounds of R samples, D sets, made on first use (and
 * remade when R changes).  Returns false when the HBM budget cannot hold
 * them (the render then iterates without wide rounds). */
static bool wide_setup(lx_render* r, uint32_t R, int D) {
    lx_context* ctx = r->ctx;
    uint64_t cap64 = (uint64_t)R * r->npix;
    if (cap64 >= (1ull << 31)) return false;
    const uint32_t cap = (uint32_t)cap64;
    try {
        if (!r->wide_counters.ptr) {
            r->wide_counters.alloc(ctx, lx_render::kMaxRounds * sizeof(LxCounters));
            LX_HIP(lx_launch_fill_u32(r->wide_counters.as<uint32_t>(),
                                      lx_render::kMaxRounds * sizeof(LxCounters) / 4, 0u, ctx->stream));
        }
    } catch (const LxError& e) {
        if (e.status != LX_ERR_BUDGET) throw;
        return false;
    }
    if (r->wide_frames != R && r->wide_sets > 0) {
        /* the previous calls' rounds end before their buffers are freed:
         * every round's merge waits for the round on the context stream */
        LX_HIP(hipStreamSynchronize(ctx->stream));
        for (int p = 0; p < lx_render::kMaxRounds; p++) r->wide[p] = lx_render::WideSet();
        r->wide_sets = 0;
    }
    r->wide_frames = R;
    if (r->wide_sets < D) {
        uint32_t tg, sg, cg;
        render_grids(ctx, cap, tg, sg, cg);
        try {
            for (int p = r->wide_sets; p < D; p++) {
                auto& W = r->wide[p];
                W.wb.alloc(ctx, cap, r->opts.dmis != 0, r->opts.emis != 0, scene_has_normal_maps(r->scene));
                W.jobs.alloc(ctx, (size_t)R * sizeof(LxJob));
                W.job_spp.alloc(ctx, (size_t)R * 4);
                W.sched.alloc(ctx, sizeof(LxSched));
                W.spill.alloc(ctx, (size_t)lx_spill_words(tg) * 4);
                W.color.alloc(ctx, (size_t)cap * 12);
                W.hj.assign(R, LxJob{});
                W.hs.assign(R, 0);
            }
        } catch (const LxError& e) {
            if (e.status != LX_ERR_BUDGET) throw;
            for (int p = r->wide_sets; p < D; p++) r->wide[p] = lx_render::WideSet();
            return false;
        }
        r->wide_sets = D;
        r->wide_grid[0] = tg;
        r->wide_grid[1] = sg;
        r->wide_grid[2] = cg;
    }
    for (int p = 0; p < D; p++) {
        if (!r->pstream[p]) LX_HIP(hipStreamCreateWithFlags(&r->pstream[p], hipStreamNonBlocking));
        if (!r->ev_done[p]) LX_HIP(hipEventCreateWithFlags(&r->ev_done[p], hipEventDisableTiming));
        if (!r->ev_free[p]) LX_HIP(hipEventCreateWithFlags(&r->ev_free[p], hipEventDisableTiming));
    }
    if (!r->ev_up) LX_HIP(hipEventCreateWithFlags(&r->ev_up, hipEventDisableTiming));
    return true;
}

/* n >= 2 samples in wide rounds of R samples: chunk c (samples first_iter +
 * c R .. + R - 1, fewer in the last) runs on stream c % D with set c % D,
 * D = min(rounds, chunks).  Copy k of set p starts at sample p R + k and
 * steps by D R, into region k of the set's colour pool; a chunk's regions
 * are added to the accumulation in sample order after the previous chunk's,
 * and chunk c + D waits for that add before it reuses the pool.  With
 * timing on, events on each chunk's stream bracket its kernels, read back
 * after a single synchronisation. */
static void iterate_wide(lx_render* r, uint32_t first_iter, uint32_t n, int rounds) {
    lx_context* ctx = r->ctx;
    hipStream_t ms = ctx->stream;
    const uint32_t R = r->wide_frames;
    const uint32_t chunks = (n + R - 1) / R;
    const int D = (int)std::min<uint32_t>((uint32_t)rounds, chunks);
    const uint32_t step = (uint32_t)D * R;
    for (int p = 0; p < D; p++) {
        auto& W = r->wide[p];
        for (uint32_t k = 0; k < R; k++) {
            LxJob& J = W.hj[k];
            const uint32_t v = (uint32_t)p * R + k;
            J = r->job;
            J.iter = (int32_t)first_iter + (int32_t)v - (int32_t)step;
            J.spp_done = 0;
            J.active = 0;
            J.accum_off = k * r->npix;
            J.flags = (J.flags & ~LX_JF_STEP_MASK) | (step << LX_JF_STEP_SHIFT);
            W.hs[k] = v < n ? (int32_t)((n - v + step - 1) / step) : 0;
        }
        LX_HIP(hipMemcpyAsync(W.jobs.ptr, W.hj.data(), (size_t)R * sizeof(LxJob), hipMemcpyHostToDevice, ms));
        LX_HIP(hipMemcpyAsync(W.job_spp.ptr, W.hs.data(), (size_t)R * 4, hipMemcpyHostToDevice, ms));
    }
    LX_HIP(hipEventRecord(r->ev_up, ms));
    for (int p = 0; p < D; p++) LX_HIP(hipStreamWaitEvent(r->pstream[p], r->ev_up, 0));
    const size_t nf = (size_t)r->npix * 3;
    LxCounters* cnt = r->wide_counters.as<LxCounters>();
    const int per = 2 * r->depth + 1;
    std::vector<hipEvent_t> ev;
    if (r->timing) {
        ev.resize((size_t)(per + 2) * chunks);
        for (auto& e : ev) LX_HIP(hipEventCreate(&e));
    }
    for (uint32_t c = 0; c < chunks; c++) {
        const int p = (int)(c % (uint32_t)D);
        auto& W = r->wide[p];
        hipStream_t s = r->pstream[p];
        const uint32_t rc = std::min(R, n - c * R); /* the copies this chunk schedules: k < rc */
        if (c >= (uint32_t)D) LX_HIP(hipStreamWaitEvent(s, r->ev_free[p], 0));
        LX_HIP(lx_launch_fill_u32(W.color.as<uint32_t>(), nf * rc, 0x80000000u, s));
        hipEvent_t* e = r->timing ? &ev[(size_t)c * (per + 2)] : nullptr;
        if (e) LX_HIP(hipEventRecord(e[0], s));
        lx_enqueue_round(ctx, W.jobs.as<LxJob>(), W.job_spp.as<int32_t>(), R, cnt + p, W.sched.as<LxSched>(),
                         R * r->npix, W.wb, W.color.as<float>(), W.spill.as<uint32_t>(), r->depth, r->wide_grid[0],
                         r->wide_grid[1], r->wide_grid[2], r->scene->dev.bvh_width,
                         r->scene->d_scene.as<LxDevScene>(), e ? e + 1 : nullptr, false, s, r->opts.dmis != 0, true);
        if (e) LX_HIP(hipEventRecord(e[per + 1], s));
        LX_HIP(hipEventRecord(r->ev_done[p], s));
        LX_HIP(hipStreamWaitEvent(ms, r->ev_done[p], 0));
        LX_HIP(lx_launch_add_regions(r->accum.as<float>(), W.color.as<float>(), nf, rc, ms));
        LX_HIP(hipEventRecord(r->ev_free[p], ms));
    }
    if (r->timing) {
        LX_HIP(hipStreamSynchronize(ms));
        const int wf = lx_wavefront_bounces(ctx, r->depth);
        for (uint32_t c = 0; c < chunks; c++) {
            hipEvent_t* e = &ev[(size_t)c * (per + 2)];
            float t;
            for (int b = 0; b < wf; b++) {
                LX_HIP(hipEventElapsedTime(&t, e[1 + 2 * b], e[2 + 2 * b]));
                r->trace_ms += t;
                LX_HIP(hipEventElapsedTime(&t, e[2 + 2 * b], e[3 + 2 * b]));
                r->shade_ms += t;
            }
            if (wf < r->depth) {
                LX_HIP(hipEventElapsedTime(&t, e[1 + 2 * wf], e[2 + 2 * wf]));
                r->tail_ms += t;
            }
            LX_HIP(hipEventElapsedTime(&t, e[0], e[1]));
            r->other_ms += t;
            LX_HIP(hipEventElapsedTime(&t, e[per], e[per + 1]));
            r->other_ms += t;
        }
        r->trace_launches += (uint32_t)wf * chunks;
        r->shade_launches += (uint32_t)wf * chunks;
        if (wf < r->depth) r->tail_launches += chunks;
        for (auto& x : ev) (void)hipEventDestroy(x);
    }
}

/* Every round set's counters. */
static void read_counters(lx_render* r, LxCounters& c) {
    LX_HIP(hipMemcpy(&c, r->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
    for (int p = 1; p < r->sets_ready; p++) {
        LxCounters o;
        LX_HIP(hipMemcpy(&o, r->extra[p - 1].counters.ptr, sizeof(o), hipMemcpyDeviceToHost));
        lx_counters_add(c, o);
    }
    if (r->wide_counters.ptr) {
        LxCounters w[lx_render::kMaxRounds];
        LX_HIP(hipMemcpy(w, r->wide_counters.ptr, sizeof(w), hipMemcpyDeviceToHost));
        for (auto& o : w) lx_counters_add(c, o);
    }
}

lx_status lx_render_iterate(lx_render* r, uint32_t first_iter, uint32_t n) {
    LX_TRY
    LX_REQUIRE(r, LX_ERR_INVALID, "null render");
    LX_REQUIRE(first_iter >= 1, LX_ERR_INVALID, "iterations are 1-based");
    if (n == 0) return LX_OK;
    lx_context* ctx = r->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    const int rounds = 1 + ctx->tuning.overlap_rounds;
    const uint32_t R = (uint32_t)std::max(1, ctx->tuning.round_frames);
    if (n >= 2 && R >= 2 && !r->counting && wide_setup(r, R, rounds)) {
        iterate_wide(r, first_iter, n, rounds);
        r->spp += n;
        return LX_OK;
    }
    if (n >= 2 && rounds >= 2 && !r->timing && !r->counting && overlap_setup(r, rounds)) {
        iterate_overlapped(r, first_iter, n, rounds);
        r->spp += n;
        return LX_OK;
    }
    upload_job(r, first_iter, n);
    if (r->timing || r->counting) {
        /* eager launches; with timing, an event before and after each
         * kernel of every round (only those the stats read), read back
         * after a single synchronisation */
        int per = 2 * r->depth + 1;
        std::vector<hipEvent_t> ev;
        if (r->timing) {
            ev.resize((size_t)per * n + 2 * n);
            for (auto& e : ev) LX_HIP(hipEventCreate(&e));
        }
        for (uint32_t k = 0; k < n; k++) {
            hipEvent_t* e = r->timing ? &ev[(size_t)k * (per + 2)] : nullptr;
            if (e) LX_HIP(hipEventRecord(e[0], ctx->stream));
            enqueue(r, e ? e + 1 : nullptr);
            if (e) LX_HIP(hipEventRecord(e[per + 1], ctx->stream));
        }
        LX_HIP(hipStreamSynchronize(ctx->stream));
        if (r->timing) {
            int wf = lx_wavefront_bounces(ctx, r->depth);
            for (uint32_t k = 0; k < n; k++) {
                hipEvent_t* e = &ev[(size_t)k * (per + 2)];
                float ms;
                for (int b = 0; b < wf; b++) {
                    LX_HIP(hipEventElapsedTime(&ms, e[1 + 2 * b], e[2 + 2 * b]));
                    r->trace_ms += ms;
                    LX_HIP(hipEventElapsedTime(&ms, e[2 + 2 * b], e[3 + 2 * b]));
                    r->shade_ms += ms;
                }
                if (wf < r->depth) {
                    LX_HIP(hipEventElapsedTime(&ms, e[1 + 2 * wf], e[2 + 2 * wf]));
                    r->tail_ms += ms;
                }
                LX_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
                r->other_ms += ms;
                LX_HIP(hipEventElapsedTime(&ms, e[per], e[per + 1]));
                r->other_ms += ms;
            }
            r->trace_launches += wf * n;
            r->shade_launches += wf * n;
            if (wf < r->depth) r->tail_launches += n;
            for (auto& e : ev) (void)hipEventDestroy(e);
        }
    } else if (ctx->tuning.use_graph) {
        if (r->graph && std::memcmp(&r->graph_tuning, &ctx->tuning, sizeof(lx_tuning)) != 0) {
            /* the launch shapes changed since the capture */
            LX_HIP(hipStreamSynchronize(ctx->stream));
            (void)hipGraphExecDestroy(r->graph);
            (void)hipGraphDestroy(r->graph_src);
            r->graph = nullptr;
            r->graph_src = nullptr;
        }
        if (!r->graph) {
            r->graph_tuning = ctx->tuning;
            LX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue(r, nullptr);
            } catch (...) {
                hipGraph_t g;
                (void)hipStreamEndCapture(ctx->stream, &g);
                throw;
            }
            LX_HIP(hipStreamEndCapture(ctx->stream, &r->graph_src));
            LX_HIP(hipGraphInstantiate(&r->graph, r->graph_src, nullptr, nullptr, 0));
        }
        for (uint32_t k = 0; k < n; k++) LX_HIP(hipGraphLaunch(r->graph, ctx->stream));
    } else {
        for (uint32_t k = 0; k < n; k++) enqueue(r, nullptr);
    }
    r->spp += n;
    LX_CATCH
}

lx_status lx_render_sync(lx_render* r) {
    LX_TRY
    LX_REQUIRE(r, LX_ERR_INVALID, "null render");
    r->ctx->set_device();
    LX_HIP(hipStreamSynchronize(r->ctx->stream));
    uint32_t ov = 0;
    LX_HIP(hipMemcpy(&ov, (uint8_t*)r->counters.ptr + offsetof(LxCounters, stack_overflow), 4, hipMemcpyDeviceToHost));
    for (int p = 1; p < r->sets_ready; p++) {
        uint32_t o = 0;
        LX_HIP(hipMemcpy(&o, (uint8_t*)r->extra[p - 1].counters.ptr + offsetof(LxCounters, stack_overflow), 4,
                         hipMemcpyDeviceToHost));
        ov |= o;
    }
    for (int p = 0; r->wide_counters.ptr && p < lx_render::kMaxRounds; p++) {
        uint32_t o = 0;
        LX_HIP(hipMemcpy(&o, (uint8_t*)r->wide_counters.ptr + p * sizeof(LxCounters) + offsetof(LxCounters, stack_overflow),
                         4, hipMemcpyDeviceToHost));
        ov |= o;
    }
    LX_REQUIRE(ov == 0, LX_ERR_STATE, "traversal stack overflow (BVH deeper than the supported stack)");
    LX_CATCH
}

lx_status lx_render_read_accum(lx_render* r, float* rgb, uint32_t* spp) {
    LX_TRY
    LX_REQUIRE(r, LX_ERR_INVALID, "null render");
    lx_status st = lx_render_sync(r);
    if (st != LX_OK) return st;
    if (rgb)
        LX_HIP(hipMemcpy(rgb, r->accum.ptr, (size_t)r->npix * 12, hipMemcpyDeviceToHost));
    if (spp) *spp = r->spp;
    LX_CATCH
}

lx_status lx_render_read_ldr(lx_render* r, uint8_t* rgba, int32_t tonemap) {
    LX_TRY
    LX_REQUIRE(r && rgba, LX_ERR_INVALID, "null argument");
    LX_REQUIRE(r->spp > 0, LX_ERR_STATE, "no samples accumulated");
    std::lock_guard<std::mutex> lk(r->ctx->mu);
    r->ctx->set_device();
    if (!r->ldr.ptr) r->ldr.alloc(r->ctx, (size_t)r->npix * 4);
    int mode = r->opts.material_debug_mode != LX_DEBUG_OFF ? -1 : tonemap;
    LX_HIP(lx_launch_tonemap(r->accum.as<float>(), r->npix, (float)r->spp, mode, r->ldr.as<uint8_t>(),
                             r->ctx->stream));
    LX_HIP(hipMemcpyAsync(rgba, r->ldr.ptr, (size_t)r->npix * 4, hipMemcpyDeviceToHost, r->ctx->stream));
    LX_HIP(hipStreamSynchronize(r->ctx->stream));
    LX_CATCH
}

lx_status lx_render_reset(lx_render* r) {
    LX_TRY
    LX_REQUIRE(r, LX_ERR_INVALID, "null render");
    std::lock_guard<std::mutex> lk(r->ctx->mu);
    r->ctx->set_device();
    LX_HIP(lx_launch_fill_u32(r->accum.as<uint32_t>(), (size_t)r->npix * 3, 0u, r->ctx->stream));
    LX_HIP(lx_launch_fill_u32(r->counters.as<uint32_t>(), sizeof(LxCounters) / 4, 0u, r->ctx->stream));
    for (int p = 1; p < r->sets_ready; p++)
        LX_HIP(lx_launch_fill_u32(r->extra[p - 1].counters.as<uint32_t>(), sizeof(LxCounters) / 4, 0u,
                                  r->ctx->stream));
    if (r->wide_counters.ptr)
        LX_HIP(lx_launch_fill_u32(r->wide_counters.as<uint32_t>(), lx_render::kMaxRounds * sizeof(LxCounters) / 4, 0u,
                                  r->ctx->stream));
    LX_HIP(hipStreamSynchronize(r->ctx->stream));
    r->spp = 0;
    r->trace_ms = r->shade_ms = r->other_ms = r->tail_ms = 0;
    r->trace_launches = r->shade_launches = r->tail_launches = 0;
    LX_CATCH
}

lx_status lx_render_set_timing(lx_render* r, int32_t flags) {
    LX_TRY
    LX_REQUIRE(r, LX_ERR_INVALID, "null render");
    r->timing = (flags & LX_TIMING_EVENTS) != 0;
    r->counting = (flags & LX_TIMING_COUNT_WORK) != 0;
    LX_CATCH
}

lx_status lx_render_stats_get(lx_render* r, lx_render_stats* out) {
    LX_TRY
    LX_REQUIRE(r && out, LX_ERR_INVALID, "null argument");
    lx_status st = lx_render_sync(r);
    if (st != LX_OK) return st;
    LxCounters c;
    read_counters(r, c);
    std::memset(out, 0, sizeof(*out));
    out->radiance_rays = c.rad_rays;
    out->dl_shadow_rays = c.dl_rays;
    out->env_shadow_rays = c.env_rays;
    out->path_bounces = c.bounces;
    out->iterations = r->spp;
    out->trace_ms = r->trace_ms;
    out->shade_ms = r->shade_ms;
    out->other_ms = r->other_ms;
    out->trace_launches = r->trace_launches;
    out->shade_launches = r->shade_launches;
    const LxWorkCnt &w0 = c.work[0], &w1 = c.work[1];
    out->closest_nodes = w0.closest_nodes + w1.closest_nodes;
    out->closest_tris = w0.closest_tris + w1.closest_tris;
    out->closest_insts = w0.closest_insts + w1.closest_insts;
    out->any_nodes = w0.any_nodes + w1.any_nodes;
    out->any_tris = w0.any_tris + w1.any_tris;
    out->any_insts = w0.any_insts + w1.any_insts;
    out->tail_closest_rays = w1.closest_rays;
    out->tail_any_rays = w1.any_rays;
    out->tail_closest_nodes = w1.closest_nodes;
    out->tail_closest_tris = w1.closest_tris;
    out->tail_closest_insts = w1.closest_insts;
    out->tail_any_nodes = w1.any_nodes;
    out->tail_any_tris = w1.any_tris;
    out->tail_any_insts = w1.any_insts;
    out->tail_ms = r->tail_ms;
    out->tail_launches = r->tail_launches;
    out->tail_bounces = c.tail_bounces;
    LX_CATCH
}

void lx_render_destroy(lx_render* r) {
    if (!r) return;
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    if (r->graph) (void)hipGraphExecDestroy(r->graph);
    if (r->graph_src) (void)hipGraphDestroy(r->graph_src);
    for (int p = 0; p < lx_render::kMaxRounds; p++) {
        if (r->pstream[p]) (void)hipStreamDestroy(r->pstream[p]);
        if (r->ev_done[p]) (void)hipEventDestroy(r->ev_done[p]);
        if (r->ev_free[p]) (void)hipEventDestroy(r->ev_free[p]);
    }
    if (r->ev_up) (void)hipEventDestroy(r->ev_up);
    lx_scene* s = r->scene;
    delete r;
    lx_scene_release(s);
}

lx_status lx_debug_bounce(lx_render* r, uint32_t iter, uint32_t* survivors, uint32_t* survivor_counts,
                          float* records, uint32_t* record_counts) {
    LX_TRY
    LX_REQUIRE(r && survivor_counts && record_counts, LX_ERR_INVALID, "null argument");
    LX_REQUIRE(iter >= 1, LX_ERR_INVALID, "iterations are 1-based");
    lx_context* ctx = r->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->set_device();
    hipStream_t s = ctx->stream;
    upload_job(r, iter, 1);
    LxJob* dj = r->jobs.as<LxJob>();
    LxCounters* dc = r->counters.as<LxCounters>();
    LX_HIP(lx_launch_begin_round(dj, r->job_spp.as<int32_t>(), 1, dc, r->sched.as<LxSched>(), r->npix, s));
    LX_HIP(lx_launch_camera(dj, r->sched.as<LxSched>(), dc, r->wb.A, r->cam_grid, s,
                            (uint32_t)r->ctx->tuning.pixel_tile, (uint32_t)r->ctx->tuning.tile_band));
    LxPathBuf* cur = &r->wb.A;
    LxPathBuf* nxt = &r->wb.B;
    size_t N = r->npix;
    std::vector<float> t(N), nx(N), ny(N), nz(N), u(N), v(N);
    std::vector<int32_t> mat(N);
    std::vector<uint32_t> pix(N);
    for (int b = 0; b < r->depth; b++) {
        LX_HIP(lx_launch_trace(dj, *cur, r->wb.R, dc, b, r->spill.as<uint32_t>(), r->trace_grid,
                               r->scene->dev.bvh_width, s));
        LX_HIP(lx_launch_shade(dj, *cur, *nxt, r->wb.R, dc, b, r->accum.as<float>(), r->shade_grid, s));
        LX_HIP(hipStreamSynchronize(s));
        uint32_t cnt[2];
        LX_HIP(hipMemcpy(cnt, (uint8_t*)r->counters.ptr + 4 * b, 8, hipMemcpyDeviceToHost));
        uint32_t nb = cnt[0];
        record_counts[b] = nb;
        if (nb) {
            LX_HIP(hipMemcpy(t.data(), r->wb.R.t, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(nx.data(), r->wb.R.nx, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(ny.data(), r->wb.R.ny, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(nz.data(), r->wb.R.nz, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(u.data(), r->wb.R.u, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(v.data(), r->wb.R.v, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(mat.data(), r->wb.R.mat, nb * 4, hipMemcpyDeviceToHost));
            LX_HIP(hipMemcpy(pix.data(), cur->pix, nb * 4, hipMemcpyDeviceToHost));
            std::vector<uint32_t> order(nb);
            std::iota(order.begin(), order.end(), 0u);
            std::sort(order.begin(), order.end(),
                      [&](uint32_t a, uint32_t c) { return (pix[a] & 0x7fffffffu) < (pix[c] & 0x7fffffffu); });
            if (records)
                for (uint32_t k = 0; k < nb; k++) {
                    uint32_t q = order[k];
                    bool hit = t[q] > 0.0f;
                    float* o = records + ((size_t)b * N + k) * 8;
                    o[0] = hit ? t[q] : -1.0f;
                    o[1] = hit ? nx[q] : 0.0f;
                    o[2] = hit ? ny[q] : 0.0f;
                    o[3] = hit ? nz[q] : 0.0f;
                    o[4] = hit ? u[q] : 0.0f;
                    o[5] = hit ? v[q] : 0.0f;
                    o[6] = hit ? (float)mat[q] : -1.0f;
                    o[7] = (float)(pix[q] & 0x7fffffffu);
                }
        }
        bool compact = b + 1 < r->depth;
        uint32_t ns = compact ? cnt[1] : 0u;
        survivor_counts[b] = ns;
        if (ns) {
            std::vector<uint32_t> sp(ns);
            LX_HIP(hipMemcpy(sp.data(), nxt->pix, ns * 4, hipMemcpyDeviceToHost));
            for (auto& x : sp) x &= 0x7fffffffu;
            std::sort(sp.begin(), sp.end());
            if (survivors) std::memcpy(survivors + (size_t)b * N, sp.data(), ns * 4);
        }
        std::swap(cur, nxt);
        if (nb == 0) {
            for (int b2 = b + 1; b2 < r->depth; b2++) survivor_counts[b2] = record_counts[b2] = 0;
            break;
        }
    }
    LX_HIP(lx_launch_end_round(dj, 1, s));
    LX_HIP(hipStreamSynchronize(s));
    r->spp += 1;
    LX_CATCH
}

} /* extern "C" */
