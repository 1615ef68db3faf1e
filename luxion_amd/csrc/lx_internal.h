/*
 * lx_internal.h -- host-side objects behind the C ABI.
 */
#ifndef LX_INTERNAL_H
#define LX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lx_kernels.h"
#include "lx_types.h"

/* bounces before the tail kernel takes over in a fresh context; negative
 * runs every bounce as its own launch */
#ifndef LX_DEFAULT_WAVEFRONT_BOUNCES
#define LX_DEFAULT_WAVEFRONT_BOUNCES 0
#endif

/* ------------------------------------------------------------- errors */
void lx_set_error(const std::string& msg);

/* C-ABI bodies: errors become a status and a message */
#define LX_TRY try {
#define LX_CATCH                                                                                           \
    }                                                                                                      \
    catch (const LxError& e) {                                                                             \
        lx_set_error(e.msg);                                                                               \
        return e.status;                                                                                   \
    }                                                                                                      \
    catch (const std::bad_alloc&) {                                                                        \
        lx_set_error("host allocation failed");                                                            \
        return LX_ERR_OOM;                                                                                 \
    }                                                                                                      \
    catch (const std::exception& e) {                                                                      \
        lx_set_error(e.what());                                                                            \
        return LX_ERR_INVALID;                                                                             \
    }                                                                                                      \
    lx_set_error("");                                                                                      \
    return LX_OK;

struct LxError {
    lx_status status;
    std::string msg;
};

#define LX_HIP(call)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw LxError{LX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};                 \
    } while (0)

#define LX_REQUIRE(cond, st, msg)                                                                          \
    do {                                                                                                   \
        if (!(cond)) throw LxError{(st), (msg)};                                                           \
    } while (0)

/* ------------------------------------------------------------ context */
struct lx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::atomic<uint64_t> in_use{0};
    int num_cus = 256;
    lx_tuning tuning{};
    std::mutex mu; /* serialises device work from the served form's thread */
    std::atomic<int> refs{1}; /* user handle + live scenes */
    /* this device's copy of the allocation-free trace's scratch array
     * (lx_noalloc_scratch_address, looked up at context creation) */
    uint32_t* noalloc_scratch = nullptr;

    void* dalloc(size_t bytes);           /* charged against the budget */
    void dfree(void* p, size_t bytes);
    void set_device() const;
};

/* RAII device buffer charged to a context. */
struct DevBuf {
    lx_context* ctx = nullptr;
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        ctx = o.ctx;
        ptr = o.ptr;
        bytes = o.bytes;
        o.ptr = nullptr;
        o.bytes = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(lx_context* c, size_t n) {
        release();
        ctx = c;
        bytes = n;
        ptr = n ? c->dalloc(n) : nullptr;
    }
    void upload(lx_context* c, const void* src, size_t n) {
        alloc(c, n);
        if (n) LX_HIP(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, c->stream));
    }
    void release() {
        if (ptr && ctx) ctx->dfree(ptr, bytes);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return (T*)ptr;
    }
};

/* ------------------------------------------------------ host threads */
/* Runs fn(k) for every k in [0, n) on up to 16 host threads, the calling
 * thread among them; indices are handed out one at a time from a shared
 * counter, so fn must be safe to run concurrently for distinct k. */
template <class Fn>
void lx_parallel_for(uint32_t n, const Fn& fn) {
    unsigned hw = std::thread::hardware_concurrency();
    uint32_t nt = std::min<uint32_t>(n, std::min<unsigned>(16u, hw ? hw : 1u));
    if (nt <= 1) {
        for (uint32_t k = 0; k < n; k++) fn(k);
        return;
    }
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (uint32_t k = next.fetch_add(1); k < n; k = next.fetch_add(1)) fn(k);
    };
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (uint32_t t = 1; t < nt; t++) pool.emplace_back(worker);
    worker();
    for (std::thread& t : pool) t.join();
}

/* ---------------------------------------------------------- BVH build */
struct BvhBuildResult {
    std::vector<LxWideNode> nodes; /* local indices */
    std::vector<float4> tris;     /* leaf-ordered (v0|prim, e1, e2) */
    uint32_t root = 0;            /* child ref (node index or leaf ref), local */
    int depth = 0;
    double sah = 0.0;
    /* A tree the GPU builder leaves on the device (nodes and tris above stay
     * empty): dev_count nodes already packed at the traversal's width with
     * absolute child refs (the node and triangle bases it was built for),
     * and its leaf-ordered triangles, tags applied. */
    DevBuf d_nodes, d_tris;
    uint32_t dev_count = 0;
    /* the boxes two levels below the root (at most 64; leaf children of the
     * root count as one), for the instance world boxes */
    std::vector<std::array<float, 6>> top_boxes;
    bool on_device() const { return d_nodes.ptr != nullptr || d_tris.ptr != nullptr; }
};

/* Host binned-SAH build + collapse to BVH4 (builder 1). */
void lx_build_blas_host(const float* pos, const uint32_t* idx, uint32_t ntri, int leaf_size, int bins, int width,
                        BvhBuildResult& out);
/* TLAS over instance world boxes (one instance per leaf). */
void lx_build_tlas_host(const std::vector<float>& boxes /* 6 per instance */, int width, BvhBuildResult& out);
/* GPU builder (builder 0; builder 2 uses it for BLASes of at least
 * LX_GPU_BUILD_MIN_TRIS triangles): top-down binned SAH on the device over
 * the Morton-ordered triangles, breadth-first levels while a node is larger
 * than one thread's share and one thread per subtree below that, with the
 * host builder's split and leaf rules; then the host builder's collapse, run
 * level by level on the device.  The tree stays on the device (out.d_nodes,
 * out.d_tris), its nodes numbered from node_base and its leaves from
 * tri_base.  `nv` is the mesh's vertex count; with ntag > 0 each triangle
 * is tagged as the scene's triangle records of a group are (its index within
 * its mesh, and its instance).  Returns false if the GPU path is
 * unavailable for this input (the caller then builds on the host). */
#define LX_GPU_BUILD_MIN_TRIS (1u << 14)
bool lx_build_blas_gpu(lx_context* ctx, const float* pos, size_t nv, const uint32_t* idx, uint32_t ntri,
                       int leaf_size, int bins, int width, uint32_t node_base, uint32_t tri_base,
                       const uint32_t* tag_off, const uint32_t* tag_inst, size_t ntag, BvhBuildResult& out);
void lx_pad_box(float lo[3], float hi[3]);

/* -------------------------------------------------------------- scene */
struct HostMesh {
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx;
    bool has_n = false, has_uv = false;
};

struct HostTex {
    uint32_t w = 0, h = 0;
    int32_t fmt = 0;
    std::vector<uint8_t> data;
};

struct lx_scene {
    lx_context* ctx = nullptr;
    std::atomic<int> refs{1}; /* user handle + live renders / jobs */
    std::vector<HostMesh> meshes;
    std::vector<lx_instance_desc> insts;
    std::vector<lx_material_desc> mats;
    std::vector<HostTex> texs;
    bool has_env = false;
    HostTex env;
    lx_camera_desc cam{};
    int trace_depth = 8;
    int cdf_mode = LX_CDF_FP64;

    /* host precompute */
    std::vector<float> inv; /* 12 per instance */
    std::vector<float> marginal, conditional;
    std::vector<uint32_t> emissive;
    std::vector<float> emissive_prefix;
    std::vector<float> tri_cdf;
    std::vector<uint32_t> tri_cdf_off;
    float total_area = 0.0f;

    /* device */
    DevBuf d_nodes, d_tris, d_inst, d_groups, d_idx, d_pos, d_nrm, d_uv, d_mats, d_texdesc, d_texels, d_env, d_marg, d_cond,
        d_emis, d_emis_prefix, d_tri_cdf, d_tri_cdf_off, d_scene;
    std::vector<DevBuf> d_tex;
    LxDevScene dev{};
    bool accel_built = false;
    lx_build_opts build_opts{}; /* of the last build (the autotuner rebuilds at other widths) */
    lx_scene_info info{};
};

void lx_scene_init(lx_scene* s, const lx_scene_desc* d); /* copies + precompute + uploads */
extern "C" void lx_scene_release(lx_scene* s);
extern "C" void lx_context_release(lx_context* c);
void lx_scene_accel(lx_scene* s, const lx_build_opts* o);
void lx_affine_inverse(const float* m, float* out);

/* ------------------------------------------------------------- render */
struct WaveBufs {
    DevBuf pathA, pathB, rec, sort;
    LxPathBuf A{}, B{};
    LxRecBuf R{};
    uint32_t *keys = nullptr, *order = nullptr, *hist = nullptr; /* queue sort stage */
    uint32_t cap = 0;
    void alloc(lx_context* ctx, uint32_t cap, bool dmis, bool emis, bool tangents);
};

struct lx_render {
    lx_scene* scene = nullptr;
    lx_context* ctx = nullptr;
    lx_render_opts opts{};
    lx_camera_desc cam{};
    int depth = 8;
    uint32_t npix = 0;
    uint32_t spp = 0;
    WaveBufs wb;
    DevBuf accum, jobs, job_spp, counters, sched, spill, ldr;
    LxJob job{};
    int32_t host_spp = 0;
    uint32_t trace_grid = 0, shade_grid = 0, cam_grid = 0;
    bool timing = false;
    bool counting = false;
    double trace_ms = 0, shade_ms = 0, other_ms = 0, tail_ms = 0;
    uint32_t trace_launches = 0, shade_launches = 0, tail_launches = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    lx_tuning graph_tuning{}; /* the tuning the graph was captured with */

    /* overlapped rounds (tuning.overlap_rounds + 1 rounds in flight, at most
     * kMaxRounds): round k runs on pstream[k % D] with set k % D -- set 0 is
     * the buffers above, set p > 0 is extra[p - 1]; its colours go to
     * color[k % D] (filled with -0.0, so the one add per pixel copies the
     * path's colour bits), added to accum in round order on the context's
     * stream */
    static constexpr int kMaxRounds = 4;
    struct RoundSet {
        WaveBufs wb;
        DevBuf jobs, job_spp, counters, sched, spill;
    } extra[kMaxRounds - 1];
    int sets_ready = 1; /* round sets allocated (set 0 always) */
    LxJob ov_job[kMaxRounds]{};
    int32_t ov_spp[kMaxRounds] = {};
    DevBuf color[kMaxRounds];
    hipStream_t pstream[kMaxRounds] = {};
    hipEvent_t ev_up = nullptr, ev_done[kMaxRounds] = {}, ev_free[kMaxRounds] = {};

    /* wide rounds (tuning.round_frames = R > 1): a round renders R
     * consecutive samples as R copies of the job (copy k into region k of
     * the set's colour pool), with the overlapped rounds' streams and
     * events; sets sized for wide_frames samples, counters kept apart so
     * that resizing the sets keeps the statistics */
    struct WideSet {
        WaveBufs wb;
        DevBuf jobs, job_spp, sched, spill, color;
        std::vector<LxJob> hj;   /* sources of async copies */
        std::vector<int32_t> hs;
    } wide[kMaxRounds];
    int wide_sets = 0;
    uint32_t wide_frames = 0;
    uint32_t wide_grid[3] = {};
    DevBuf wide_counters; /* kMaxRounds LxCounters */
};

/* Records one round (begin, camera, bounces, end) on the stream. */
void lx_enqueue_round(lx_context* ctx, LxJob* d_jobs, const int32_t* d_spp, uint32_t njobs, LxCounters* d_cnt,
                      LxSched* d_sched, uint32_t path_cap, WaveBufs& wb, float* d_accum, uint32_t* d_spill,
                      int max_depth, uint32_t trace_grid, uint32_t shade_grid, uint32_t cam_grid,
                      int width /* BVH width of every referenced scene */,
                      const LxDevScene* uscene /* device scene shared by every job, or NULL */,
                      hipEvent_t* ev /* 2*max_depth+1 slots or NULL: [2b], [2b+1] bracket the trace and
                                        shade of bounce b < wavefront_bounces, [2w], [2w+1] the tail kernel,
                                        [2*max_depth] the round's end; other slots are not recorded */,
                      bool count = false, hipStream_t stream = nullptr /* NULL: the context's stream */,
                      bool dl_query = true /* some job may have -dmis: the tail kernel whose direct-light
                                              rays are occlusion queries */,
                      bool colour_fresh = false /* d_accum was filled with -0.0 for this round (one path
                                                   per pixel lands in it): the tail kernel stores colours */);
uint32_t lx_spill_words(uint32_t trace_grid);
/* a += b for the cumulative fields (ray, bounce and work counts, overflow flag) */
void lx_counters_add(LxCounters& a, const LxCounters& b);
/* bounces launched one kernel each before the tail kernel takes over
 * (tuning.wavefront_bounces; negative or above max_depth = all of them,
 * 0 = the tail kernel runs every bounce) */
int lx_wavefront_bounces(const lx_context* ctx, int max_depth);
/* Whether a round under the context's tuning starts with the tail kernel at
 * bounce 0 in slot order, so that it makes the camera paths itself: such a
 * round reads and writes no path buffers (WaveBufs stay unallocated). */
bool lx_tail_makes_camera(const lx_context* ctx, int max_depth);



/* ------------------------------------------------------ batched / served */
void lx_tuning_check(const lx_tuning& t); /* throws LX_ERR_INVALID */
void lx_validate_job(const lx_job& j);
/* Rejects an unknown material debug view or shading model. */
void lx_validate_views(int32_t debug_mode, int32_t shading_model);
/* Bounces a job traces: one for a material debug view, else max_depth or the
 * scene's trace depth. */
int lx_job_depth(int32_t max_depth, int trace_depth, int32_t debug_mode);
uint32_t lx_job_flags(int32_t debug_mode, int32_t shading_model);
LxJob lx_make_dev_job(const lx_job& j, uint32_t accum_off);
int lx_simulate_rounds(const std::vector<uint32_t>& npix, const std::vector<int32_t>& spp, uint32_t cap,
                       std::vector<int>* finish_round);
/* Paths one round of a batch may hold: the budget's share for path buffers
 * when the round keeps paths in memory, else every pixel of the batch. */
uint32_t lx_batch_path_cap(lx_context* ctx, const std::vector<uint32_t>& npix, bool dmis, uint64_t extra_bytes,
                           bool paths_in_memory = true);
#endif
