/*
 * lx_kernels.hip -- the wavefront bounce loop on gfx950.
 *
 * One round = one sample for every active job:
 *   k_begin_round   device-side job scheduling: active jobs, path offsets,
 *                   counters (no host round trip, so the round can be a
 *                   captured hipGraph)
 *   k_camera        generateRayFromCamera (src/kernels/generate_ray_from_camera.cu:8-39)
 *   bounces b < B (wavefront):
 *     k_trace       sampleDirectLight + raygen rays + hit attributes
 *                   (src/kernels/sample_direct_light.cu, src/optixshaders/optix_triangle.cu)
 *     k_shade       shadePath + RR + compaction (src/kernels/shade_path.cu:94-364) fused
 *                   with finalGather (src/kernels/image_kernels.cu:49-60): a path adds its
 *                   colour to the accumulation buffer when it stops
 *   bounces b >= B (tail):
 *     k_finish      each lane takes a live path and carries it through all its
 *                   remaining bounces (trace + shade in registers), refilling from
 *                   the queue when its path stops; the per-bounce launch barrier,
 *                   which makes every late bounce as long as its slowest path, is gone
 *   k_end_round     spp bookkeeping
 * Every kernel reads its path count from device memory and grid-strides over
 * fixed launch shapes.  Paths are compacted by copying their state to the next
 * buffer at a slot from one atomic per wave64 (ballot + mbcnt).  A path's
 * arithmetic does not depend on the bounce at which the tail kernel takes
 * over, so every choice of B gives the same bits.
 */
#include "lx_path_dev.h"
#include "lx_kernels.h"

#include <mutex>
#include <vector>

#define LX_WAVE 64

LXD uint32_t lx_lane() { return __lane_id(); }

LXD void lx_wave_add(unsigned long long* dst, uint32_t v) {
    /* wave64 reduction, one atomic per wave */
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, LX_WAVE);
    if (lx_lane() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

/* ------------------------------------------------------------ scheduling */
__global__ void k_begin_round(LxJob* __restrict__ jobs, const int32_t* __restrict__ job_spp, uint32_t njobs,
                              LxCounters* __restrict__ C, LxSched* __restrict__ sched, uint32_t path_cap) {
    /* single thread: jobs that still need samples take consecutive path
     * ranges until path_cap paths (or LX_SCHED_MAX jobs) are scheduled; the
     * rest wait for a later round */
    __shared__ uint32_t s_total;
    if (threadIdx.x == 0) {
        uint32_t off = 0, na = 0;
        for (uint32_t j = 0; j < njobs; j++) {
            LxJob& J = jobs[j];
            bool want = J.spp_done < job_spp[j];
            if (want && na < LX_SCHED_MAX && off + J.num_pix <= path_cap) {
                J.active = 1;
                J.path_off = off;
                uint32_t step = (J.flags & LX_JF_STEP_MASK) >> LX_JF_STEP_SHIFT;
                J.iter = J.iter + (step ? (int32_t)step : 1);
                sched->idx[na] = j;
                sched->off[na] = off;
                na++;
                off += J.num_pix;
            } else {
                J.active = 0;
            }
        }
        sched->n_active = na;
        s_total = off;
        C->tail_next = 0;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i <= LX_MAX_DEPTH; i += blockDim.x) C->count[i] = (i == 0) ? s_total : 0u;
}

__global__ void k_end_round(LxJob* __restrict__ jobs, uint32_t njobs) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs && jobs[j].active) jobs[j].spp_done += 1;
}

/* The pixel (x, y) of a job's q-th camera path.  tile <= 1: row-major.
 * Otherwise (tile a power of two) the image's whole tile x tile squares come
 * first, with pixels in Z (Morton) order inside a square -- an 8x8 square is
 * one wave64, so a wave's camera rays and their first hits are compact --
 * then the right strip and the bottom strip the squares leave, row-major.
 * Squares follow each other row by row (band 0), or, with band > 0, in
 * horizontal bands of `band` square rows walked column by column, so that
 * the waves running at one time cover a compact region of the image.  A
 * bijection of [0, W*H) onto the pixels for any W, H; shifts stand for the
 * divisions by the tile size, so at most two integer divisions remain. */
LXD uint32_t lx_even_bits(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}
LXD void lx_tile_xy(uint32_t q, uint32_t W, uint32_t H, uint32_t tile, uint32_t band, uint32_t& x, uint32_t& y) {
    if (tile <= 1u) {
        y = q / W;
        x = q - y * W;
        return;
    }
    const uint32_t lt = (uint32_t)__builtin_ctz(tile);
    const uint32_t tiles_x = W >> lt, tiles_y = H >> lt;
    const uint32_t Wt = tiles_x << lt, Ht = tiles_y << lt;
    const uint32_t nt = Wt * Ht;
    if (q < nt) {
        const uint32_t t = q >> (2u * lt), w = q & ((1u << (2u * lt)) - 1u);
        uint32_t tx, ty;
        if (band == 0u) {
            ty = t / tiles_x;
            tx = t - ty * tiles_x;
        } else {
            const uint32_t per_band = tiles_x * band; /* every band but the last is full */
            const uint32_t b = t / per_band;
            const uint32_t rows = min(band, tiles_y - b * band);
            const uint32_t u = t - b * per_band;
            tx = u / rows;
            ty = b * band + (u - tx * rows);
        }
        x = (tx << lt) + lx_even_bits(w);
        y = (ty << lt) + lx_even_bits(w >> 1);
        return;
    }
    uint32_t r = q - nt;
    const uint32_t rw = W - Wt, nr = rw * Ht; /* right strip: columns Wt.., rows ..Ht */
    if (r < nr) {
        y = r / rw;
        x = Wt + (r - y * rw);
        return;
    }
    r -= nr; /* bottom strip: rows Ht.., every column */
    const uint32_t yy = r / W;
    y = Ht + yy;
    x = r - yy * W;
}

/* generateRayFromCamera, src/kernels/generate_ray_from_camera.cu:8-39: the
 * pixel of job J's q-th camera path (returned) and its ray direction. */
LXD uint32_t lx_camera_ray(const LxJob& J, uint32_t q, uint32_t tile, uint32_t band, f3& dir) {
    const uint32_t W = (uint32_t)J.cam.width;
    uint32_t ux, uy;
    lx_tile_xy(q, W, (uint32_t)J.cam.height, tile, band, ux, uy);
    const uint32_t index = uy * W + ux;
    int x = (int)ux, y = (int)uy;
    LxRng rng = lx_make_rng(J.iter, (int)index, J.depth, J.seed);
    float x1 = rng.u01() - 0.5f;
    float x2 = rng.u01() - 0.5f;
    float pX = (float(x) + x1 + 0.5f) - (float)J.cam.width * 0.5f;
    float pY = (float(y) + x2 + 0.5f) - (float)J.cam.height * 0.5f;
    f3 view = mk3(J.cam.view[0], J.cam.view[1], J.cam.view[2]);
    f3 right = mk3(J.cam.right[0], J.cam.right[1], J.cam.right[2]);
    f3 up = mk3(J.cam.up[0], J.cam.up[1], J.cam.up[2]);
    dir = normalize((view - ((right * J.cam.pixel_length[0]) * pX)) - ((up * J.cam.pixel_length[1]) * pY));
    return index;
}

/* The scheduled job whose path range holds slot i: binary search over the
 * active jobs' first slots (off[0] = 0, increasing). */
LXD uint32_t lx_sched_job(const uint32_t* off, const uint32_t* idx, uint32_t nact, uint32_t i) {
    uint32_t lo = 0, hi = nact - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return idx[lo];
}

__global__ __launch_bounds__(256) void k_camera(const LxJob* __restrict__ jobs, const LxSched* __restrict__ sched,
                                                const LxCounters* __restrict__ C, LxPathBuf P, uint32_t tile,
                                                uint32_t band) {
    __shared__ uint32_t s_off[LX_SCHED_MAX];
    __shared__ uint32_t s_idx[LX_SCHED_MAX];
    uint32_t nact = sched->n_active;
    for (uint32_t k = threadIdx.x; k < nact; k += blockDim.x) {
        s_off[k] = sched->off[k];
        s_idx[k] = sched->idx[k];
    }
    __syncthreads();
    uint32_t total = C->count[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        uint32_t j = lx_sched_job(s_off, s_idx, nact, i);
        const LxJob& J = jobs[j];
        f3 dir;
        uint32_t index = lx_camera_ray(J, i - J.path_off, tile, band, dir);
        P.ox[i] = J.cam.position[0];
        P.oy[i] = J.cam.position[1];
        P.oz[i] = J.cam.position[2];
        P.dx[i] = dir.x;
        P.dy[i] = dir.y;
        P.dz[i] = dir.z;
        P.tx[i] = 1.0f;
        P.ty[i] = 1.0f;
        P.tz[i] = 1.0f;
        P.cx[i] = 0.0f;
        P.cy[i] = 0.0f;
        P.cz[i] = 0.0f;
        P.pdf[i] = -1.0f;
        P.pix[i] = index;
        P.job[i] = j;
    }
}

/* ------------------------------------------------------------ counters */
template <bool CNT>
LXD void lx_flush_counts(LxCounters* C, LxWorkCnt* W, const LxRayCount& rc, uint32_t nsh, const LxWork& wc,
                         const LxWork& wa) {
    lx_wave_add(&C->rad_rays, rc.rad);
    lx_wave_add(&C->dl_rays, rc.dl);
    lx_wave_add(&C->env_rays, rc.env);
    lx_wave_add(&C->bounces, nsh);
    if (CNT) {
        lx_wave_add(&W->closest_rays, rc.rad + rc.dl);
        lx_wave_add(&W->any_rays, rc.env);
        lx_wave_add(&W->closest_nodes, wc.nodes);
        lx_wave_add(&W->closest_tris, wc.tris);
        lx_wave_add(&W->closest_insts, wc.insts);
        lx_wave_add(&W->any_nodes, wa.nodes);
        lx_wave_add(&W->any_tris, wa.tris);
        lx_wave_add(&W->any_insts, wa.insts);
    }
}

/* ------------------------------------------------------------ trace */
/* record sink writing straight to the SoA record buffers */
struct LxRecStore {
    const LxRecBuf& R;
    uint32_t i;
    LXD void hit(float t, f3 N, f2 uv, int mat, bool tan, f3 Tg) {
        R.t[i] = t;
        R.nx[i] = N.x;
        R.ny[i] = N.y;
        R.nz[i] = N.z;
        R.u[i] = uv.x;
        R.v[i] = uv.y;
        R.mat[i] = mat;
        if (tan) {
            R.tgx[i] = Tg.x;
            R.tgy[i] = Tg.y;
            R.tgz[i] = Tg.z;
        }
    }
    LXD void miss() { R.t[i] = -1.0f; }
    LXD void dl(float t, f3 N, f2 uv, int mat) {
        R.dl_t[i] = t;
        R.dl_nx[i] = N.x;
        R.dl_ny[i] = N.y;
        R.dl_nz[i] = N.z;
        R.dl_u[i] = uv.x;
        R.dl_v[i] = uv.y;
        R.dl_mat[i] = mat;
    }
    LXD void dl_miss() { R.dl_t[i] = -1.0f; }
    LXD void light(f3 lp) {
        R.lpx[i] = lp.x;
        R.lpy[i] = lp.y;
        R.lpz[i] = lp.z;
    }
    LXD void env_dir(f3 ed) {
        R.edx[i] = ed.x;
        R.edy[i] = ed.y;
        R.edz[i] = ed.z;
    }
    LXD void env(float env_t) { R.env_t[i] = env_t; }
};

/* CNT = true is the counting replay: identical work plus per-ray node,
 * triangle and instance visit counts for the algorithmic-bytes model. */
template <bool CNT, int W>
__global__ __launch_bounds__(LX_TRACE_BLOCK) void k_trace(const LxJob* __restrict__ jobs, LxPathBuf P, LxRecBuf R,
                                                           LxCounters* __restrict__ C, int bounce,
                                                           uint32_t* __restrict__ spill,
                                                           const uint32_t* __restrict__ order) {
    __shared__ uint32_t s_stack[LX_STACK * LX_TRACE_BLOCK];
    uint32_t n = C->count[bounce];
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gsz = gridDim.x * blockDim.x;
    LxStack st;
    st.lds = (LX_LDS uint32_t*)(s_stack + threadIdx.x);
    st.stride = LX_TRACE_BLOCK;
    st.spill = (LX_GLOBAL uint32_t*)(spill + gtid);
    st.sstride = gsz;
    LxRayCount rc = {0, 0, 0};
    LxWork wc = {0, 0, 0}, wa = {0, 0, 0};
    for (uint32_t q = gtid; q < n; q += gsz) {
        uint32_t i = order ? order[q] : q; /* slot of the q-th queued path */
        const LxJob& J = jobs[P.job[i]];
        const LxDevScene& S = *J.scene;
        uint32_t pix = P.pix[i] & 0x7fffffffu;
        LxRecStore out{R, i};
        lx_trace_path<CNT, W>(J, S, mk3(P.ox[i], P.oy[i], P.oz[i]), mk3(P.dx[i], P.dy[i], P.dz[i]), pix, bounce, st,
                           &C->stack_overflow, out, rc, &wc, &wa);
    }
    lx_flush_counts<CNT>(C, &C->work[0], rc, 0, wc, wa);
}

/* ------------------------------------------------------------ shade */
LXD void lx_load_rec(const LxJob& J, const LxDevScene& S, const LxRecBuf& R, uint32_t i, LxRec& r) {
    r.t = R.t[i];
    r.env_t = R.env_t[i];
    if (r.t > 0.0f) {
        r.N = mk3(R.nx[i], R.ny[i], R.nz[i]);
        r.uv = mk2(R.u[i], R.v[i]);
        r.mat = R.mat[i];
        if (S.mats[r.mat].normal_tex >= 0) r.Tg = mk3(R.tgx[i], R.tgy[i], R.tgz[i]);
        if (J.emis) r.ed = mk3(R.edx[i], R.edy[i], R.edz[i]);
    }
    if (J.dmis) {
        r.dl_t = R.dl_t[i];
        if (r.dl_t > 0.0f) {
            r.dlN = mk3(R.dl_nx[i], R.dl_ny[i], R.dl_nz[i]);
            r.dluv = mk2(R.dl_u[i], R.dl_v[i]);
            r.dlmat = R.dl_mat[i];
        }
        r.lp = mk3(R.lpx[i], R.lpy[i], R.lpz[i]);
    }
}

/* store: the buffer was filled with -0.0 for this round and receives one
 * path per pixel, so the add -0 + c would give c's bits: the colour is
 * written without reading the buffer */
LXD void lx_accumulate(float* __restrict__ accum, const LxJob& J, uint32_t pix, f3 col, bool store = false) {
    /* finalGather: the path's colour is final */
    LX_GLOBAL float* a = (LX_GLOBAL float*)(accum + 3u * (size_t)(J.accum_off + pix));
    if (store) {
        a[0] = col.x;
        a[1] = col.y;
        a[2] = col.z;
    } else {
        a[0] += col.x;
        a[1] += col.y;
        a[2] += col.z;
    }
}

__global__ __launch_bounds__(LX_SHADE_BLOCK) void k_shade(const LxJob* __restrict__ jobs, LxPathBuf P, LxPathBuf Q,
                                                           LxRecBuf R, LxCounters* __restrict__ C, int bounce,
                                                           float* __restrict__ accum,
                                                           const uint32_t* __restrict__ order) {
    uint32_t n = C->count[bounce];
    uint32_t wave_base = blockIdx.x * blockDim.x + (threadIdx.x & ~(LX_WAVE - 1));
    uint32_t gsz = gridDim.x * blockDim.x;
    uint32_t lane = lx_lane();
    int depth = bounce + 1;
    uint32_t nsh = 0;
    for (uint32_t base = wave_base; base < n; base += gsz) {
        uint32_t q = base + lane;
        bool active = q < n;
        uint32_t i = active && order ? order[q] : q; /* slot of the q-th queued path */
        bool survive = false;
        f3 o, d, thr, col;
        float pdf = 0.0f;
        uint32_t pixw = 0, job = 0;
        bool spec = false;
        if (active) {
            job = P.job[i];
            const LxJob& J = jobs[job];
            const LxDevScene& S = *J.scene;
            pixw = P.pix[i];
            spec = (pixw >> 31) != 0u;
            uint32_t pix = pixw & 0x7fffffffu;
            o = mk3(P.ox[i], P.oy[i], P.oz[i]);
            d = mk3(P.dx[i], P.dy[i], P.dz[i]);
            thr = mk3(P.tx[i], P.ty[i], P.tz[i]);
            col = mk3(P.cx[i], P.cy[i], P.cz[i]);
            pdf = P.pdf[i];
            LxRec r;
            lx_load_rec(J, S, R, i, r);
            bool compact = depth < J.depth;
            survive = lx_shade_path(J, S, r, depth, compact, o, d, thr, col, pdf, spec, pix);
            if (!survive) lx_accumulate(accum, J, pix, col);
            pixw = pix | (spec ? 0x80000000u : 0u);
            nsh++;
        }
        uint64_t mask = __ballot(survive);
        uint32_t cnt = (uint32_t)__popcll(mask);
        uint32_t obase = 0;
        if (cnt) {
            if (lane == 0) obase = atomicAdd(&C->count[bounce + 1], cnt);
            obase = __shfl(obase, 0, LX_WAVE);
        }
        if (survive) {
            uint32_t slot = obase + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            Q.ox[slot] = o.x;
            Q.oy[slot] = o.y;
            Q.oz[slot] = o.z;
            Q.dx[slot] = d.x;
            Q.dy[slot] = d.y;
            Q.dz[slot] = d.z;
            Q.tx[slot] = thr.x;
            Q.ty[slot] = thr.y;
            Q.tz[slot] = thr.z;
            Q.cx[slot] = col.x;
            Q.cy[slot] = col.y;
            Q.cz[slot] = col.z;
            Q.pdf[slot] = pdf;
            Q.pix[slot] = pixw;
            Q.job[slot] = job;
        }
    }
    LxRayCount rc = {0, 0, 0};
    LxWork z = {0, 0, 0};
    lx_flush_counts<false>(C, nullptr, rc, nsh, z, z);
}

/* ------------------------------------------------------------ tail */
/* One lane's column of tail-kernel path state in LDS (stride BLOCK floats):
 * throughput, colour, last pdf, the pending environment colour term and the
 * pending environment shadow ray (origin, direction). */
#define LX_TAIL_LDS_STATE 16
template <int BLOCK>
struct LxTailColumn {
    LX_LDS float* p;
    LXD float get(int k) const { return p[k * BLOCK]; }
    LXD void set(int k, float v) const { p[k * BLOCK] = v; }
    LXD f3 thr() const { return mk3(get(0), get(1), get(2)); }
    LXD f3 col() const { return mk3(get(3), get(4), get(5)); }
    LXD float pdf() const { return get(6); }
    LXD f3 pend_col() const { return mk3(get(7), get(8), get(9)); }
    LXD void put_thr(f3 v) const { set(0, v.x); set(1, v.y); set(2, v.z); }
    LXD void put_col(f3 v) const { set(3, v.x); set(4, v.y); set(5, v.z); }
    LXD void put_pdf(float v) const { set(6, v); }
    LXD void put_pend_col(f3 v) const { set(7, v.x); set(8, v.y); set(9, v.z); }
    LXD f3 pend_o() const { return mk3(get(10), get(11), get(12)); }
    LXD f3 pend_d() const { return mk3(get(13), get(14), get(15)); }
    LXD void put_pend_ray(f3 o, f3 d) const {
        set(10, o.x); set(11, o.y); set(12, o.z);
        set(13, d.x); set(14, d.y); set(15, d.z);
    }
};

/* BLOCK threads per workgroup; OCC > 0 asks the register allocator for at
 * least OCC waves per SIMD (spilling to scratch if it must). */
/* UNI 1: every job of the launch renders the scene at `uscene`, which then
 * lives in scalar registers instead of one copy per lane; UNI 2: in
 * addition the launch has a single job (jobs[0]), also kept in scalars. */
/* DLQ: direct-light shadow rays answered by lx_trace_dl (the kernel for jobs
 * with -dmis; without it the kernel carries no occlusion-query code). */
template <bool CNT, int BLOCK, int OCC, int W, int UNI, bool DLQ>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1)))
void k_finish(const LxJob* __restrict__ jobs, LxPathBuf P, LxCounters* __restrict__ C, int bounce0,
              uint32_t* __restrict__ spill, float* __restrict__ accum, uint32_t chunk,
              const LxDevScene* __restrict__ uscene, const uint32_t* __restrict__ order,
              const LxSched* __restrict__ cam_sched, uint32_t cam_tile, uint32_t cam_band, uint32_t acc_store) {
    __shared__ uint32_t s_stack[LX_STACK * BLOCK];
    /* per-lane path state that only shading reads: it waits in an LDS
     * column while the lane traverses, instead of holding registers */
    __shared__ float s_path[LX_TAIL_LDS_STATE * BLOCK];
    LxTailColumn<BLOCK> ps{(LX_LDS float*)(s_path + threadIdx.x)};
    uint32_t n = C->count[bounce0];
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gsz = gridDim.x * blockDim.x;
    uint32_t lane = lx_lane();
    LxStack st;
    st.lds = (LX_LDS uint32_t*)(s_stack + threadIdx.x);
    st.stride = BLOCK;
    st.spill = (LX_GLOBAL uint32_t*)(spill + gtid);
    st.sstride = gsz;
    LxRayCount rc = {0, 0, 0};
    LxWork wc = {0, 0, 0}, wa = {0, 0, 0};
    uint32_t nsh = 0;
    bool have = false, exhausted = false;
    uint32_t job = 0, pix = 0;
    int b = bounce0;
    f3 o = mk3(0, 0, 0), d = mk3(0, 0, 0);
    bool spec = false;
    /* the environment shadow ray of the last shaded bounce, traced together
     * with the next radiance ray; the ray and its colour term wait in the
     * LDS column */
    bool pend = false, pend_add = false, flush = false;
    while (true) {
        /* refill idle lanes from the live-path queue, one atomic per wave,
         * once at least `chunk` lanes are idle (or all of them are) */
        uint64_t idle = __ballot(!have && !flush);
        if (!exhausted && idle && ((uint32_t)__popcll(idle) >= chunk || idle == ~0ull)) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&C->tail_next, (uint32_t)__popcll(idle));
            base = __shfl(base, 0, LX_WAVE);
            if (base >= n) exhausted = true;
            if (!have && !flush) {
                uint32_t my = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (my < n) {
                    if (cam_sched) {
                        /* camera paths made here (the camera kernel's ray for
                         * slot my) instead of loaded from the path buffer */
                        job = UNI == 2 ? 0u
                                       : lx_sched_job(cam_sched->off, cam_sched->idx, cam_sched->n_active, my);
                        const LxJob& Jc = UNI == 2 ? jobs[0] : jobs[job];
                        pix = lx_camera_ray(Jc, my - Jc.path_off, cam_tile, cam_band, d);
                        o = mk3(Jc.cam.position[0], Jc.cam.position[1], Jc.cam.position[2]);
                        spec = false;
                        ps.put_thr(mk3(1.0f, 1.0f, 1.0f));
                        ps.put_col(mk3(0.0f, 0.0f, 0.0f));
                        ps.put_pdf(-1.0f);
                    } else {
                        if (order) my = lx_gptr(order)[my]; /* slot of the my-th queued path */
                        job = lx_gptr(P.job)[my];
                        uint32_t pw = lx_gptr(P.pix)[my];
                        spec = (pw >> 31) != 0u;
                        pix = pw & 0x7fffffffu;
                        o = mk3(lx_gptr(P.ox)[my], lx_gptr(P.oy)[my], lx_gptr(P.oz)[my]);
                        d = mk3(lx_gptr(P.dx)[my], lx_gptr(P.dy)[my], lx_gptr(P.dz)[my]);
                        ps.put_thr(mk3(lx_gptr(P.tx)[my], lx_gptr(P.ty)[my], lx_gptr(P.tz)[my]));
                        ps.put_col(mk3(lx_gptr(P.cx)[my], lx_gptr(P.cy)[my], lx_gptr(P.cz)[my]));
                        ps.put_pdf(lx_gptr(P.pdf)[my]);
                    }
                    b = bounce0;
                    have = true;
                }
            }
        }
        if (__ballot(have || flush) == 0ull) break;
        if (have || flush) {
            const LxJob& J = UNI == 2 ? jobs[0] : jobs[job];
            const LxDevScene& S = UNI ? *uscene : *J.scene;
            LxRec r{};
            LxRecSink sink{r};
            bool occluded = false;
            lx_trace_path_d<CNT, W, DLQ>(J, S, o, d, pix, b, st, &C->stack_overflow, pend, ps.pend_o(), ps.pend_d(),
                                         have, occluded, sink, rc, &wc, &wa);
            f3 col = ps.col();
            if (pend) {
                if (!occluded && pend_add) col = col + ps.pend_col();
                pend = false;
            }
            if (flush) {
                lx_accumulate(accum, J, pix, col, acc_store != 0u);
                flush = false;
            } else {
                int depth = b + 1;
                f3 thr = ps.thr(), pend_col = mk3(0, 0, 0);
                float pdf = ps.pdf();
                bool cont = lx_shade_path<true>(J, S, r, depth, depth < J.depth, o, d, thr, col, pdf, spec, pix,
                                                &pend_add, &pend_col);
                nsh++;
                if (J.emis && r.t > 0.0f) {
                    /* this bounce's environment shadow ray, traced next time */
                    pend = true;
                    ps.put_pend_ray(r.so, r.ed);
                }
                if (cont) {
                    b++;
                    ps.put_thr(thr);
                    ps.put_pdf(pdf);
                    ps.put_col(col);
                    ps.put_pend_col(pend_col);
                } else {
                    have = false;
                    if (pend) {
                        flush = true;
                        ps.put_col(col);
                        ps.put_pend_col(pend_col);
                    } else {
                        lx_accumulate(accum, J, pix, col, acc_store != 0u);
                    }
                }
            }
        }
    }
    lx_flush_counts<CNT>(C, &C->work[1], rc, nsh, wc, wa);
    lx_wave_add(&C->tail_bounces, nsh);
}

/* ------------------------------------------------------------ sort stages */
/* Queue reordering for coherence (the reference sorts path indices with
 * thrust::sort_by_key: Morton codes of the rays, src/pathtrace.cu:157-185,
 * src/kernels/intersection_precompute.cu; material ids, pathtrace.cu:266-276).
 * Keys here have 12 bits -- a 16^3 Morton cell of a point one reach along the
 * ray inside the scene box, or the hit material (misses in one more bucket)
 * -- so a bucket sort (histogram, scan, scatter) gives the permutation in
 * three small launches with no temporary storage.  The order inside a bucket
 * is the atomics' order; it only decides which lane takes which path. */
#define LX_SORT_BINS 4097u

LXD uint32_t lx_cell_key(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t k = 0;
    for (int b = 3; b >= 0; b--) k = (k << 3) | (((x >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((z >> b) & 1u);
    return k;
}
LXD uint32_t lx_cell(float p, float lo, float inv) {
    float c = (p - lo) * inv * 16.0f;
    c = fminf(fmaxf(c, 0.0f), 15.0f);
    return (uint32_t)c;
}

/* Lanes of a wave that share a key add to its counter with one atomic: the
 * keys of neighbouring paths mostly agree, and same-address atomics from one
 * wave would otherwise serialise.  Returns this lane's offset from the
 * counter's old value. */
template <class Add>
LXD uint32_t lx_wave_bucket_add(uint32_t key, bool active, Add add) {
    uint32_t lane = lx_lane();
    uint32_t mine = 0;
    uint64_t todo = __ballot(active);
    while (todo) {
        uint32_t leader = (uint32_t)__builtin_ctzll(todo);
        uint32_t lkey = __shfl(key, (int)leader, LX_WAVE);
        uint64_t same = __ballot(active && key == lkey) & todo;
        uint32_t base = 0;
        if (lane == leader) base = add(lkey, (uint32_t)__popcll(same));
        base = __shfl(base, (int)leader, LX_WAVE);
        if ((same >> lane) & 1ull) mine = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return mine;
}

__global__ __launch_bounds__(256) void k_sort_keys(const LxJob* __restrict__ jobs, LxPathBuf P, LxRecBuf R,
                                                   const LxCounters* __restrict__ C, int bounce, int mode,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ hist) {
    __shared__ uint32_t lh[LX_SORT_BINS];
    for (uint32_t k = threadIdx.x; k < LX_SORT_BINS; k += blockDim.x) lh[k] = 0u;
    __syncthreads();
    uint32_t n = C->count[bounce];
    uint32_t gsz = gridDim.x * blockDim.x;
    /* whole waves iterate together so the wave-wide aggregation sees every lane */
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~(LX_WAVE - 1u)); base < n; base += gsz) {
        uint32_t i = base + lx_lane();
        bool active = i < n;
        uint32_t key = 0;
        if (!active) {
        } else if (mode == LX_SORT_MATERIALS) {
            key = R.t[i] > 0.0f ? min((uint32_t)R.mat[i], 4095u) : 4096u;
        } else {
            const LxDevScene& S = *jobs[P.job[i]].scene;
            float r = S.sort_reach;
            key = lx_cell_key(lx_cell(P.ox[i] + P.dx[i] * r, S.box_lo[0], S.box_inv[0]),
                              lx_cell(P.oy[i] + P.dy[i] * r, S.box_lo[1], S.box_inv[1]),
                              lx_cell(P.oz[i] + P.dz[i] * r, S.box_lo[2], S.box_inv[2]));
        }
        if (active) keys[i] = key;
        lx_wave_bucket_add(key, active, [&](uint32_t k, uint32_t c) { return atomicAdd(&lh[k], c); });
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < LX_SORT_BINS; k += blockDim.x)
        if (lh[k]) atomicAdd(&hist[k], lh[k]);
}

/* exclusive prefix sum of the histogram, in place (one workgroup) */
__global__ __launch_bounds__(256) void k_sort_scan(uint32_t* __restrict__ hist) {
    __shared__ uint32_t part[256];
    const uint32_t per = (LX_SORT_BINS + 255u) / 256u;
    uint32_t b0 = threadIdx.x * per, b1 = min(b0 + per, LX_SORT_BINS);
    uint32_t sum = 0;
    for (uint32_t k = b0; k < b1; k++) sum += hist[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 256u; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t k = b0; k < b1; k++) {
        uint32_t h = hist[k];
        hist[k] = run;
        run += h;
    }
}

/* Each workgroup places a chunk of 16 keys per thread: ranks within the chunk
 * from an LDS histogram, then one global atomic per distinct key of the
 * chunk reserves its range -- neighbouring paths mostly share keys, and
 * per-wave atomics on the same few counters would serialise. */
#define LX_SORT_ITEMS 16
__global__ __launch_bounds__(256) void k_sort_scatter(const LxCounters* __restrict__ C, int bounce,
                                                      const uint32_t* __restrict__ keys, uint32_t* __restrict__ offs,
                                                      uint32_t* __restrict__ order) {
    __shared__ uint32_t lcount[LX_SORT_BINS];
    __shared__ uint32_t lbase[LX_SORT_BINS];
    const uint32_t n = C->count[bounce];
    const uint32_t chunk = 256u * LX_SORT_ITEMS;
    for (uint32_t c0 = blockIdx.x * chunk; c0 < n; c0 += gridDim.x * chunk) {
        for (uint32_t b = threadIdx.x; b < LX_SORT_BINS; b += 256u) lcount[b] = 0u;
        __syncthreads();
        uint32_t key[LX_SORT_ITEMS], rank[LX_SORT_ITEMS];
#pragma unroll
        for (int j = 0; j < LX_SORT_ITEMS; j++) {
            uint32_t i = c0 + (uint32_t)j * 256u + threadIdx.x;
            bool active = i < n;
            key[j] = active ? keys[i] : 0u;
            rank[j] = lx_wave_bucket_add(key[j], active,
                                         [&](uint32_t k, uint32_t c) { return atomicAdd(&lcount[k], c); });
        }
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < LX_SORT_BINS; b += 256u)
            if (lcount[b]) lbase[b] = atomicAdd(&offs[b], lcount[b]);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < LX_SORT_ITEMS; j++) {
            uint32_t i = c0 + (uint32_t)j * 256u + threadIdx.x;
            if (i < n) order[lbase[key[j]] + rank[j]] = i;
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------ test hooks */
template <int W>
__global__ __launch_bounds__(LX_TRACE_BLOCK) void k_trace_rays(const LxDevScene* __restrict__ Sp,
                                                                const lx_ray* __restrict__ rays, uint32_t n,
                                                                int any, int alpha, uint32_t iteration,
                                                                lx_hit* __restrict__ hits, uint8_t* __restrict__ occ,
                                                                uint32_t* __restrict__ spill,
                                                                uint32_t* __restrict__ overflow) {
    __shared__ uint32_t s_stack[LX_STACK * LX_TRACE_BLOCK];
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gsz = gridDim.x * blockDim.x;
    LxStack st;
    st.lds = (LX_LDS uint32_t*)(s_stack + threadIdx.x);
    st.stride = LX_TRACE_BLOCK;
    st.spill = (LX_GLOBAL uint32_t*)(spill + gtid);
    st.sstride = gsz;
    LxDevScene S = *Sp;
    uint32_t amask = alpha ? LX_INST_ALPHA : 0u;
    for (uint32_t i = gtid; i < n; i += gsz) {
        lx_ray r = rays[i];
        f3 o = mk3(r.origin[0], r.origin[1], r.origin[2]);
        f3 d = mk3(r.dir[0], r.dir[1], r.dir[2]);
        if (any == 1) {
            LxHitD h = lx_trace<true, false, W>(S, o, d, r.tmax, i, iteration, amask, st, overflow);
            occ[i] = h.inst != 0xffffffffu ? 1 : 0;
        } else {
            /* any == 2: the tail kernel's direct-light query (alpha rules on) */
            LxHitD h = any == 2 ? lx_trace_dl<false, W>(S, o, d, r.tmax, i, iteration, st, overflow, nullptr)
                                : lx_trace<false, false, W>(S, o, d, r.tmax, i, iteration, amask, st, overflow);
            lx_hit o2;
            o2.t = h.inst == 0xffffffffu ? -1.0f : h.t;
            o2.instance = h.inst;
            o2.prim = h.prim;
            o2.u = h.u;
            o2.v = h.v;
            hits[i] = o2;
        }
    }
}

/* sendImageToPBO, src/kernels/image_kernels.cu:6-46: average, tonemap,
 * gamma 0.45, RGBA8.  Tonemaps from src/tonemapping.cu:3-30. */
LXD f3 lx_aces(f3 x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, dd = 0.59f, e = 0.14f;
    f3 num = x * ((a * x) + mk3(b, b, b));
    f3 den = (x * ((c * x) + mk3(dd, dd, dd))) + mk3(e, e, e);
    f3 r = mk3(num.x / den.x, num.y / den.y, num.z / den.z);
    return mk3(lxm_clamp(r.x, 0.0f, 1.0f), lxm_clamp(r.y, 0.0f, 1.0f), lxm_clamp(r.z, 0.0f, 1.0f));
}
LXD f3 lx_agx(f3 v) {
    v = mk3(lxm_max(v.x, 1e-6f), lxm_max(v.y, 1e-6f), lxm_max(v.z, 1e-6f));
    v.x = (lxm_log2(v.x) + 10.0f) / 16.0f;
    v.y = (lxm_log2(v.y) + 10.0f) / 16.0f;
    v.z = (lxm_log2(v.z) + 10.0f) / 16.0f;
    v = mk3(lxm_clamp(v.x, 0.0f, 1.0f), lxm_clamp(v.y, 0.0f, 1.0f), lxm_clamp(v.z, 0.0f, 1.0f));
    f3 x2 = v * v;
    f3 x4 = x2 * x2;
    v = (((((15.5f * x4) * v - 40.14f * x4) + (31.96f * x2) * v) - 6.86f * x2) + 0.429f * v) + mk3(0.0322f, 0.0322f, 0.0322f);
    f3 r;
    r.x = v.x * 1.1968790f - v.y * 0.0980208f - v.z * 0.0990297f;
    r.y = -v.x * 0.0528968f + v.y * 1.1519031f - v.z * 0.0989611f;
    r.z = -v.x * 0.0529716f - v.y * 0.0980434f + v.z * 1.1510736f;
    return mk3(lxm_clamp(r.x, 0.0f, 1.0f), lxm_clamp(r.y, 0.0f, 1.0f), lxm_clamp(r.z, 0.0f, 1.0f));
}

__global__ void k_tonemap(const float* __restrict__ accum, uint32_t npix, float spp, int mode, uint8_t* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    f3 p = mk3(accum[3 * i], accum[3 * i + 1], accum[3 * i + 2]);
    float inv = 1.0f / spp;
    p = p * inv;
    if (mode >= 0) { /* mode < 0: a material debug view, shown as is (image_kernels.cu:20) */
        if (mode == LX_TONEMAP_REINHARD) p = mk3(p.x / (p.x + 1.0f), p.y / (p.y + 1.0f), p.z / (p.z + 1.0f));
        else if (mode == LX_TONEMAP_AGX) p = lx_agx(p);
        else p = lx_aces(p);
        p = mk3(lxm_pow(p.x, 0.45f), lxm_pow(p.y, 0.45f), lxm_pow(p.z, 0.45f));
    }
    int r = (int)((double)p.x * 255.0), g = (int)((double)p.y * 255.0), b = (int)((double)p.z * 255.0);
    r = r < 0 ? 0 : (r > 255 ? 255 : r);
    g = g < 0 ? 0 : (g > 255 ? 255 : g);
    b = b < 0 ? 0 : (b > 255 ? 255 : b);
    out[4 * i + 0] = (uint8_t)r;
    out[4 * i + 1] = (uint8_t)g;
    out[4 * i + 2] = (uint8_t)b;
    out[4 * i + 3] = 0;
}

__global__ void k_math_eval(int fn, const float* __restrict__ x, const float* __restrict__ y, float* __restrict__ out,
                            uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = x[i], b = y ? y[i] : 0.0f, r = 0.0f;
    switch (fn) {
    case 0: r = lxm_sin(a); break;
    case 1: r = lxm_cos(a); break;
    case 2: r = lxm_atan(a); break;
    case 3: r = lxm_atan2(a, b); break;
    case 4: r = lxm_acos(a); break;
    case 5: r = lxm_log2(a); break;
    case 6: r = lxm_exp2(a); break;
    case 7: r = lxm_pow(a, b); break;
    case 8: r = lxm_sqrt(a); break;
    case 9: r = a / b; break;
    default: r = 0.0f;
    }
    out[i] = r;
}

__global__ void k_copy16(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

/* Pixel ranges of an RGB pool: range y is (first pixel r[2y], count r[2y+1]);
 * fill sets its floats' bits to v, add does dst += src over them. */
__global__ void k_fill_ranges(uint32_t* __restrict__ p, const uint32_t* __restrict__ r, uint32_t v) {
    size_t base = 3ull * r[2 * blockIdx.y], n = 3ull * r[2 * blockIdx.y + 1];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[base + i] = v;
}
__global__ void k_add_ranges(float* __restrict__ dst, const float* __restrict__ src, const uint32_t* __restrict__ r) {
    size_t base = 3ull * r[2 * blockIdx.y], n = 3ull * r[2 * blockIdx.y + 1];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[base + i] = dst[base + i] + src[base + i];
}

/* dst += src, element by element (one round's colours into the accumulation) */
__global__ void k_add_f32(float* __restrict__ dst, const float* __restrict__ src, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = dst[i] + src[i];
}

/* The regions of a wide round's colour pool added in sample order: the same
 * float sums as one add per region. */
__global__ void k_add_regions(float* __restrict__ dst, const float* __restrict__ src, size_t n, uint32_t regions) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float a = dst[i];
        for (uint32_t k = 0; k < regions; k++) a = a + src[(size_t)k * n + i];
        dst[i] = a;
    }
}

/* ------------------------------------------------------------ launchers */
hipError_t lx_launch_begin_round(LxJob* jobs, const int32_t* job_spp, uint32_t njobs, LxCounters* C, LxSched* sched,
                                 uint32_t cap, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_round, dim3(1), dim3(64), 0, s, jobs, job_spp, njobs, C, sched, cap);
    return hipGetLastError();
}
hipError_t lx_launch_end_round(LxJob* jobs, uint32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_end_round, dim3((njobs + 255) / 256), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}
hipError_t lx_launch_camera(const LxJob* jobs, const LxSched* sched, const LxCounters* C, const LxPathBuf& P,
                            uint32_t grid, hipStream_t s, uint32_t tile, uint32_t band) {
    hipLaunchKernelGGL(k_camera, dim3(grid), dim3(256), 0, s, jobs, sched, C, P, tile, band);
    return hipGetLastError();
}
hipError_t lx_launch_trace(const LxJob* jobs, const LxPathBuf& P, const LxRecBuf& R, LxCounters* C, int bounce,
                           uint32_t* spill, uint32_t grid, int width, hipStream_t s, bool count,
                           const uint32_t* order) {
    dim3 g(grid), b(LX_TRACE_BLOCK);
    if (width == 8) {
        if (count) hipLaunchKernelGGL((k_trace<true, 8>), g, b, 0, s, jobs, P, R, C, bounce, spill, order);
        else hipLaunchKernelGGL((k_trace<false, 8>), g, b, 0, s, jobs, P, R, C, bounce, spill, order);
    } else {
        if (count) hipLaunchKernelGGL((k_trace<true, 4>), g, b, 0, s, jobs, P, R, C, bounce, spill, order);
        else hipLaunchKernelGGL((k_trace<false, 4>), g, b, 0, s, jobs, P, R, C, bounce, spill, order);
    }
    return hipGetLastError();
}
hipError_t lx_launch_shade(const LxJob* jobs, const LxPathBuf& P, const LxPathBuf& Q, const LxRecBuf& R, LxCounters* C,
                           int bounce, float* accum, uint32_t grid, hipStream_t s, const uint32_t* order) {
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(LX_SHADE_BLOCK), 0, s, jobs, P, Q, R, C, bounce, accum, order);
    return hipGetLastError();
}
hipError_t lx_launch_sort(const LxJob* jobs, const LxPathBuf& P, const LxRecBuf& R, const LxCounters* C, int bounce,
                          int mode, uint32_t* keys, uint32_t* hist, uint32_t* order, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3((LX_SORT_BINS + 255) / 256), dim3(256), 0, s, hist, (size_t)LX_SORT_BINS, 0u);
    hipLaunchKernelGGL(k_sort_keys, dim3(grid), dim3(256), 0, s, jobs, P, R, C, bounce, mode, keys, hist);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(256), 0, s, hist);
    hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(256), 0, s, C, bounce, keys, hist, order);
    return hipGetLastError();
}
struct LxFinishArgs {
    const LxJob* jobs;
    const LxPathBuf* P;
    LxCounters* C;
    int bounce0;
    uint32_t* spill;
    float* accum;
    uint32_t chunk;
    const LxDevScene* us;
    const uint32_t* order;
    const LxSched* cam;
    uint32_t tile, band;
    bool dlq;
    bool store; /* colour buffer filled with -0.0 for this round */
};
template <int BLOCK, int OCC, int W, int UNI>
static void lx_finish_at(const LxFinishArgs& a, uint32_t grid, hipStream_t s) {
    if (OCC > 0 && a.dlq)
        hipLaunchKernelGGL((k_finish<false, BLOCK, OCC, W, UNI, true>), dim3(grid), dim3(BLOCK), 0, s, a.jobs, *a.P,
                           a.C, a.bounce0, a.spill, a.accum, a.chunk, a.us, a.order, a.cam, a.tile, a.band,
                           (uint32_t)a.store);
    else
        hipLaunchKernelGGL((k_finish<false, BLOCK, OCC, W, UNI, false>), dim3(grid), dim3(BLOCK), 0, s, a.jobs, *a.P,
                           a.C, a.bounce0, a.spill, a.accum, a.chunk, a.us, a.order, a.cam, a.tile, a.band,
                           (uint32_t)a.store);
}
template <int W, int UNI>
static void lx_finish_w(const LxFinishArgs& a, uint32_t threads, int block, int occ, hipStream_t s) {
    if (occ >= 4) {
        if (block == 256 && threads >= 256) lx_finish_at<256, 4, W, UNI>(a, threads / 256, s);
        else lx_finish_at<128, 4, W, UNI>(a, threads / 128, s);
    } else {
        lx_finish_at<128, 0, W, UNI>(a, threads / 128, s);
    }
}
hipError_t lx_launch_finish(const LxJob* jobs, const LxPathBuf& P, LxCounters* C, int bounce0, uint32_t* spill,
                            float* accum, uint32_t threads, int block, int occ, int width, int chunk,
                            const LxDevScene* uscene, bool single_job, hipStream_t s, bool count,
                            const uint32_t* order, const LxSched* cam_sched, uint32_t tile, uint32_t band,
                            bool dl_query, bool colour_fresh) {
    /* threads = grid * block lanes, the same for every shape, so one spill
     * allocation serves them all */
    uint32_t c = chunk < 1 ? 1u : (chunk > 64 ? 64u : (uint32_t)chunk);
    LxFinishArgs a{jobs, &P, C, bounce0, spill, accum, c, uscene, order, cam_sched, tile, band, dl_query, colour_fresh};
    const uint32_t st = colour_fresh ? 1u : 0u;
    dim3 g(threads / 128), b(128);
    if (count) {
        /* the counting replay: same paths and rays, per-lane scene and job */
        if (width == 8 && dl_query)
            hipLaunchKernelGGL((k_finish<true, 128, 0, 8, 0, true>), g, b, 0, s, jobs, P, C, bounce0, spill, accum, c,
                               uscene, order, cam_sched, tile, band, st);
        else if (width == 8)
            hipLaunchKernelGGL((k_finish<true, 128, 0, 8, 0, false>), g, b, 0, s, jobs, P, C, bounce0, spill, accum, c,
                               uscene, order, cam_sched, tile, band, st);
        else if (dl_query)
            hipLaunchKernelGGL((k_finish<true, 128, 0, 4, 0, true>), g, b, 0, s, jobs, P, C, bounce0, spill, accum, c,
                               uscene, order, cam_sched, tile, band, st);
        else
            hipLaunchKernelGGL((k_finish<true, 128, 0, 4, 0, false>), g, b, 0, s, jobs, P, C, bounce0, spill, accum, c,
                               uscene, order, cam_sched, tile, band, st);
    } else if (uscene && single_job) {
        if (width == 8) lx_finish_w<8, 2>(a, threads, block, occ, s);
        else lx_finish_w<4, 2>(a, threads, block, occ, s);
    } else if (uscene) {
        if (width == 8) lx_finish_w<8, 1>(a, threads, block, occ, s);
        else lx_finish_w<4, 1>(a, threads, block, occ, s);
    } else {
        a.us = nullptr;
        if (width == 8) lx_finish_w<8, 0>(a, threads, block, occ, s);
        else lx_finish_w<4, 0>(a, threads, block, occ, s);
    }
    return hipGetLastError();
}
hipError_t lx_launch_trace_rays(const LxDevScene* S, const lx_ray* rays, uint32_t n, int any, int alpha,
                                uint32_t iteration, lx_hit* hits, uint8_t* occ, uint32_t* spill, uint32_t* overflow,
                                uint32_t grid, int width, hipStream_t s) {
    if (width == 8)
        hipLaunchKernelGGL(k_trace_rays<8>, dim3(grid), dim3(LX_TRACE_BLOCK), 0, s, S, rays, n, any, alpha, iteration,
                           hits, occ, spill, overflow);
    else
        hipLaunchKernelGGL(k_trace_rays<4>, dim3(grid), dim3(LX_TRACE_BLOCK), 0, s, S, rays, n, any, alpha, iteration,
                           hits, occ, spill, overflow);
    return hipGetLastError();
}
/* one workgroup per item: strided min / max over the item's vertices, then
 * a reduction in LDS */
__global__ __launch_bounds__(256) void k_inst_bounds(const float* __restrict__ pos,
                                                     const LxBoundsItem* __restrict__ items, double* __restrict__ out) {
    __shared__ double s_b[6][256];
    const LxBoundsItem& it = items[blockIdx.x];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t v = threadIdx.x; v < it.count; v += blockDim.x) {
        const float* p = pos + 3u * (size_t)(it.first + v);
        const double x = p[0], y = p[1], z = p[2];
        for (int r = 0; r < 3; r++) {
            const double w = (double)it.m[4 * r] * x + (double)it.m[4 * r + 1] * y + (double)it.m[4 * r + 2] * z +
                             (double)it.m[4 * r + 3];
            lo[r] = fmin(lo[r], w);
            hi[r] = fmax(hi[r], w);
        }
    }
    for (int r = 0; r < 3; r++) {
        s_b[r][threadIdx.x] = lo[r];
        s_b[3 + r][threadIdx.x] = hi[r];
    }
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int r = 0; r < 3; r++) {
                s_b[r][threadIdx.x] = fmin(s_b[r][threadIdx.x], s_b[r][threadIdx.x + h]);
                s_b[3 + r][threadIdx.x] = fmax(s_b[3 + r][threadIdx.x], s_b[3 + r][threadIdx.x + h]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6u * blockIdx.x + threadIdx.x] = s_b[threadIdx.x][0];
}
hipError_t lx_launch_inst_bounds(const float* pos, const LxBoundsItem* items, uint32_t n, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_inst_bounds, dim3(n), dim3(256), 0, s, pos, items, out);
    return hipGetLastError();
}
hipError_t lx_launch_tonemap(const float* accum, uint32_t npix, float spp, int mode, uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_tonemap, dim3((npix + 255) / 256), dim3(256), 0, s, accum, npix, spp, mode, out);
    return hipGetLastError();
}
hipError_t lx_launch_math_eval(int fn, const float* x, const float* y, float* out, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_math_eval, dim3((n + 255) / 256), dim3(256), 0, s, fn, x, y, out, n);
    return hipGetLastError();
}
hipError_t lx_launch_copy16(const void* src, void* dst, size_t n16, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, (const float4*)src, (float4*)dst, n16);
    return hipGetLastError();
}
hipError_t lx_launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t lx_launch_fill_ranges(uint32_t* p, const uint32_t* ranges, uint32_t nranges, uint32_t v, hipStream_t s) {
    if (nranges == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_ranges, dim3(64, nranges), dim3(256), 0, s, p, ranges, v);
    return hipGetLastError();
}
hipError_t lx_launch_add_ranges(float* dst, const float* src, const uint32_t* ranges, uint32_t nranges, hipStream_t s) {
    if (nranges == 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_ranges, dim3(64, nranges), dim3(256), 0, s, dst, src, ranges);
    return hipGetLastError();
}
hipError_t lx_launch_add_f32(float* dst, const float* src, size_t n, hipStream_t s) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_add_f32, dim3((uint32_t)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}
hipError_t lx_launch_add_regions(float* dst, const float* src, size_t n, uint32_t regions, hipStream_t s) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_add_regions, dim3((uint32_t)blocks), dim3(256), 0, s, dst, src, n, regions);
    return hipGetLastError();
}

/* Scratch of lx_trace_closest_noalloc, the trace entry point that allocates
 * nothing: the traversal-stack spill columns of LX_NOALLOC_MAX_RAYS lanes,
 * then one 256-byte line whose first word is the stack-overflow flag.  Each
 * device holds its own copy, at its own address, until the process ends.
 * In host code the name `scratch` is only a host-side stand-in: host code
 * uses the address hipGetSymbolAddress returns for the current device. */
static __device__ __align__(256) uint32_t scratch[LX_NOALLOC_SCRATCH_WORDS];

hipError_t lx_noalloc_scratch_address(int device, uint32_t** out) {
    static std::mutex mu;
    static std::vector<uint32_t*> by_device;
    *out = nullptr;
    if (device < 0) return hipErrorInvalidDevice;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if ((size_t)device >= by_device.size()) by_device.resize((size_t)device + 1, nullptr);
    if (!by_device[device]) {
        uint32_t* p = nullptr;
        e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        by_device[device] = p;
    }
    *out = by_device[device];
    return hipSuccess;
}
