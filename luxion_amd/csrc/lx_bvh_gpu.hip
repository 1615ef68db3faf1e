/*
 * lx_bvh_gpu.hip -- GPU BLAS builder (builder 0): top-down binned SAH on the
 * device over Morton-ordered triangles, collapsed to 4- or 8-wide nodes.
 *
 * Replaces the per-mesh OptiX GAS build (src/myoptix/optix_accel.cpp:9-110),
 * which runs on the GPU in the reference.  Steps, one mesh at a time:
 *   1. per triangle: bounds and centroid; the mesh's triangle and centroid
 *      bounds (block reductions + order-preserving integer atomics);
 *   2. 30-bit Morton codes of the centroids, radix-sorted with their
 *      triangle indices (hipCUB): the initial triangle order, so that every
 *      node's triangles, and the passes over them, stay spatially compact;
 *   3. breadth-first binned SAH for nodes of more than kSmallNode triangles,
 *      one level per round of launches: bin the triangles of every active
 *      node (per-workgroup bins in LDS when a workgroup's triangles belong to
 *      one node, else global atomics), choose each node's split (one thread
 *      per node: the sweep and cost of the host builder), partition the
 *      triangles into the children's ranges (atomic cursors) while reducing
 *      the children's triangle and centroid bounds;
 *   4. every node of at most kSmallNode triangles is built to its leaves by
 *      one thread with the host builder's recursion (binned SAH, the same
 *      leaf rule, in-place partition, split by index when all centroids
 *      coincide);
 *   5. the host builder's collapse to 4- or 8-wide nodes, one level of
 *      wide nodes per round of launches (children chosen per node, the next
 *      level numbered by a prefix sum), written packed at the traversal's
 *      width with absolute child refs; the triangles are written in leaf
 *      order as (v0 | prim, e1, e2) and, for a group of instances sharing
 *      one BLAS, tagged with their mesh-local index and instance.
 * Nothing of the tree goes back to the host but the root's first two levels
 * of boxes (for the instance world boxes), the depth and the SAH cost.  The
 * tree has the host builder's SAH quality; child boxes get the same
 * conservative padding, so the traversal's results do not depend on which
 * builder made the tree.
 */
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "lx_internal.h"

namespace {

/* a load that sees other compute units' stores (bypasses the per-CU cache) */
__device__ __forceinline__ float ld_agent(const float* p) {
    return __uint_as_float(__hip_atomic_load((uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ uint32_t f2key(float f) {
    /* order-preserving map of float to uint32 */
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__global__ void k_tri_bounds(const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n,
                             float* __restrict__ lo, float* __restrict__ hi, uint32_t* __restrict__ cbox) {
    __shared__ uint32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = (threadIdx.x < 3) ? 0xffffffffu : 0u;
    __syncthreads();
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        for (int a = 0; a < 3; a++) {
            float v0 = pos[3 * idx[3 * t] + a], v1 = pos[3 * idx[3 * t + 1] + a], v2 = pos[3 * idx[3 * t + 2] + a];
            float l = fminf(v0, fminf(v1, v2)), h = fmaxf(v0, fmaxf(v1, v2));
            lo[3 * t + a] = l;
            hi[3 * t + a] = h;
            float c = 0.5f * (l + h);
            atomicMin(&s[a], f2key(c));
            atomicMax(&s[3 + a], f2key(c));
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&cbox[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&cbox[threadIdx.x], s[threadIdx.x]);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton(const float* __restrict__ lo, const float* __restrict__ hi, uint32_t n,
                         const uint32_t* __restrict__ cbox, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float cl = key2f(cbox[a]), ch = key2f(cbox[3 + a]);
        float c = 0.5f * (lo[3 * t + a] + hi[3 * t + a]);
        float ext = ch - cl;
        float x = ext > 0.0f ? (c - cl) / ext : 0.0f;
        int v = (int)(x * 1024.0f);
        q[a] = (uint32_t)(v < 0 ? 0 : (v > 1023 ? 1023 : v));
    }
    keys[t] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[t] = t;
}


/* ------------------------------------------------------- binned SAH build */
constexpr uint32_t kSmallNode = 64; /* nodes of at most this many triangles: one thread each */
constexpr int kMaxBins = 32;
constexpr uint32_t kKeyEmptyLo = 0xffffffffu, kKeyEmptyHi = 0u;

struct SahNode {
    uint32_t blo[3], bhi[3]; /* triangle bounds, order-preserving keys */
    uint32_t clo[3], chi[3]; /* centroid bounds, keys */
    int32_t left, right;     /* children, -1 for a leaf */
    uint32_t first, count;   /* range in the triangle order */
    uint32_t lcur, rcur, lcount;
    int32_t axis, split;     /* axis -1: split by index */
    int32_t slot;            /* index in the level's active list, -1 otherwise */
};

struct SahBin {
    uint32_t cnt, lo[3], hi[3];
};

struct SahCounters {
    uint32_t nodes, next, small, error;
};

__device__ __forceinline__ float sah_area(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

/* the bin of centroid c on an axis with centroid bounds [cl, ch] (the host
 * builder's expression); both the binning and the partition call this, so
 * their counts agree exactly */
__device__ __forceinline__ int sah_bin_of(float c, float cl, float ch, int nb) {
    float ext = ch - cl;
    float scale = (float)nb / ext;
    int k = (int)((c - cl) * scale);
    return k < 0 ? 0 : (k >= nb ? nb - 1 : k);
}

__device__ __forceinline__ void sah_init_node(SahNode& d, uint32_t first, uint32_t count) {
    for (int a = 0; a < 3; a++) {
        d.blo[a] = d.clo[a] = kKeyEmptyLo;
        d.bhi[a] = d.chi[a] = kKeyEmptyHi;
    }
    d.left = d.right = -1;
    d.first = first;
    d.count = count;
    d.lcur = d.rcur = d.lcount = 0;
    d.axis = -1;
    d.split = 0;
    d.slot = -1;
}

/* the mesh's triangle and centroid bounds into node 0 */
__global__ void k_sah_root(const float* __restrict__ tlo, const float* __restrict__ thi, uint32_t n,
                           SahNode* __restrict__ nodes) {
    __shared__ uint32_t s[12];
    if (threadIdx.x < 12) s[threadIdx.x] = (threadIdx.x % 6) < 3 ? kKeyEmptyLo : kKeyEmptyHi;
    __syncthreads();
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        for (int a = 0; a < 3; a++) {
            float l = tlo[3 * t + a], h = thi[3 * t + a], c = 0.5f * (l + h);
            atomicMin(&s[a], f2key(l));
            atomicMax(&s[3 + a], f2key(h));
            atomicMin(&s[6 + a], f2key(c));
            atomicMax(&s[9 + a], f2key(c));
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&nodes[0].blo[threadIdx.x], s[threadIdx.x]);
        atomicMax(&nodes[0].bhi[threadIdx.x], s[3 + threadIdx.x]);
        atomicMin(&nodes[0].clo[threadIdx.x], s[6 + threadIdx.x]);
        atomicMax(&nodes[0].chi[threadIdx.x], s[9 + threadIdx.x]);
    }
}

__global__ void k_sah_set_slots(const uint32_t* __restrict__ list, uint32_t nlist, SahNode* __restrict__ nodes,
                                int32_t on) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nlist) nodes[list[i]].slot = on ? (int32_t)i : -1;
}

__global__ void k_sah_clear_bins(SahBin* __restrict__ bins, uint32_t nbins) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbins) return;
    SahBin b;
    b.cnt = 0;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = kKeyEmptyLo;
        b.hi[a] = kKeyEmptyHi;
    }
    bins[i] = b;
}

/* the workgroup's active slot if all its positions belong to one active
 * node, else -1 (block-uniform) */
__device__ __forceinline__ int32_t sah_block_slot(int32_t slot, bool live, int32_t* s_mm) {
    if (threadIdx.x == 0) {
        s_mm[0] = 0x7fffffff;
        s_mm[1] = -0x7fffffff - 1;
    }
    __syncthreads();
    if (live) {
        atomicMin(&s_mm[0], slot);
        atomicMax(&s_mm[1], slot);
    }
    __syncthreads();
    int32_t r = (s_mm[0] == s_mm[1] && s_mm[0] >= 0) ? s_mm[0] : -1;
    __syncthreads();
    return r;
}

/* bins of every active node: 3 axes x nb bins of (count, triangle bounds) */
__global__ void k_sah_bin(uint32_t n, const uint32_t* __restrict__ ord, const uint32_t* __restrict__ posnode,
                          const SahNode* __restrict__ nodes, const float* __restrict__ tlo,
                          const float* __restrict__ thi, SahBin* __restrict__ bins, int nb) {
    __shared__ SahBin sb[3 * kMaxBins];
    __shared__ int32_t s_mm[2];
    uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t slot = -1;
    uint32_t node = 0;
    if (pos < n) {
        node = posnode[pos];
        slot = nodes[node].slot;
    }
    int32_t uni = sah_block_slot(slot, pos < n, s_mm);
    if (uni >= 0) {
        for (uint32_t i = threadIdx.x; i < 3u * (uint32_t)nb; i += blockDim.x) {
            sb[i].cnt = 0;
            for (int a = 0; a < 3; a++) {
                sb[i].lo[a] = kKeyEmptyLo;
                sb[i].hi[a] = kKeyEmptyHi;
            }
        }
        __syncthreads();
    }
    if (slot >= 0) {
        uint32_t t = ord[pos];
        const SahNode& d = nodes[node];
        float lo[3], hi[3];
        for (int c = 0; c < 3; c++) {
            lo[c] = tlo[3 * t + c];
            hi[c] = thi[3 * t + c];
        }
        for (int a = 0; a < 3; a++) {
            float cl = key2f(d.clo[a]), ch = key2f(d.chi[a]);
            if (!(ch - cl > 0.0f)) continue;
            int k = sah_bin_of(0.5f * (lo[a] + hi[a]), cl, ch, nb);
            SahBin* b = uni >= 0 ? &sb[a * nb + k] : &bins[((size_t)slot * 3 + a) * nb + k];
            atomicAdd(&b->cnt, 1u);
            for (int c = 0; c < 3; c++) {
                atomicMin(&b->lo[c], f2key(lo[c]));
                atomicMax(&b->hi[c], f2key(hi[c]));
            }
        }
    }
    if (uni >= 0) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < 3u * (uint32_t)nb; i += blockDim.x) {
            if (sb[i].cnt == 0) continue;
            SahBin* g = &bins[(size_t)uni * 3 * nb + i];
            atomicAdd(&g->cnt, sb[i].cnt);
            for (int c = 0; c < 3; c++) {
                atomicMin(&g->lo[c], sb[i].lo[c]);
                atomicMax(&g->hi[c], sb[i].hi[c]);
            }
        }
    }
}

/* one thread per active node: the host builder's sweep over the bins of
 * each axis (cost = left area x count + right area x count), then two
 * children; nodes this large always split */
__global__ void k_sah_split(const uint32_t* __restrict__ list, uint32_t nlist, SahNode* __restrict__ nodes,
                            const SahBin* __restrict__ bins, int nb, uint32_t capacity,
                            SahCounters* __restrict__ cnt, uint32_t* __restrict__ next_list,
                            uint32_t* __restrict__ small_list) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlist) return;
    uint32_t me = list[i];
    SahNode& d = nodes[me];
    float best_cost = 3.402823466e+38f;
    int best_axis = -1, best_split = -1;
    float rarea[kMaxBins];
    uint32_t rcnt[kMaxBins];
    for (int a = 0; a < 3; a++) {
        float cl = key2f(d.clo[a]), ch = key2f(d.chi[a]);
        if (!(ch - cl > 0.0f)) continue;
        const SahBin* b = bins + ((size_t)i * 3 + a) * nb;
        float rl[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
        float rh[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
        uint32_t rc = 0;
        for (int k = nb - 1; k >= 1; k--) {
            rc += b[k].cnt;
            if (b[k].cnt)
                for (int c = 0; c < 3; c++) {
                    rl[c] = fminf(rl[c], key2f(b[k].lo[c]));
                    rh[c] = fmaxf(rh[c], key2f(b[k].hi[c]));
                }
            rcnt[k] = rc;
            rarea[k] = sah_area(rl, rh);
        }
        float ll[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
        float lh[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
        uint32_t lc = 0;
        for (int k = 0; k < nb - 1; k++) {
            lc += b[k].cnt;
            if (b[k].cnt)
                for (int c = 0; c < 3; c++) {
                    ll[c] = fminf(ll[c], key2f(b[k].lo[c]));
                    lh[c] = fmaxf(lh[c], key2f(b[k].hi[c]));
                }
            if (lc == 0 || rcnt[k + 1] == 0) continue;
            float cost = sah_area(ll, lh) * (float)lc + rarea[k + 1] * (float)rcnt[k + 1];
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = a;
                best_split = k;
            }
        }
    }
    uint32_t lcount = d.count / 2; /* all centroids coincide: split by index */
    if (best_axis >= 0) {
        const SahBin* b = bins + ((size_t)i * 3 + best_axis) * nb;
        lcount = 0;
        for (int k = 0; k <= best_split; k++) lcount += b[k].cnt;
    }
    uint32_t base = atomicAdd(&cnt->nodes, 2u);
    if (base + 2 > capacity || lcount == 0 || lcount >= d.count) {
        atomicOr(&cnt->error, 1u);
        return;
    }
    sah_init_node(nodes[base], d.first, lcount);
    sah_init_node(nodes[base + 1], d.first + lcount, d.count - lcount);
    d.left = (int32_t)base;
    d.right = (int32_t)(base + 1);
    d.axis = best_axis;
    d.split = best_split;
    d.lcount = lcount;
    d.lcur = d.rcur = 0;
    for (uint32_t c = 0; c < 2; c++) {
        uint32_t cc = c ? d.count - lcount : lcount;
        if (cc > kSmallNode) next_list[atomicAdd(&cnt->next, 1u)] = base + c;
        else small_list[atomicAdd(&cnt->small, 1u)] = base + c;
    }
}

/* triangles of every active node to their child's range; the children's
 * triangle and centroid bounds by atomics (per workgroup in LDS when the
 * workgroup's triangles all belong to one node).  Positions of inactive
 * nodes are copied. */
__global__ void k_sah_partition(uint32_t n, const uint32_t* __restrict__ ord, const uint32_t* __restrict__ posnode,
                                uint32_t* __restrict__ ord2, uint32_t* __restrict__ posnode2,
                                SahNode* __restrict__ nodes, const float* __restrict__ tlo,
                                const float* __restrict__ thi, int nb, SahCounters* __restrict__ cnt) {
    __shared__ uint32_t sk[2][12];
    __shared__ int32_t s_mm[2];
    uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t slot = -1;
    uint32_t node = 0;
    if (pos < n) {
        node = posnode[pos];
        slot = nodes[node].slot;
    }
    int32_t uni = sah_block_slot(slot, pos < n, s_mm);
    if (uni >= 0) {
        if (threadIdx.x < 24) sk[threadIdx.x / 12][threadIdx.x % 12] = (threadIdx.x % 12) < 3 ||
                                                                             ((threadIdx.x % 12) >= 6 &&
                                                                              (threadIdx.x % 12) < 9)
                                                                             ? kKeyEmptyLo
                                                                             : kKeyEmptyHi;
        __syncthreads();
    }
    uint32_t t = 0;
    if (pos < n) t = ord[pos];
    if (pos < n && slot < 0) {
        ord2[pos] = t;
        posnode2[pos] = node;
    } else if (slot >= 0) {
        SahNode& d = nodes[node];
        float lo[3], hi[3];
        for (int c = 0; c < 3; c++) {
            lo[c] = tlo[3 * t + c];
            hi[c] = thi[3 * t + c];
        }
        bool left;
        if (d.axis < 0) {
            left = pos - d.first < d.lcount;
        } else {
            int a = d.axis;
            left = sah_bin_of(0.5f * (lo[a] + hi[a]), key2f(d.clo[a]), key2f(d.chi[a]), nb) <= d.split;
        }
        uint32_t k;
        if (uni >= 0) {
            /* every live lane of the workgroup is in this node: one atomic
             * per wave and side, lanes ranked by the ballot */
            const uint64_t ml = __ballot(left), mr = __ballot(!left);
            const uint64_t below = (1ull << __lane_id()) - 1ull;
            const uint32_t lead = (uint32_t)__ffsll((unsigned long long)(ml | mr)) - 1u;
            uint32_t bl = 0, br = 0;
            if (__lane_id() == lead) {
                if (ml) bl = atomicAdd(&d.lcur, (uint32_t)__popcll(ml));
                if (mr) br = atomicAdd(&d.rcur, (uint32_t)__popcll(mr));
            }
            bl = __shfl(bl, (int)lead);
            br = __shfl(br, (int)lead);
            k = left ? bl + (uint32_t)__popcll(ml & below) : br + (uint32_t)__popcll(mr & below);
        } else {
            k = left ? atomicAdd(&d.lcur, 1u) : atomicAdd(&d.rcur, 1u);
        }
        uint32_t lim = left ? d.lcount : d.count - d.lcount;
        if (k >= lim) {
            atomicOr(&cnt->error, 2u);
        } else {
            uint32_t dst = d.first + (left ? 0u : d.lcount) + k;
            uint32_t child = (uint32_t)(left ? d.left : d.right);
            ord2[dst] = t;
            posnode2[dst] = child;
            if (uni >= 0) {
                uint32_t* q = sk[left ? 0 : 1];
                for (int c = 0; c < 3; c++) {
                    float cc = 0.5f * (lo[c] + hi[c]);
                    atomicMin(&q[c], f2key(lo[c]));
                    atomicMax(&q[3 + c], f2key(hi[c]));
                    atomicMin(&q[6 + c], f2key(cc));
                    atomicMax(&q[9 + c], f2key(cc));
                }
            } else {
                SahNode& cd = nodes[child];
                for (int c = 0; c < 3; c++) {
                    float cc = 0.5f * (lo[c] + hi[c]);
                    atomicMin(&cd.blo[c], f2key(lo[c]));
                    atomicMax(&cd.bhi[c], f2key(hi[c]));
                    atomicMin(&cd.clo[c], f2key(cc));
                    atomicMax(&cd.chi[c], f2key(cc));
                }
            }
        }
    }
    if (uni >= 0) {
        __syncthreads();
        if (threadIdx.x < 24) {
            int side = threadIdx.x / 12, c = threadIdx.x % 12;
            /* the node of this workgroup: the node of its first position */
            SahNode& d = nodes[posnode[blockIdx.x * blockDim.x]];
            SahNode& cd = nodes[side ? d.right : d.left];
            uint32_t v = sk[side][c];
            uint32_t* dst = c < 3 ? &cd.blo[c] : c < 6 ? &cd.bhi[c - 3] : c < 9 ? &cd.clo[c - 6] : &cd.chi[c - 9];
            if ((c < 3) || (c >= 6 && c < 9)) atomicMin(dst, v);
            else atomicMax(dst, v);
        }
    }
}

/* one thread per node of at most kSmallNode triangles: the host builder's
 * recursion (bounds, binned SAH over three axes, leaf rule, in-place
 * partition, split by index when it would leave one side empty) down to the
 * leaves, with an explicit stack */
__global__ void k_sah_small(const uint32_t* __restrict__ list, uint32_t nlist, SahNode* __restrict__ nodes,
                            uint32_t* __restrict__ ord, const float* __restrict__ tlo,
                            const float* __restrict__ thi, int nb, int leaf_size, uint32_t capacity,
                            SahCounters* __restrict__ cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlist) return;
    uint32_t stack[kSmallNode];
    int sp = 0;
    stack[sp++] = list[i];
    uint32_t bcnt[kMaxBins];
    float blo[kMaxBins][3], bhi[kMaxBins][3];
    float rarea[kMaxBins];
    uint32_t rcnt[kMaxBins];
    while (sp > 0) {
        uint32_t me = stack[--sp];
        SahNode& d = nodes[me];
        uint32_t b = d.first, e = d.first + d.count, n = d.count;
        float lo[3], hi[3], clo[3], chi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = clo[a] = 3.402823466e+38f;
            hi[a] = chi[a] = -3.402823466e+38f;
        }
        for (uint32_t p = b; p < e; p++) {
            uint32_t t = ord[p];
            for (int a = 0; a < 3; a++) {
                float l = tlo[3 * t + a], h = thi[3 * t + a], c = 0.5f * (l + h);
                lo[a] = fminf(lo[a], l);
                hi[a] = fmaxf(hi[a], h);
                clo[a] = fminf(clo[a], c);
                chi[a] = fmaxf(chi[a], c);
            }
        }
        for (int a = 0; a < 3; a++) {
            d.blo[a] = f2key(lo[a]);
            d.bhi[a] = f2key(hi[a]);
        }
        d.left = d.right = -1;
        if (n <= 1) continue;
        float best_cost = 3.402823466e+38f;
        int best_axis = -1, best_split = -1;
        for (int a = 0; a < 3; a++) {
            if (!(chi[a] - clo[a] > 0.0f)) continue;
            for (int k = 0; k < nb; k++) {
                bcnt[k] = 0;
                for (int c = 0; c < 3; c++) {
                    blo[k][c] = 3.402823466e+38f;
                    bhi[k][c] = -3.402823466e+38f;
                }
            }
            for (uint32_t p = b; p < e; p++) {
                uint32_t t = ord[p];
                float l[3], h[3];
                for (int c = 0; c < 3; c++) {
                    l[c] = tlo[3 * t + c];
                    h[c] = thi[3 * t + c];
                }
                int k = sah_bin_of(0.5f * (l[a] + h[a]), clo[a], chi[a], nb);
                bcnt[k]++;
                for (int c = 0; c < 3; c++) {
                    blo[k][c] = fminf(blo[k][c], l[c]);
                    bhi[k][c] = fmaxf(bhi[k][c], h[c]);
                }
            }
            float rl[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
            float rh[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
            uint32_t rc = 0;
            for (int k = nb - 1; k >= 1; k--) {
                rc += bcnt[k];
                for (int c = 0; c < 3; c++) {
                    rl[c] = fminf(rl[c], blo[k][c]);
                    rh[c] = fmaxf(rh[c], bhi[k][c]);
                }
                rcnt[k] = rc;
                rarea[k] = sah_area(rl, rh);
            }
            float ll[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
            float lh[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
            uint32_t lc = 0;
            for (int k = 0; k < nb - 1; k++) {
                lc += bcnt[k];
                for (int c = 0; c < 3; c++) {
                    ll[c] = fminf(ll[c], blo[k][c]);
                    lh[c] = fmaxf(lh[c], bhi[k][c]);
                }
                if (lc == 0 || rcnt[k + 1] == 0) continue;
                float cost = sah_area(ll, lh) * (float)lc + rarea[k + 1] * (float)rcnt[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = k;
                }
            }
        }
        float parea = sah_area(lo, hi);
        /* leaf if small enough and splitting does not pay (C_trav = 1, C_tri = 1 per area unit) */
        if ((int)n <= leaf_size && (best_axis < 0 || parea * (float)n <= parea + best_cost)) continue;
        uint32_t mid = b + n / 2;
        if (best_axis >= 0) {
            /* in-place partition: bin <= best_split to the front */
            uint32_t lo_i = b, hi_i = e;
            while (lo_i < hi_i) {
                uint32_t t = ord[lo_i];
                float c = 0.5f * (tlo[3 * t + best_axis] + thi[3 * t + best_axis]);
                if (sah_bin_of(c, clo[best_axis], chi[best_axis], nb) <= best_split) {
                    lo_i++;
                } else {
                    hi_i--;
                    ord[lo_i] = ord[hi_i];
                    ord[hi_i] = t;
                }
            }
            mid = lo_i;
            if (mid == b || mid == e) mid = b + n / 2;
        }
        uint32_t base = atomicAdd(&cnt->nodes, 2u);
        if (base + 2 > capacity || sp + 2 > (int)kSmallNode) {
            atomicOr(&cnt->error, 4u);
            return;
        }
        sah_init_node(nodes[base], b, mid - b);
        sah_init_node(nodes[base + 1], mid, e - mid);
        d.left = (int32_t)base;
        d.right = (int32_t)(base + 1);
        stack[sp++] = base + 1;
        stack[sp++] = base;
    }
}

/* ------------------------------------------------ collapse on the device */
/* The host builder's collapse (lx_bvh_host.cpp), one level of wide nodes per
 * round of launches: each wide node of the level takes its binary node's two
 * children and replaces the largest-area internal one by its children until
 * it has `W`; its internal children become the next level, numbered in
 * level order by a prefix sum over the counts, so the numbering does not
 * depend on scheduling.  Nodes are written packed at the traversal's width
 * (LxNode4 / LxNode8) with absolute child refs and the host collapse's
 * padded child boxes. */

__device__ __forceinline__ void sah_box(const SahNode& d, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; a++) {
        lo[a] = key2f(d.blo[a]);
        hi[a] = key2f(d.bhi[a]);
    }
}

__device__ __forceinline__ float sah_node_area(const SahNode& d) {
    float lo[3], hi[3];
    sah_box(d, lo, hi);
    return sah_area(lo, hi);
}

/* lx_pad_box, the same operations in the same order */
__device__ __forceinline__ void wide_pad_box(float lo[3], float hi[3]) {
    for (int a = 0; a < 3; a++) {
        float m = fmaxf(fabsf(lo[a]), fabsf(hi[a]));
        float pad = (m + (hi[a] - lo[a])) * 1.9073486e-6f + 8.6736174e-19f;
        lo[a] = lo[a] - pad;
        hi[a] = hi[a] + pad;
    }
}

/* per wide node of the level: its children (binary node ids, -1 past the
 * last), how many are internal, and its share of the SAH cost (its own box
 * area plus each leaf child's area times its triangle count, over the root
 * box area) */
__global__ void k_wide_kids(const uint32_t* __restrict__ level, uint32_t n, const SahNode* __restrict__ nodes, int W,
                            int32_t* __restrict__ kids, uint32_t* __restrict__ n_internal, double* __restrict__ sah_part) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SahNode& b = nodes[level[i]];
    int32_t k8[8] = {b.left, b.right, -1, -1, -1, -1, -1, -1};
    float karea[8] = {sah_node_area(nodes[b.left]), sah_node_area(nodes[b.right]), 0, 0, 0, 0, 0, 0};
    int nk = 2;
    while (nk < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; k++) {
            if (nodes[k8[k]].left < 0) continue;
            if (karea[k] > ba) {
                ba = karea[k];
                best = k;
            }
        }
        if (best < 0) break;
        const SahNode& c = nodes[k8[best]];
        int32_t cl = c.left, cr = c.right;
        k8[best] = cl;
        karea[best] = sah_node_area(nodes[cl]);
        k8[nk] = cr;
        karea[nk++] = sah_node_area(nodes[cr]);
    }
    const double inv_ra = 1.0 / fmax(1e-30, (double)sah_node_area(nodes[0]));
    double part = (double)sah_node_area(b) * inv_ra;
    uint32_t ni = 0;
    for (int k = 0; k < 8; k++) {
        kids[8 * (size_t)i + k] = k8[k];
        if (k >= nk) continue;
        const SahNode& c = nodes[k8[k]];
        if (c.left >= 0) ni++;
        else part += (double)karea[k] * inv_ra * (double)c.count;
    }
    n_internal[i] = ni;
    sah_part[i] = part;
}

/* the next level's size: last offset + last count */
__global__ void k_wide_total(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ n_internal, uint32_t n,
                             uint32_t* __restrict__ total) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *total = offs[n - 1] + n_internal[n - 1];
}

template <int W, class N>
__global__ void k_wide_emit(const uint32_t* __restrict__ level, uint32_t n, const SahNode* __restrict__ nodes,
                            const int32_t* __restrict__ kids, const uint32_t* __restrict__ offs, uint32_t level_base,
                            uint32_t next_base, uint32_t node_base, uint32_t tri_base, N* __restrict__ out,
                            uint32_t* __restrict__ next_level) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lox[W], loy[W], loz[W], hix[W], hiy[W], hiz[W];
    uint32_t ch[W];
    uint32_t slot = offs[i], nk = 0;
    for (int k = 0; k < W; k++) {
        int32_t kid = kids[8 * (size_t)i + k];
        if (kid < 0) {
            /* an empty slot's box is inverted (lo = +inf, hi = -inf): every slab test misses it */
            lox[k] = loy[k] = loz[k] = __builtin_inff();
            hix[k] = hiy[k] = hiz[k] = -__builtin_inff();
            ch[k] = LX_CHILD_EMPTY;
            continue;
        }
        nk++;
        const SahNode& c = nodes[kid];
        float lo[3], hi[3];
        sah_box(c, lo, hi);
        wide_pad_box(lo, hi);
        lox[k] = lo[0];
        loy[k] = lo[1];
        loz[k] = lo[2];
        hix[k] = hi[0];
        hiy[k] = hi[1];
        hiz[k] = hi[2];
        if (c.left >= 0) {
            next_level[slot] = (uint32_t)kid;
            ch[k] = node_base + next_base + slot;
            slot++;
        } else {
            ch[k] = LX_LEAF_BIT | ((c.count - 1u) << 28) | (c.first + tri_base);
        }
    }
    N& o = out[level_base + i];
    float* plox = (float*)&o.lox;
    float* phix = (float*)&o.hix;
    float* ploy = (float*)&o.loy;
    float* phiy = (float*)&o.hiy;
    float* ploz = (float*)&o.loz;
    float* phiz = (float*)&o.hiz;
    uint32_t* pch = (uint32_t*)&o.child;
    int* meta = (int*)&o.meta;
    for (int k = 0; k < W; k++) {
        plox[k] = lox[k];
        phix[k] = hix[k];
        ploy[k] = loy[k];
        phiy[k] = hiy[k];
        ploz[k] = loz[k];
        phiz[k] = hiz[k];
        pch[k] = ch[k];
        meta[k] = 0;
    }
    meta[1] = (int)nk; /* meta.x = 0: a BLAS node */
}

/* group triangle index -> (index within its mesh, instance): the scene's
 * tagging of a group's triangle records (tag_off ascending, tag_off[0] = 0) */
__global__ void k_tag_tris(float4* __restrict__ tris, uint32_t n, const uint32_t* __restrict__ tag_off,
                           const uint32_t* __restrict__ tag_inst, uint32_t ntag) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t t = __float_as_uint(tris[3 * k].w);
    uint32_t lo = 0, hi = ntag; /* the last m with tag_off[m] <= t */
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (tag_off[mid] <= t) lo = mid;
        else hi = mid;
    }
    tris[3 * k].w = __uint_as_float(t - tag_off[lo]);
    tris[3 * k + 1].w = __uint_as_float(tag_inst[lo]);
}

__global__ void k_reorder(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                          const uint32_t* __restrict__ order, uint32_t n, float4* __restrict__ tris) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t t = order[k];
    const float* a = pos + 3 * idx[3 * t];
    const float* b = pos + 3 * idx[3 * t + 1];
    const float* c = pos + 3 * idx[3 * t + 2];
    tris[3 * k + 0] = make_float4(a[0], a[1], a[2], __uint_as_float(t));
    tris[3 * k + 1] = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.0f);
    tris[3 * k + 2] = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.0f);
}

} // namespace

namespace {

constexpr int kMaxWideLevels = 1024;

/* the collapse's levels: wide nodes in level order, numbered from 0 (the
 * root) at node_base; returns the node count (0 if the tree is deeper than
 * kMaxWideLevels), the depth and the SAH cost */
template <int W, class N>
uint32_t collapse_levels(lx_context* ctx, const DevBuf& d_nodes, uint32_t nbin, uint32_t node_base, uint32_t tri_base,
                         DevBuf& d_out, int& depth, double& sah) {
    hipStream_t s = ctx->stream;
    const uint32_t cap = nbin / 2 + 1; /* at most one wide node per internal binary node */
    d_out.alloc(ctx, (size_t)cap * sizeof(N));
    DevBuf d_lv[2], d_kids, d_ni, d_offs, d_part, d_sum, d_tot, d_tmp;
    d_lv[0].alloc(ctx, (size_t)cap * 4);
    d_lv[1].alloc(ctx, (size_t)cap * 4);
    d_kids.alloc(ctx, (size_t)cap * 32);
    d_ni.alloc(ctx, (size_t)cap * 4);
    d_offs.alloc(ctx, (size_t)cap * 4);
    d_part.alloc(ctx, (size_t)cap * 8);
    d_sum.alloc(ctx, kMaxWideLevels * 8);
    d_tot.alloc(ctx, 4);
    size_t scan_bytes = 0, red_bytes = 0;
    LX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, d_ni.as<uint32_t>(), d_offs.as<uint32_t>(), (int)cap,
                                            s));
    LX_HIP(hipcub::DeviceReduce::Sum(nullptr, red_bytes, d_part.as<double>(), d_sum.as<double>(), (int)cap, s));
    d_tmp.alloc(ctx, std::max(scan_bytes, red_bytes) + 16);
    static const uint32_t kRootId = 0;
    LX_HIP(hipMemcpyAsync(d_lv[0].ptr, &kRootId, 4, hipMemcpyHostToDevice, s));
    uint32_t n = 1, base = 0;
    int level = 0, cur = 0;
    while (n > 0) {
        if (level >= kMaxWideLevels || base + n > cap) return 0;
        const uint32_t g = (n + 255) / 256;
        hipLaunchKernelGGL(k_wide_kids, dim3(g), dim3(256), 0, s, d_lv[cur].as<uint32_t>(), n, d_nodes.as<SahNode>(),
                           W, d_kids.as<int32_t>(), d_ni.as<uint32_t>(), d_part.as<double>());
        LX_HIP(hipGetLastError());
        size_t tb = d_tmp.bytes;
        LX_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp.ptr, tb, d_ni.as<uint32_t>(), d_offs.as<uint32_t>(), (int)n, s));
        tb = d_tmp.bytes;
        LX_HIP(hipcub::DeviceReduce::Sum(d_tmp.ptr, tb, d_part.as<double>(), d_sum.as<double>() + level, (int)n, s));
        hipLaunchKernelGGL(k_wide_total, dim3(1), dim3(64), 0, s, d_offs.as<uint32_t>(), d_ni.as<uint32_t>(), n,
                           d_tot.as<uint32_t>());
        hipLaunchKernelGGL((k_wide_emit<W, N>), dim3(g), dim3(256), 0, s, d_lv[cur].as<uint32_t>(), n,
                           d_nodes.as<SahNode>(), d_kids.as<int32_t>(), d_offs.as<uint32_t>(), base, base + n,
                           node_base, tri_base, d_out.as<N>(), d_lv[cur ^ 1].as<uint32_t>());
        LX_HIP(hipGetLastError());
        uint32_t next = 0;
        LX_HIP(hipMemcpyAsync(&next, d_tot.ptr, 4, hipMemcpyDeviceToHost, s));
        LX_HIP(hipStreamSynchronize(s));
        base += n;
        n = next;
        cur ^= 1;
        level++;
    }
    std::vector<double> sums((size_t)level);
    LX_HIP(hipMemcpyAsync(sums.data(), d_sum.ptr, (size_t)level * 8, hipMemcpyDeviceToHost, s));
    LX_HIP(hipStreamSynchronize(s));
    sah = 0.0;
    for (double v : sums) sah += v;
    depth = level; /* the deepest wide nodes' children are one level further down */
    return base;
}

/* the boxes two levels below the root of a packed tree (root = node 0) */
template <int W, class N>
void top_boxes(lx_context* ctx, const DevBuf& d_out, uint32_t count, uint32_t node_base,
               std::vector<std::array<float, 6>>& boxes) {
    uint32_t m = std::min<uint32_t>(count, 1 + W);
    std::vector<N> h(m);
    LX_HIP(hipMemcpyAsync(h.data(), d_out.ptr, (size_t)m * sizeof(N), hipMemcpyDeviceToHost, ctx->stream));
    LX_HIP(hipStreamSynchronize(ctx->stream));
    auto child_box = [&](const N& n, int k) {
        const float *lox = (const float*)&n.lox, *loy = (const float*)&n.loy, *loz = (const float*)&n.loz;
        const float *hix = (const float*)&n.hix, *hiy = (const float*)&n.hiy, *hiz = (const float*)&n.hiz;
        return std::array<float, 6>{lox[k], loy[k], loz[k], hix[k], hiy[k], hiz[k]};
    };
    const uint32_t* rch = (const uint32_t*)&h[0].child;
    for (int k = 0; k < W; k++) {
        uint32_t c = rch[k];
        if (c == LX_CHILD_EMPTY) continue;
        if (c & LX_LEAF_BIT) {
            boxes.push_back(child_box(h[0], k));
            continue;
        }
        uint32_t local = c - node_base;
        if (local >= m) throw LxError{LX_ERR_STATE, "GPU BVH: a root child outside the first level"};
        const uint32_t* cch = (const uint32_t*)&h[local].child;
        for (int q = 0; q < W; q++)
            if (cch[q] != LX_CHILD_EMPTY) boxes.push_back(child_box(h[local], q));
    }
}

} // namespace

bool lx_build_blas_gpu(lx_context* ctx, const float* pos, size_t nv, const uint32_t* idx, uint32_t ntri,
                       int leaf_size, int bins, int width, uint32_t node_base, uint32_t tri_base,
                       const uint32_t* tag_off, const uint32_t* tag_inst, size_t ntag, BvhBuildResult& out) {
    if (ntri == 0 || ntri >= (1u << 28) - 1 || nv == 0) return false;
    if (leaf_size < 1) leaf_size = 1;
    if (leaf_size > 8) leaf_size = 8;
    int nb = std::max(4, std::min(kMaxBins, bins > 0 ? bins : 16));
    hipStream_t s = ctx->stream;
    const uint32_t n = ntri;
    const uint32_t blocks = (n + 255) / 256;
    DevBuf d_pos, d_idx, d_lo, d_hi, d_cbox, d_keys, d_vals, d_keys2, d_ord, d_tmp, d_ord2, d_pn, d_pn2, d_nodes,
        d_bins, d_list[2], d_small, d_cnt;
    d_pos.upload(ctx, pos, nv * 12); /* indices are below nv: checked when the scene is created */
    d_idx.upload(ctx, idx, (size_t)ntri * 12);
    d_lo.alloc(ctx, (size_t)n * 12);
    d_hi.alloc(ctx, (size_t)n * 12);
    uint32_t cinit[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    d_cbox.upload(ctx, cinit, sizeof(cinit));
    hipLaunchKernelGGL(k_tri_bounds, dim3(blocks), dim3(256), 0, s, d_pos.as<float>(), d_idx.as<uint32_t>(), n,
                       d_lo.as<float>(), d_hi.as<float>(), d_cbox.as<uint32_t>());
    LX_HIP(hipGetLastError());
    /* Morton order of the centroids: the initial triangle order */
    d_keys.alloc(ctx, (size_t)n * 4);
    d_vals.alloc(ctx, (size_t)n * 4);
    d_keys2.alloc(ctx, (size_t)n * 4);
    d_ord.alloc(ctx, (size_t)n * 4);
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, s, d_lo.as<float>(), d_hi.as<float>(), n,
                       d_cbox.as<uint32_t>(), d_keys.as<uint32_t>(), d_vals.as<uint32_t>());
    LX_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    LX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_keys.as<uint32_t>(), d_keys2.as<uint32_t>(),
                                              d_vals.as<uint32_t>(), d_ord.as<uint32_t>(), (int)n, 0, 30, s));
    d_tmp.alloc(ctx, tmp_bytes + 16);
    LX_HIP(hipcub::DeviceRadixSort::SortPairs(d_tmp.ptr, tmp_bytes, d_keys.as<uint32_t>(), d_keys2.as<uint32_t>(),
                                              d_vals.as<uint32_t>(), d_ord.as<uint32_t>(), (int)n, 0, 30, s));
    /* the binary tree: at most 2n - 1 nodes */
    const uint32_t capacity = 2 * n;
    d_nodes.alloc(ctx, (size_t)capacity * sizeof(SahNode));
    SahNode root;
    for (int a = 0; a < 3; a++) {
        root.blo[a] = root.clo[a] = kKeyEmptyLo;
        root.bhi[a] = root.chi[a] = kKeyEmptyHi;
    }
    root.left = root.right = -1;
    root.first = 0;
    root.count = n;
    root.lcur = root.rcur = root.lcount = 0;
    root.axis = -1;
    root.split = 0;
    root.slot = -1;
    LX_HIP(hipMemcpyAsync(d_nodes.ptr, &root, sizeof(root), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sah_root, dim3(blocks), dim3(256), 0, s, d_lo.as<float>(), d_hi.as<float>(), n,
                       d_nodes.as<SahNode>());
    LX_HIP(hipGetLastError());
    d_cnt.alloc(ctx, sizeof(SahCounters));
    d_small.alloc(ctx, (size_t)n * 4);
    SahCounters hc = {1u, 0u, 0u, 0u};
    uint32_t nsmall = 0;
    if (n > kSmallNode) {
        /* breadth-first levels over the nodes too large for one thread;
         * active nodes are disjoint and hold more than kSmallNode triangles
         * each, so at most n / kSmallNode are active at once */
        const uint32_t max_active = n / kSmallNode + 1;
        d_list[0].alloc(ctx, (size_t)max_active * 4);
        d_list[1].alloc(ctx, (size_t)max_active * 4);
        d_bins.alloc(ctx, (size_t)max_active * 3 * nb * sizeof(SahBin));
        d_ord2.alloc(ctx, (size_t)n * 4);
        d_pn.alloc(ctx, (size_t)n * 4);
        d_pn2.alloc(ctx, (size_t)n * 4);
        LX_HIP(lx_launch_fill_u32(d_pn.as<uint32_t>(), n, 0u, s));
        static const uint32_t kRootId = 0;
        LX_HIP(hipMemcpyAsync(d_list[0].ptr, &kRootId, 4, hipMemcpyHostToDevice, s));
        uint32_t nact = 1;
        int cur = 0;
        while (nact > 0) {
            LX_HIP(hipMemcpyAsync(d_cnt.ptr, &hc, sizeof(hc), hipMemcpyHostToDevice, s));
            uint32_t lb = (nact + 255) / 256, nbins = nact * 3 * (uint32_t)nb;
            hipLaunchKernelGGL(k_sah_set_slots, dim3(lb), dim3(256), 0, s, d_list[cur].as<uint32_t>(), nact,
                               d_nodes.as<SahNode>(), 1);
            hipLaunchKernelGGL(k_sah_clear_bins, dim3((nbins + 255) / 256), dim3(256), 0, s, d_bins.as<SahBin>(),
                               nbins);
            hipLaunchKernelGGL(k_sah_bin, dim3(blocks), dim3(256), 0, s, n, d_ord.as<uint32_t>(),
                               d_pn.as<uint32_t>(), d_nodes.as<SahNode>(), d_lo.as<float>(), d_hi.as<float>(),
                               d_bins.as<SahBin>(), nb);
            hipLaunchKernelGGL(k_sah_split, dim3(lb), dim3(256), 0, s, d_list[cur].as<uint32_t>(), nact,
                               d_nodes.as<SahNode>(), d_bins.as<SahBin>(), nb, capacity, d_cnt.as<SahCounters>(),
                               d_list[cur ^ 1].as<uint32_t>(), d_small.as<uint32_t>());
            hipLaunchKernelGGL(k_sah_partition, dim3(blocks), dim3(256), 0, s, n, d_ord.as<uint32_t>(),
                               d_pn.as<uint32_t>(), d_ord2.as<uint32_t>(), d_pn2.as<uint32_t>(),
                               d_nodes.as<SahNode>(), d_lo.as<float>(), d_hi.as<float>(), nb,
                               d_cnt.as<SahCounters>());
            hipLaunchKernelGGL(k_sah_set_slots, dim3(lb), dim3(256), 0, s, d_list[cur].as<uint32_t>(), nact,
                               d_nodes.as<SahNode>(), 0);
            LX_HIP(hipGetLastError());
            LX_HIP(hipMemcpyAsync(&hc, d_cnt.ptr, sizeof(hc), hipMemcpyDeviceToHost, s));
            LX_HIP(hipStreamSynchronize(s));
            if (hc.error || hc.next > max_active) return false;
            std::swap(d_ord, d_ord2);
            std::swap(d_pn, d_pn2);
            nact = hc.next;
            hc.next = 0;
            cur ^= 1;
        }
        nsmall = hc.small;
    } else {
        static const uint32_t kRootId = 0;
        LX_HIP(hipMemcpyAsync(d_small.ptr, &kRootId, 4, hipMemcpyHostToDevice, s));
        nsmall = 1;
    }
    LX_HIP(hipMemcpyAsync(d_cnt.ptr, &hc, sizeof(hc), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sah_small, dim3((nsmall + 63) / 64), dim3(64), 0, s, d_small.as<uint32_t>(), nsmall,
                       d_nodes.as<SahNode>(), d_ord.as<uint32_t>(), d_lo.as<float>(), d_hi.as<float>(), nb,
                       leaf_size, capacity, d_cnt.as<SahCounters>());
    LX_HIP(hipGetLastError());
    LX_HIP(hipMemcpyAsync(&hc, d_cnt.ptr, sizeof(hc), hipMemcpyDeviceToHost, s));
    LX_HIP(hipStreamSynchronize(s));
    if (hc.error || hc.nodes > capacity) return false;
    const uint32_t nn = hc.nodes;
    /* the triangles in leaf order, tagged for a group */
    out.d_tris.alloc(ctx, (size_t)n * 48);
    hipLaunchKernelGGL(k_reorder, dim3(blocks), dim3(256), 0, s, d_pos.as<float>(), d_idx.as<uint32_t>(),
                       d_ord.as<uint32_t>(), n, out.d_tris.as<float4>());
    LX_HIP(hipGetLastError());
    if (ntag) {
        DevBuf d_toff, d_tinst;
        d_toff.upload(ctx, tag_off, ntag * 4);
        d_tinst.upload(ctx, tag_inst, ntag * 4);
        hipLaunchKernelGGL(k_tag_tris, dim3(blocks), dim3(256), 0, s, out.d_tris.as<float4>(), n,
                           d_toff.as<uint32_t>(), d_tinst.as<uint32_t>(), (uint32_t)ntag);
        LX_HIP(hipGetLastError());
        LX_HIP(hipStreamSynchronize(s)); /* before the tag arrays are freed */
    }
    /* the collapse, or a single leaf */
    SahNode root_node;
    LX_HIP(hipMemcpyAsync(&root_node, d_nodes.ptr, sizeof(SahNode), hipMemcpyDeviceToHost, s));
    LX_HIP(hipStreamSynchronize(s));
    out.nodes.clear();
    out.tris.clear();
    out.top_boxes.clear();
    if (root_node.left < 0) {
        out.root = LX_LEAF_BIT | ((root_node.count - 1u) << 28) | 0u;
        out.depth = 0;
        out.sah = (double)root_node.count;
        out.dev_count = 0;
        return true;
    }
    if (width == 8) {
        out.dev_count = collapse_levels<8, LxNode8>(ctx, d_nodes, nn, node_base, tri_base, out.d_nodes, out.depth,
                                                     out.sah);
        if (out.dev_count == 0) return false;
        top_boxes<8, LxNode8>(ctx, out.d_nodes, out.dev_count, node_base, out.top_boxes);
    } else {
        out.dev_count = collapse_levels<4, LxNode4>(ctx, d_nodes, nn, node_base, tri_base, out.d_nodes, out.depth,
                                                     out.sah);
        if (out.dev_count == 0) return false;
        top_boxes<4, LxNode4>(ctx, out.d_nodes, out.dev_count, node_base, out.top_boxes);
    }
    out.root = 0;
    return true;
}
