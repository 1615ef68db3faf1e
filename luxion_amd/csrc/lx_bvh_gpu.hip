/*
 * lx_bvh_gpu.hip -- GPU BLAS builder (builder 0): linear BVH from Morton
 * codes, collapsed to 4- or 8-wide nodes on the device.
 *
 * Replaces the per-mesh OptiX GAS build (src/myoptix/optix_accel.cpp:9-110),
 * which runs on the GPU in the reference.  Steps, one mesh at a time:
 *   1. per triangle: bounds and centroid;
 *   2. centroid bounds (block reduction + order-preserving integer atomics);
 *   3. 30-bit Morton codes of the centroids, radix-sorted with their
 *      triangle indices (hipCUB);
 *   4. the binary radix tree over the sorted codes (Karras, "Maximizing
 *      parallelism in the construction of BVHs, octrees, and k-d trees",
 *      HPG 2012), with the index as tie-break for equal codes;
 *   5. node bounds bottom-up: the second child to arrive at a node computes
 *      its box (atomic arrival counters);
 *   6. top-down collapse, one launch per level: each wide node adopts the
 *      children of its largest-area internal children until it has W; a
 *      subtree of at most leaf_size triangles becomes one leaf (its
 *      triangles are contiguous in Morton order);
 *   7. triangles written in leaf (sorted) order as (v0 | prim, e1, e2).
 * Child boxes get the same conservative padding as the host builder, so the
 * traversal's results do not depend on which builder made the tree.
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

/* common-prefix length of sorted keys i and j (index tie-break), -1 outside */
__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

/* internal node i (0..n-2): children, leaf range; node ids: internal i,
 * leaf k = n - 1 + k */
__global__ void k_karras(const uint32_t* __restrict__ keys, int n, int2* __restrict__ child, int2* __restrict__ range,
                         int* __restrict__ parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : (gamma + 1);
    child[i] = make_int2(left, right);
    range[i] = make_int2(lo, hi);
    parent[left] = i;
    parent[right] = i;
}

/* bottom-up bounds: node boxes in nlo / nhi (6 floats per node, both
 * internal and leaf), arrival counters in flag */
__global__ void k_bounds(const uint32_t* __restrict__ order, const float* __restrict__ tlo,
                         const float* __restrict__ thi, int n, const int2* __restrict__ child,
                         const int* __restrict__ parent, float* nlo, float* nhi, uint32_t* flag) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int node = n - 1 + k;
    uint32_t tri = order[k];
    for (int a = 0; a < 3; a++) {
        nlo[3 * node + a] = tlo[3 * tri + a];
        nhi[3 * node + a] = thi[3 * tri + a];
    }
    if (n == 1) return;
    __threadfence();
    int p = parent[node];
    while (p >= 0) {
        if (atomicAdd(&flag[p], 1u) == 0u) return; /* the sibling finishes this node */
        __threadfence();
        int2 c = child[p];
        for (int a = 0; a < 3; a++) {
            float l0 = ld_agent(&nlo[3 * c.x + a]), l1 = ld_agent(&nlo[3 * c.y + a]);
            float h0 = ld_agent(&nhi[3 * c.x + a]), h1 = ld_agent(&nhi[3 * c.y + a]);
            nlo[3 * p + a] = fminf(l0, l1);
            nhi[3 * p + a] = fmaxf(h0, h1);
        }
        __threadfence();
        p = parent[p];
    }
}

__device__ __forceinline__ float area3(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ int count_of(int node, int n, const int2* range) {
    if (node >= n - 1) return 1;
    int2 r = range[node];
    return r.y - r.x + 1;
}

/* SAH leaf test for a subtree of at most leaf_size triangles: it stays a
 * leaf unless splitting it once (a box test, relative cost 0.5 of a
 * triangle test, then each child's triangles weighted by its area share)
 * is cheaper than testing all its triangles */
__device__ __forceinline__ bool make_leaf(int c, int n, int leaf_size, const int2* __restrict__ child,
                                          const int2* __restrict__ range, const float* __restrict__ nlo,
                                          const float* __restrict__ nhi);

__device__ __forceinline__ bool make_leaf(int c, int n, int leaf_size, const int2* __restrict__ child,
                                          const int2* __restrict__ range, const float* __restrict__ nlo,
                                          const float* __restrict__ nhi) {
    if (c >= n - 1) return true;
    int cnt = count_of(c, n, range);
    if (cnt > leaf_size) return false;
    float a = area3(nlo + 3 * c, nhi + 3 * c);
    if (!(a > 0.0f)) return true;
    int2 k = child[c];
    float split = 0.5f + (area3(nlo + 3 * k.x, nhi + 3 * k.x) * (float)count_of(k.x, n, range) +
                          area3(nlo + 3 * k.y, nhi + 3 * k.y) * (float)count_of(k.y, n, range)) / a;
    return !(split < (float)cnt);
}

/* one level of the collapse: frontier entries (binary node, wide slot) */
template <int W>
__global__ void k_collapse(const int2* __restrict__ fin, uint32_t nin, int n, int leaf_size,
                           const int2* __restrict__ child, const int2* __restrict__ range,
                           const float* __restrict__ nlo, const float* __restrict__ nhi, LxWideNode* __restrict__ out,
                           int2* __restrict__ fout, uint32_t* __restrict__ counters /* [0] slots, [1] next frontier */) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nin) return;
    int2 f = fin[e];
    int kids[8];
    int nk = 2;
    kids[0] = child[f.x].x;
    kids[1] = child[f.x].y;
    while (nk < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; k++) {
            int c = kids[k];
            if (make_leaf(c, n, leaf_size, child, range, nlo, nhi)) continue;
            float a = area3(nlo + 3 * c, nhi + 3 * c);
            if (a > ba) {
                ba = a;
                best = k;
            }
        }
        if (best < 0) break;
        int c = kids[best];
        kids[best] = child[c].x;
        kids[nk++] = child[c].y;
    }
    LxWideNode w;
    w.n = (uint32_t)nk;
    for (int k = 0; k < 8; k++) {
        if (k >= nk) {
            for (int a = 0; a < 3; a++) w.lo[k][a] = w.hi[k][a] = 0.0f;
            w.child[k] = LX_CHILD_EMPTY;
            continue;
        }
        int c = kids[k];
        for (int a = 0; a < 3; a++) {
            float l = nlo[3 * c + a], h = nhi[3 * c + a];
            /* the host builder's padding: (max |coord| + extent) * 2^-19 */
            float m = fmaxf(fabsf(l), fabsf(h));
            float pad = (m + (h - l)) * 1.9073486e-6f + 1e-30f;
            w.lo[k][a] = l - pad;
            w.hi[k][a] = h + pad;
        }
        int cnt = count_of(c, n, range);
        if (make_leaf(c, n, leaf_size, child, range, nlo, nhi)) {
            int first = c >= n - 1 ? c - (n - 1) : range[c].x;
            w.child[k] = LX_LEAF_BIT | ((uint32_t)(cnt - 1) << 28) | (uint32_t)first;
        } else {
            uint32_t slot = atomicAdd(&counters[0], 1u);
            uint32_t q = atomicAdd(&counters[1], 1u);
            fout[q] = make_int2(c, (int)slot);
            w.child[k] = slot;
        }
    }
    out[f.y] = w;
}

/* a cluster record: a maximal subtree of at most K triangles */
struct LxCluster {
    int node, count, first, pad;
    float lo[3], hi[3];
};

__global__ void k_cluster_roots(int n, int K, const int2* __restrict__ range, const int* __restrict__ parent,
                                const float* __restrict__ nlo, const float* __restrict__ nhi,
                                LxCluster* __restrict__ out, uint32_t* __restrict__ counter) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * n - 1) return;
    int c = count_of(v, n, range);
    int p = parent[v];
    if (c > K || (p >= 0 && count_of(p, n, range) <= K)) return;
    uint32_t q = atomicAdd(counter, 1u);
    LxCluster r;
    r.node = v;
    r.count = c;
    r.first = v >= n - 1 ? v - (n - 1) : range[v].x;
    r.pad = 0;
    for (int a = 0; a < 3; a++) {
        r.lo[a] = nlo[3 * v + a];
        r.hi[a] = nhi[3 * v + a];
    }
    out[q] = r;
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

const uint32_t kClusterMark = 0x40000000u; /* child ref placeholder: frontier index */

/* Top of the tree over clusters: binned SAH with triangle-count weights. */
struct TopNode {
    float lo[3], hi[3];
    int left = -1, right = -1; /* top node ids; -1 for a cluster leaf */
    int cluster = -1;
};

float top_area(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

int build_top(std::vector<TopNode>& T, std::vector<LxCluster>& C, int b, int e, int bins) {
    int me = (int)T.size();
    T.emplace_back();
    TopNode nd;
    float clo[3], chi[3];
    for (int a = 0; a < 3; a++) {
        nd.lo[a] = clo[a] = 3.4e38f;
        nd.hi[a] = chi[a] = -3.4e38f;
    }
    for (int i = b; i < e; i++)
        for (int a = 0; a < 3; a++) {
            nd.lo[a] = std::min(nd.lo[a], C[i].lo[a]);
            nd.hi[a] = std::max(nd.hi[a], C[i].hi[a]);
            float c = 0.5f * (C[i].lo[a] + C[i].hi[a]);
            clo[a] = std::min(clo[a], c);
            chi[a] = std::max(chi[a], c);
        }
    if (e - b == 1) {
        nd.cluster = b;
        T[me] = nd;
        return me;
    }
    /* binned SAH: cost of a split = sum over sides of area * triangles */
    double best = 1e300;
    int bax = -1, bbin = -1;
    std::vector<double> cnt(bins), racc(bins);
    std::vector<float> blo(3 * bins), bhi(3 * bins);
    for (int a = 0; a < 3; a++) {
        float ext = chi[a] - clo[a];
        if (!(ext > 0.0f)) continue;
        std::fill(cnt.begin(), cnt.end(), 0.0);
        for (int k = 0; k < bins; k++)
            for (int q = 0; q < 3; q++) {
                blo[3 * k + q] = 3.4e38f;
                bhi[3 * k + q] = -3.4e38f;
            }
        for (int i = b; i < e; i++) {
            float c = 0.5f * (C[i].lo[a] + C[i].hi[a]);
            int k = std::min(bins - 1, std::max(0, (int)((c - clo[a]) / ext * bins)));
            cnt[k] += C[i].count;
            for (int q = 0; q < 3; q++) {
                blo[3 * k + q] = std::min(blo[3 * k + q], C[i].lo[q]);
                bhi[3 * k + q] = std::max(bhi[3 * k + q], C[i].hi[q]);
            }
        }
        /* right-to-left sweep */
        float rlo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, rhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        double rc = 0;
        for (int k = bins - 1; k > 0; k--) {
            for (int q = 0; q < 3; q++) {
                rlo[q] = std::min(rlo[q], blo[3 * k + q]);
                rhi[q] = std::max(rhi[q], bhi[3 * k + q]);
            }
            rc += cnt[k];
            racc[k] = rc > 0 ? top_area(rlo, rhi) * rc : 0.0;
        }
        float llo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, lhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        double lc = 0;
        for (int k = 0; k < bins - 1; k++) {
            for (int q = 0; q < 3; q++) {
                llo[q] = std::min(llo[q], blo[3 * k + q]);
                lhi[q] = std::max(lhi[q], bhi[3 * k + q]);
            }
            lc += cnt[k];
            if (lc == 0 || rc - lc + cnt[0] * 0 <= 0) continue;
            double cost = top_area(llo, lhi) * lc + racc[k + 1];
            double total_right = 0;
            for (int q = k + 1; q < bins; q++) total_right += cnt[q];
            if (total_right == 0) continue;
            if (cost < best) {
                best = cost;
                bax = a;
                bbin = k;
            }
        }
    }
    int mid;
    if (bax < 0) {
        mid = (b + e) / 2; /* all centroids equal: split the list */
    } else {
        float ext = chi[bax] - clo[bax];
        auto it = std::partition(C.begin() + b, C.begin() + e, [&](const LxCluster& c) {
            float cc = 0.5f * (c.lo[bax] + c.hi[bax]);
            int k = std::min(bins - 1, std::max(0, (int)((cc - clo[bax]) / ext * bins)));
            return k <= bbin;
        });
        mid = (int)(it - C.begin());
        if (mid == b || mid == e) mid = (b + e) / 2;
    }
    nd.left = build_top(T, C, b, mid, bins);
    nd.right = build_top(T, C, mid, e, bins);
    T[me] = nd;
    return me;
}

int top_count(const std::vector<TopNode>& T, const std::vector<LxCluster>& C, int t) {
    return T[t].cluster >= 0 ? C[T[t].cluster].count : top_count(T, C, T[t].left) + top_count(T, C, T[t].right);
}

/* Collapse the top tree to W-wide nodes on the host.  Clusters are not
 * expanded: a small one becomes a leaf, a larger one a wide node the device
 * collapse fills in (frontier entry: Karras node, slot). */
uint32_t collapse_top(const std::vector<TopNode>& T, const std::vector<LxCluster>& C, int t, int W, int leaf_size,
                      std::vector<LxWideNode>& top, std::vector<int2>& frontier, int depth, int& max_depth) {
    max_depth = std::max(max_depth, depth);
    uint32_t me = (uint32_t)top.size();
    top.emplace_back();
    int kids[8] = {T[t].left, T[t].right, -1, -1, -1, -1, -1, -1};
    int nk = 2;
    while (nk < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; k++) {
            const TopNode& c = T[kids[k]];
            if (c.cluster >= 0) continue;
            float a = top_area(c.lo, c.hi);
            if (a > ba) {
                ba = a;
                best = k;
            }
        }
        if (best < 0) break;
        int c = kids[best];
        kids[best] = T[c].left;
        kids[nk++] = T[c].right;
    }
    LxWideNode w;
    std::memset(&w, 0, sizeof(w));
    w.n = (uint32_t)nk;
    for (int k = 0; k < 8; k++) {
        if (k >= nk) {
            w.child[k] = LX_CHILD_EMPTY;
            continue;
        }
        const TopNode& c = T[kids[k]];
        for (int a = 0; a < 3; a++) {
            w.lo[k][a] = c.lo[a];
            w.hi[k][a] = c.hi[a];
        }
        lx_pad_box(w.lo[k], w.hi[k]);
        if (c.cluster < 0) {
            w.child[k] = collapse_top(T, C, kids[k], W, leaf_size, top, frontier, depth + 1, max_depth);
        } else {
            const LxCluster& cl = C[c.cluster];
            if (cl.count <= leaf_size) {
                w.child[k] = LX_LEAF_BIT | ((uint32_t)(cl.count - 1) << 28) | (uint32_t)cl.first;
            } else {
                /* a cluster wide node: numbered after the top nodes later */
                frontier.push_back(make_int2(cl.node, -1));
                w.child[k] = kClusterMark | (uint32_t)(frontier.size() - 1);
            }
        }
    }
    top[me] = w;
    return me;
}

} // namespace

bool lx_build_blas_gpu(lx_context* ctx, const float* pos, const uint32_t* idx, uint32_t ntri, int leaf_size, int bins,
                       int width, BvhBuildResult& out) {
    (void)bins;
    if (ntri == 0 || ntri >= (1u << 28) - 1) return false;
    if (leaf_size < 1) leaf_size = 1;
    if (leaf_size > 8) leaf_size = 8;
    int W = width == 8 ? 8 : 4;
    hipStream_t s = ctx->stream;
    uint32_t nv = 0;
    for (size_t k = 0; k < 3 * (size_t)ntri; k++) nv = std::max(nv, idx[k] + 1);
    int n = (int)ntri;
    DevBuf d_pos, d_idx, d_lo, d_hi, d_cbox, d_keys, d_vals, d_keys2, d_vals2, d_tmp, d_child, d_range, d_parent,
        d_nlo, d_nhi, d_flag, d_nodes, d_front[2], d_cnt, d_tris;
    d_pos.upload(ctx, pos, (size_t)nv * 12);
    d_idx.upload(ctx, idx, (size_t)ntri * 12);
    d_lo.alloc(ctx, (size_t)ntri * 12);
    d_hi.alloc(ctx, (size_t)ntri * 12);
    uint32_t cinit[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    d_cbox.upload(ctx, cinit, sizeof(cinit));
    uint32_t blocks = (ntri + 255) / 256;
    hipLaunchKernelGGL(k_tri_bounds, dim3(blocks), dim3(256), 0, s, d_pos.as<float>(), d_idx.as<uint32_t>(), ntri,
                       d_lo.as<float>(), d_hi.as<float>(), d_cbox.as<uint32_t>());
    LX_HIP(hipGetLastError());
    d_keys.alloc(ctx, (size_t)ntri * 4);
    d_vals.alloc(ctx, (size_t)ntri * 4);
    d_keys2.alloc(ctx, (size_t)ntri * 4);
    d_vals2.alloc(ctx, (size_t)ntri * 4);
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, s, d_lo.as<float>(), d_hi.as<float>(), ntri,
                       d_cbox.as<uint32_t>(), d_keys.as<uint32_t>(), d_vals.as<uint32_t>());
    LX_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    LX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_keys.as<uint32_t>(), d_keys2.as<uint32_t>(),
                                              d_vals.as<uint32_t>(), d_vals2.as<uint32_t>(), n, 0, 30, s));
    d_tmp.alloc(ctx, tmp_bytes + 16);
    LX_HIP(hipcub::DeviceRadixSort::SortPairs(d_tmp.ptr, tmp_bytes, d_keys.as<uint32_t>(), d_keys2.as<uint32_t>(),
                                              d_vals.as<uint32_t>(), d_vals2.as<uint32_t>(), n, 0, 30, s));
    int nint = n - 1;
    size_t nall = (size_t)(2 * n - 1);
    d_child.alloc(ctx, (size_t)std::max(1, nint) * 8);
    d_range.alloc(ctx, (size_t)std::max(1, nint) * 8);
    d_parent.alloc(ctx, nall * 4);
    d_nlo.alloc(ctx, nall * 12);
    d_nhi.alloc(ctx, nall * 12);
    d_flag.alloc(ctx, (size_t)std::max(1, nint) * 4);
    LX_HIP(hipMemsetAsync(d_flag.ptr, 0, (size_t)std::max(1, nint) * 4, s));
    LX_HIP(hipMemsetAsync(d_parent.ptr, 0xff, nall * 4, s));
    if (nint > 0) {
        hipLaunchKernelGGL(k_karras, dim3((nint + 255) / 256), dim3(256), 0, s, d_keys2.as<uint32_t>(), n,
                           d_child.as<int2>(), d_range.as<int2>(), d_parent.as<int>());
        LX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, s, d_vals2.as<uint32_t>(), d_lo.as<float>(),
                       d_hi.as<float>(), n, d_child.as<int2>(), d_parent.as<int>(), d_nlo.as<float>(),
                       d_nhi.as<float>(), d_flag.as<uint32_t>());
    LX_HIP(hipGetLastError());
    d_tris.alloc(ctx, (size_t)ntri * 48);
    hipLaunchKernelGGL(k_reorder, dim3(blocks), dim3(256), 0, s, d_pos.as<float>(), d_idx.as<uint32_t>(),
                       d_vals2.as<uint32_t>(), ntri, d_tris.as<float4>());
    LX_HIP(hipGetLastError());
    out.tris.resize(3 * (size_t)ntri);
    out.nodes.clear();
    out.depth = 0;
    out.sah = 0.0;
    if (n <= leaf_size) {
        /* one leaf holds the whole mesh */
        out.root = LX_LEAF_BIT | ((uint32_t)(n - 1) << 28);
        LX_HIP(hipMemcpyAsync(out.tris.data(), d_tris.ptr, (size_t)ntri * 48, hipMemcpyDeviceToHost, s));
        LX_HIP(hipStreamSynchronize(s));
        return true;
    }
    d_nodes.alloc(ctx, (size_t)std::max(1, nint) * sizeof(LxWideNode));
    d_front[0].alloc(ctx, (size_t)std::max(1, nint) * 8);
    d_front[1].alloc(ctx, (size_t)std::max(1, nint) * 8);
    d_cnt.alloc(ctx, 16);
    /* binned-SAH refinement of the top: clusters = maximal subtrees of at
     * most kCluster triangles; the tree above them is rebuilt with SAH */
    const int kCluster = 64;
    std::vector<LxWideNode> top;
    std::vector<int2> frontier;
    uint32_t next_slot = 1;
    int top_depth = 0;
    if (n > kCluster) {
        DevBuf d_cl, d_ccnt;
        d_cl.alloc(ctx, nall * sizeof(LxCluster));
        static const uint32_t kZero4[4] = {0, 0, 0, 0};
        d_ccnt.upload(ctx, kZero4, 16);
        hipLaunchKernelGGL(k_cluster_roots, dim3((uint32_t)((nall + 255) / 256)), dim3(256), 0, s, n, kCluster,
                           d_range.as<int2>(), d_parent.as<int>(), d_nlo.as<float>(), d_nhi.as<float>(),
                           d_cl.as<LxCluster>(), d_ccnt.as<uint32_t>());
        LX_HIP(hipGetLastError());
        uint32_t ncl = 0;
        LX_HIP(hipMemcpyAsync(&ncl, d_ccnt.ptr, 4, hipMemcpyDeviceToHost, s));
        LX_HIP(hipStreamSynchronize(s));
        std::vector<LxCluster> C(ncl);
        LX_HIP(hipMemcpyAsync(C.data(), d_cl.ptr, (size_t)ncl * sizeof(LxCluster), hipMemcpyDeviceToHost, s));
        LX_HIP(hipStreamSynchronize(s));
        std::sort(C.begin(), C.end(), [](const LxCluster& a, const LxCluster& b) { return a.first < b.first; });
        if (ncl >= 2) {
            std::vector<TopNode> T;
            T.reserve(2 * ncl);
            int troot = build_top(T, C, 0, (int)ncl, 32);
            collapse_top(T, C, troot, W, leaf_size, top, frontier, 0, top_depth);
            /* top nodes take slots 0 .. top.size() - 1, cluster wide nodes
             * the next frontier.size() slots, the device collapse the rest */
            uint32_t base = (uint32_t)top.size();
            for (size_t k = 0; k < frontier.size(); k++) frontier[k].y = (int)(base + k);
            for (auto& w : top)
                for (int k = 0; k < 8; k++) {
                    uint32_t c = w.child[k];
                    if (c == LX_CHILD_EMPTY || (c & LX_LEAF_BIT) || !(c & kClusterMark)) continue;
                    w.child[k] = base + (c & ~kClusterMark);
                }
            next_slot = base + (uint32_t)frontier.size();
        }
    }
    int cur = 0, depth = 0;
    uint32_t nin;
    uint32_t cnt[4] = {1u, 0u, 0u, 0u};
    if (top.empty()) {
        static const int2 kRoot = make_int2(0, 0);
        LX_HIP(hipMemcpyAsync(d_front[0].ptr, &kRoot, 8, hipMemcpyHostToDevice, s));
        nin = 1;
    } else {
        if (!frontier.empty())
            LX_HIP(hipMemcpyAsync(d_front[0].ptr, frontier.data(), frontier.size() * 8, hipMemcpyHostToDevice, s));
        nin = (uint32_t)frontier.size();
        cnt[0] = next_slot;
        depth = top_depth + 1;
    }
    LX_HIP(hipMemcpyAsync(d_cnt.ptr, cnt, 16, hipMemcpyHostToDevice, s));
    LX_HIP(hipStreamSynchronize(s));
    static const uint32_t kZero = 0;
    while (nin > 0) {
        LX_HIP(hipMemcpyAsync(d_cnt.as<uint32_t>() + 1, &kZero, 4, hipMemcpyHostToDevice, s));
        if (W == 8)
            hipLaunchKernelGGL(k_collapse<8>, dim3((nin + 127) / 128), dim3(128), 0, s, d_front[cur].as<int2>(), nin, n,
                               leaf_size, d_child.as<int2>(), d_range.as<int2>(), d_nlo.as<float>(),
                               d_nhi.as<float>(), d_nodes.as<LxWideNode>(), d_front[cur ^ 1].as<int2>(),
                               d_cnt.as<uint32_t>());
        else
            hipLaunchKernelGGL(k_collapse<4>, dim3((nin + 127) / 128), dim3(128), 0, s, d_front[cur].as<int2>(), nin, n,
                               leaf_size, d_child.as<int2>(), d_range.as<int2>(), d_nlo.as<float>(),
                               d_nhi.as<float>(), d_nodes.as<LxWideNode>(), d_front[cur ^ 1].as<int2>(),
                               d_cnt.as<uint32_t>());
        LX_HIP(hipGetLastError());
        LX_HIP(hipMemcpyAsync(cnt, d_cnt.ptr, 16, hipMemcpyDeviceToHost, s));
        LX_HIP(hipStreamSynchronize(s));
        nin = cnt[1];
        cur ^= 1;
        depth++;
    }
    uint32_t nodes = cnt[0];
    out.nodes.resize(nodes);
    LX_HIP(hipMemcpyAsync(out.nodes.data(), d_nodes.ptr, (size_t)nodes * sizeof(LxWideNode), hipMemcpyDeviceToHost, s));
    LX_HIP(hipMemcpyAsync(out.tris.data(), d_tris.ptr, (size_t)ntri * 48, hipMemcpyDeviceToHost, s));
    LX_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < top.size(); k++) out.nodes[k] = top[k];
    out.root = 0;
    out.depth = depth;
    lx_wide_dfs_order(out.nodes, out.root);
    out.sah = lx_wide_sah(out.nodes, out.root);
    return true;
}
