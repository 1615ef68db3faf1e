/*
 * lx_bvh_host.cpp -- host binned-SAH builder and the binary -> 4- or
 * 8-wide collapse shared by the host and GPU builders.
 *
 * Replaces the reference's per-mesh OptiX GAS build (src/myoptix/optix_accel.cpp:9-110)
 * and IAS build (src/myoptix/optix_ias.cpp:9-63).  The reference's own
 * software BVH (src/mesh/bvh_build.cu:15-318, binary, one triangle per
 * leaf) is restated in the oracle; this builder produces the wide,
 * multi-triangle-leaf layout the gfx950 traversal reads.
 */
#include "lx_internal.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cfloat>
#include <thread>

namespace {

struct Prim {
    float lo[3], hi[3], c[3];
    uint32_t id;
};

struct BNode {
    float lo[3], hi[3];
    int32_t left, right; /* -1 for leaves */
    uint32_t first, count;
};

static inline float area(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct Builder {
    std::vector<Prim>& prims;
    std::vector<BNode> nodes;
    std::atomic<int32_t> next{0};
    int leaf_size;
    int bins;
    std::atomic<int> max_depth{0};

    Builder(std::vector<Prim>& p, int ls, int nb) : prims(p), leaf_size(ls), bins(nb) {
        nodes.resize(std::max<size_t>(1, 2 * p.size()));
    }

    int32_t alloc() { return next.fetch_add(1); }

    void bounds(uint32_t b, uint32_t e, float* lo, float* hi, float* clo, float* chi) {
        for (int a = 0; a < 3; a++) {
            lo[a] = clo[a] = FLT_MAX;
            hi[a] = chi[a] = -FLT_MAX;
        }
        for (uint32_t i = b; i < e; i++) {
            const Prim& p = prims[i];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], p.lo[a]);
                hi[a] = std::max(hi[a], p.hi[a]);
                clo[a] = std::min(clo[a], p.c[a]);
                chi[a] = std::max(chi[a], p.c[a]);
            }
        }
    }

    void make_leaf(int32_t ni, uint32_t b, uint32_t e, const float* lo, const float* hi) {
        BNode& n = nodes[ni];
        for (int a = 0; a < 3; a++) {
            n.lo[a] = lo[a];
            n.hi[a] = hi[a];
        }
        n.left = n.right = -1;
        n.first = b;
        n.count = e - b;
    }

    /* Builds [b, e) into node ni.  Returns nothing; children recurse. */
    void build(int32_t ni, uint32_t b, uint32_t e, int depth) {
        int md = max_depth.load();
        while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
        }
        float lo[3], hi[3], clo[3], chi[3];
        bounds(b, e, lo, hi, clo, chi);
        uint32_t n = e - b;
        if (n <= 1) {
            make_leaf(ni, b, e, lo, hi);
            return;
        }
        /* binned SAH over all three axes */
        const int NB = bins;
        float best_cost = FLT_MAX;
        int best_axis = -1, best_split = -1;
        std::vector<float> blo(NB * 3), bhi(NB * 3);
        std::vector<uint32_t> bcnt(NB);
        std::vector<float> rarea(NB);
        std::vector<uint32_t> rcnt(NB);
        for (int a = 0; a < 3; a++) {
            float ext = chi[a] - clo[a];
            if (!(ext > 0.0f)) continue;
            float scale = (float)NB / ext;
            for (int k = 0; k < NB; k++) {
                bcnt[k] = 0;
                for (int c = 0; c < 3; c++) {
                    blo[k * 3 + c] = FLT_MAX;
                    bhi[k * 3 + c] = -FLT_MAX;
                }
            }
            for (uint32_t i = b; i < e; i++) {
                const Prim& p = prims[i];
                int k = (int)((p.c[a] - clo[a]) * scale);
                k = k < 0 ? 0 : (k >= NB ? NB - 1 : k);
                bcnt[k]++;
                for (int c = 0; c < 3; c++) {
                    blo[k * 3 + c] = std::min(blo[k * 3 + c], p.lo[c]);
                    bhi[k * 3 + c] = std::max(bhi[k * 3 + c], p.hi[c]);
                }
            }
            float rl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint32_t rc = 0;
            for (int k = NB - 1; k >= 1; k--) {
                rc += bcnt[k];
                for (int c = 0; c < 3; c++) {
                    rl[c] = std::min(rl[c], blo[k * 3 + c]);
                    rh[c] = std::max(rh[c], bhi[k * 3 + c]);
                }
                rcnt[k] = rc;
                rarea[k] = area(rl, rh);
            }
            float ll[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint32_t lc = 0;
            for (int k = 0; k < NB - 1; k++) {
                lc += bcnt[k];
                for (int c = 0; c < 3; c++) {
                    ll[c] = std::min(ll[c], blo[k * 3 + c]);
                    lh[c] = std::max(lh[c], bhi[k * 3 + c]);
                }
                if (lc == 0 || rcnt[k + 1] == 0) continue;
                float cost = area(ll, lh) * (float)lc + rarea[k + 1] * (float)rcnt[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = k;
                }
            }
        }
        float parea = area(lo, hi);
        /* leaf if small enough and splitting does not pay (C_trav = 1, C_tri = 1 per area unit) */
        if ((int)n <= leaf_size && (best_axis < 0 || parea * (float)n <= parea + best_cost)) {
            make_leaf(ni, b, e, lo, hi);
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = b + n / 2; /* all centroids coincide: split by index */
        } else {
            float ext = chi[best_axis] - clo[best_axis];
            float scale = (float)NB / ext;
            auto it = std::partition(prims.begin() + b, prims.begin() + e, [&](const Prim& p) {
                int k = (int)((p.c[best_axis] - clo[best_axis]) * scale);
                k = k < 0 ? 0 : (k >= NB ? NB - 1 : k);
                return k <= best_split;
            });
            mid = (uint32_t)(it - prims.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        }
        BNode& nd = nodes[ni];
        for (int a = 0; a < 3; a++) {
            nd.lo[a] = lo[a];
            nd.hi[a] = hi[a];
        }
        int32_t l = alloc(), r = alloc();
        nd.left = l;
        nd.right = r;
        nd.first = 0;
        nd.count = 0;
        if (n > 65536 && depth < 5) {
            std::thread t([&, l, b, mid, depth]() { build(l, b, mid, depth + 1); });
            build(r, mid, e, depth + 1);
            t.join();
        } else {
            build(l, b, mid, depth + 1);
            build(r, mid, e, depth + 1);
        }
    }
};

} // namespace

void lx_pad_box(float lo[3], float hi[3]) {
    /* conservative padding: covers the slab test's rounding and the
     * triangle test's epsilon-inflated acceptance region; the 2^-60 floor
     * keeps a flat box's slab wider than 2^-100 t for the traversal's
     * clamped reciprocal of a zero direction component */
    for (int a = 0; a < 3; a++) {
        float m = std::max(std::fabs(lo[a]), std::fabs(hi[a]));
        float pad = (m + (hi[a] - lo[a])) * 1.9073486e-6f + 8.6736174e-19f; /* 2^-19, 2^-60 */
        lo[a] = lo[a] - pad;
        hi[a] = hi[a] + pad;
    }
}

/* The wide node over binary internal node b: starting from b's two
 * children, the largest-area internal child is replaced by its own two
 * children until there are W (or no internal child is left).  Returns the
 * number of children. */
static int wide_kids(const std::vector<BNode>& bn, const BNode& b, int W, int32_t kids[8]) {
    float karea[8];
    kids[0] = b.left;
    kids[1] = b.right;
    karea[0] = area(bn[b.left].lo, bn[b.left].hi);
    karea[1] = area(bn[b.right].lo, bn[b.right].hi);
    int nk = 2;
    while (nk < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; k++) {
            if (bn[kids[k]].left < 0) continue;
            if (karea[k] > ba) {
                ba = karea[k];
                best = k;
            }
        }
        if (best < 0) break;
        const BNode& c = bn[kids[best]];
        kids[best] = c.left;
        karea[best] = area(bn[c.left].lo, bn[c.left].hi);
        kids[nk] = c.right;
        karea[nk++] = area(bn[c.right].lo, bn[c.right].hi);
    }
    return nk;
}

/* Starts wide node `me` of `out` for binary internal node b: child count,
 * padded child boxes, empty slots; the child refs of used slots are left to
 * the caller. */
static int begin_wide(const std::vector<BNode>& bn, const BNode& b, int W, LxWideNode& n, int32_t kids[8]) {
    int nk = wide_kids(bn, b, W, kids);
    n.n = (uint32_t)nk;
    for (int k = 0; k < 8; k++) {
        if (k < nk) {
            const BNode& c = bn[kids[k]];
            for (int a = 0; a < 3; a++) {
                n.lo[k][a] = c.lo[a];
                n.hi[k][a] = c.hi[a];
            }
            lx_pad_box(n.lo[k], n.hi[k]);
        } else {
            for (int a = 0; a < 3; a++) n.lo[k][a] = n.hi[k][a] = 0.0f;
            n.child[k] = LX_CHILD_EMPTY;
        }
    }
    return nk;
}

struct CollapseSums {
    int max_depth = 0;
    double sah = 0.0;
};

/* Appends the wide subtree of binary node bi to `out` in depth-first
 * pre-order (a node, then its children's subtrees in slot order); returns
 * its child ref (a node index of `out`, or a leaf ref). */
template <class LeafRef>
static uint32_t collapse_rec(const std::vector<BNode>& bn, int32_t bi, int W, std::vector<LxWideNode>& out,
                             const LeafRef& leaf_ref, int depth, CollapseSums& acc, double inv_root_area) {
    if (depth > acc.max_depth) acc.max_depth = depth;
    const BNode& b = bn[bi];
    if (b.left < 0) {
        acc.sah += (double)area(b.lo, b.hi) * inv_root_area * (double)b.count;
        return leaf_ref(b.first, b.count);
    }
    acc.sah += (double)area(b.lo, b.hi) * inv_root_area;
    uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    int32_t kids[8];
    int nk = begin_wide(bn, b, W, out[me], kids);
    for (int k = 0; k < nk; k++) {
        uint32_t ref = collapse_rec(bn, kids[k], W, out, leaf_ref, depth + 1, acc, inv_root_area);
        out[me].child[k] = ref; /* `out` may have grown: index again */
    }
    return me;
}

/* The top of a large tree, collapsed serially down to wide depth `cut`; the
 * subtrees below the cut become tasks.  `seq` lists the top nodes and tasks
 * in the depth-first pre-order the serial collapse would write them in
 * (k >= 0: top node k; k < 0: task -k - 1). */
struct CollapseTop {
    std::vector<LxWideNode> nodes;
    std::vector<std::array<int32_t, 8>> task_of; /* per slot: the task it refers to, or -1 */
    std::vector<std::pair<int32_t, int>> tasks;  /* binary node, wide depth */
    std::vector<int64_t> seq;
    CollapseSums acc;
};

template <class LeafRef>
static uint32_t collapse_top(const std::vector<BNode>& bn, int32_t bi, int W, const LeafRef& leaf_ref, int depth,
                             int cut, CollapseTop& T, double inv_root_area, int32_t& task) {
    task = -1;
    const BNode& b = bn[bi];
    if (b.left >= 0 && depth >= cut) {
        task = (int32_t)T.tasks.size();
        T.tasks.push_back({bi, depth});
        T.seq.push_back(-(int64_t)task - 1);
        return 0u;
    }
    if (depth > T.acc.max_depth) T.acc.max_depth = depth;
    if (b.left < 0) {
        T.acc.sah += (double)area(b.lo, b.hi) * inv_root_area * (double)b.count;
        return leaf_ref(b.first, b.count);
    }
    T.acc.sah += (double)area(b.lo, b.hi) * inv_root_area;
    uint32_t me = (uint32_t)T.nodes.size();
    T.nodes.emplace_back();
    T.task_of.emplace_back();
    T.task_of[me].fill(-1);
    T.seq.push_back((int64_t)me);
    int32_t kids[8];
    int nk = begin_wide(bn, b, W, T.nodes[me], kids);
    for (int k = 0; k < nk; k++) {
        int32_t t;
        uint32_t ref = collapse_top(bn, kids[k], W, leaf_ref, depth + 1, cut, T, inv_root_area, t);
        T.nodes[me].child[k] = ref;
        T.task_of[me][k] = t;
    }
    return me;
}

/* Collapse the binary tree into W-wide nodes (W = 4 or 8), nodes in
 * depth-first pre-order.  Large trees are collapsed on host threads: the
 * top levels serially, the subtrees below them concurrently, each into a
 * node array of its own, then the arrays are placed where the serial
 * pre-order puts them and their child refs offset -- the result is the
 * serial collapse's node for node. */
template <class LeafRef>
static uint32_t collapse(const std::vector<BNode>& bn, int32_t root, int W, std::vector<LxWideNode>& out,
                         const LeafRef& leaf_ref, int& max_depth, double& sah, double root_area) {
    const double inv_ra = 1.0 / root_area;
    out.clear();
    if (bn.size() < (1u << 17)) {
        out.reserve(bn.size() / 2 + 1); /* at most one wide node per internal binary node */
        CollapseSums acc;
        uint32_t r = collapse_rec(bn, root, W, out, leaf_ref, 0, acc, inv_ra);
        max_depth = acc.max_depth;
        sah = acc.sah;
        return r;
    }
    CollapseTop T;
    const int cut = W == 8 ? 3 : 4; /* up to 512 or 256 tasks */
    int32_t root_task;
    uint32_t r = collapse_top(bn, root, W, leaf_ref, 0, cut, T, inv_ra, root_task);
    const size_t nt = T.tasks.size();
    std::vector<std::vector<LxWideNode>> sub(nt);
    std::vector<CollapseSums> sums(nt);
    lx_parallel_for((uint32_t)nt, [&](uint32_t t) {
        sub[t].reserve(64);
        collapse_rec(bn, T.tasks[t].first, W, sub[t], leaf_ref, T.tasks[t].second, sums[t], inv_ra);
    });
    /* final positions, in the serial pre-order */
    std::vector<uint32_t> top_at(T.nodes.size()), task_at(nt);
    uint64_t pos = 0;
    for (int64_t e : T.seq) {
        if (e >= 0) top_at[(size_t)e] = (uint32_t)pos++;
        else {
            task_at[(size_t)(-e - 1)] = (uint32_t)pos;
            pos += sub[(size_t)(-e - 1)].size();
        }
    }
    LX_REQUIRE(pos < LX_LEAF_BIT, LX_ERR_INVALID, "BVH too large for the node encoding");
    out.resize((size_t)pos);
    for (size_t k = 0; k < T.nodes.size(); k++) {
        LxWideNode n = T.nodes[k];
        for (int c = 0; c < 8; c++) {
            if (n.child[c] == LX_CHILD_EMPTY) continue;
            if (T.task_of[k][c] >= 0) n.child[c] = task_at[(size_t)T.task_of[k][c]]; /* a task's root is its node 0 */
            else if (!(n.child[c] & LX_LEAF_BIT)) n.child[c] = top_at[n.child[c]];
        }
        out[top_at[k]] = n;
    }
    lx_parallel_for((uint32_t)nt, [&](uint32_t t) {
        const uint32_t base = task_at[t];
        const std::vector<LxWideNode>& v = sub[t];
        for (size_t k = 0; k < v.size(); k++) {
            LxWideNode n = v[k];
            for (int c = 0; c < 8; c++)
                if (n.child[c] != LX_CHILD_EMPTY && !(n.child[c] & LX_LEAF_BIT)) n.child[c] += base;
            out[base + k] = n;
        }
    });
    max_depth = T.acc.max_depth;
    sah = T.acc.sah;
    for (size_t t = 0; t < nt; t++) {
        max_depth = std::max(max_depth, sums[t].max_depth);
        sah += sums[t].sah;
    }
    if (root_task >= 0) return task_at[(size_t)root_task]; /* cut == 0 only */
    return (r & LX_LEAF_BIT) ? r : top_at[r];
}

static uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return LX_LEAF_BIT | ((count - 1u) << 28) | first;
}

void lx_build_blas_host(const float* pos, const uint32_t* idx, uint32_t ntri, int leaf_size, int bins, int width,
                        BvhBuildResult& out) {
    LX_REQUIRE(ntri > 0, LX_ERR_INVALID, "mesh has no triangles");
    LX_REQUIRE(ntri < (1u << 28) - 1, LX_ERR_INVALID, "mesh too large for the leaf encoding");
    if (leaf_size < 1) leaf_size = 1;
    if (leaf_size > 8) leaf_size = 8;
    if (bins < 4) bins = 4;
    if (bins > 64) bins = 64;
    std::vector<Prim> prims(ntri);
    for (uint32_t t = 0; t < ntri; t++) {
        Prim& p = prims[t];
        for (int a = 0; a < 3; a++) {
            float v0 = pos[3 * idx[3 * t] + a], v1 = pos[3 * idx[3 * t + 1] + a], v2 = pos[3 * idx[3 * t + 2] + a];
            p.lo[a] = std::min(v0, std::min(v1, v2));
            p.hi[a] = std::max(v0, std::max(v1, v2));
            p.c[a] = 0.5f * (p.lo[a] + p.hi[a]);
        }
        p.id = t;
    }
    Builder B(prims, leaf_size, bins);
    int32_t root = B.alloc();
    B.build(root, 0, ntri, 0);
    B.nodes.resize(B.next.load());
    /* leaf-ordered triangles: v0 | prim, e1 = v1 - v0, e2 = v2 - v0 */
    out.tris.resize(3 * (size_t)ntri);
    for (uint32_t k = 0; k < ntri; k++) {
        uint32_t t = prims[k].id;
        const float* a = pos + 3 * idx[3 * t];
        const float* b = pos + 3 * idx[3 * t + 1];
        const float* c = pos + 3 * idx[3 * t + 2];
        float pw;
        std::memcpy(&pw, &t, 4);
        out.tris[3 * k + 0] = float4{a[0], a[1], a[2], pw};
        out.tris[3 * k + 1] = float4{b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.0f};
        out.tris[3 * k + 2] = float4{c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.0f};
    }
    out.nodes.clear();
    int md = 0;
    double sah = 0.0;
    double ra = std::max(1e-30, (double)area(B.nodes[root].lo, B.nodes[root].hi));
    out.root = collapse(B.nodes, root, width == 8 ? 8 : 4, out.nodes, make_leaf_ref, md, sah, ra);
    out.depth = md;
    out.sah = sah;
}

void lx_build_tlas_host(const std::vector<float>& boxes, int width, BvhBuildResult& out) {
    uint32_t n = (uint32_t)(boxes.size() / 6);
    LX_REQUIRE(n > 0, LX_ERR_INVALID, "scene has no instances");
    std::vector<Prim> prims(n);
    for (uint32_t i = 0; i < n; i++) {
        Prim& p = prims[i];
        for (int a = 0; a < 3; a++) {
            p.lo[a] = boxes[6 * i + a];
            p.hi[a] = boxes[6 * i + 3 + a];
            p.c[a] = 0.5f * (p.lo[a] + p.hi[a]);
        }
        p.id = i;
    }
    Builder B(prims, 1, 64); /* instances are few: finer bins are cheap */
    int32_t root = B.alloc();
    B.build(root, 0, n, 0);
    B.nodes.resize(B.next.load());
    out.nodes.clear();
    out.tris.clear();
    int md = 0;
    double sah = 0.0;
    double ra = std::max(1e-30, (double)area(B.nodes[root].lo, B.nodes[root].hi));
    auto ref = [&](uint32_t first, uint32_t count) {
        (void)count;
        return make_leaf_ref(prims[first].id, 1);
    };
    out.root = collapse(B.nodes, root, width == 8 ? 8 : 4, out.nodes, ref, md, sah, ra);
    out.depth = md;
    out.sah = sah;
}
