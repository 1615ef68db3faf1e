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
#include <atomic>
#include <cmath>
#include <cfloat>
#include <functional>
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
     * triangle test's epsilon-inflated acceptance region */
    for (int a = 0; a < 3; a++) {
        float m = std::max(std::fabs(lo[a]), std::fabs(hi[a]));
        float pad = (m + (hi[a] - lo[a])) * 1.9073486e-6f + 1e-30f; /* 2^-19 */
        lo[a] = lo[a] - pad;
        hi[a] = hi[a] + pad;
    }
}

/* Collapse the binary tree into W-wide nodes (W = 4 or 8): each wide node
 * adopts the children of its largest-area internal children until it has W. */
static uint32_t collapse(const std::vector<BNode>& bn, int32_t bi, int W, std::vector<LxWideNode>& out,
                         const std::function<uint32_t(uint32_t, uint32_t)>& leaf_ref, int depth, int& max_depth,
                         double& sah, double root_area) {
    if (depth > max_depth) max_depth = depth;
    const BNode& b = bn[bi];
    if (b.left < 0) {
        sah += (double)area(b.lo, b.hi) / root_area * (double)b.count;
        return leaf_ref(b.first, b.count);
    }
    sah += (double)area(b.lo, b.hi) / root_area;
    int32_t kids[8] = {b.left, b.right, -1, -1, -1, -1, -1, -1};
    int nk = 2;
    while (nk < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; k++) {
            const BNode& c = bn[kids[k]];
            if (c.left < 0) continue;
            float a = area(c.lo, c.hi);
            if (a > ba) {
                ba = a;
                best = k;
            }
        }
        if (best < 0) break;
        int32_t c = kids[best];
        kids[best] = bn[c].left;
        kids[nk++] = bn[c].right;
    }
    uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    LxWideNode n;
    std::memset(&n, 0, sizeof(n));
    n.n = (uint32_t)nk;
    for (int k = 0; k < 8; k++) {
        if (k < nk) {
            const BNode& c = bn[kids[k]];
            for (int a = 0; a < 3; a++) {
                n.lo[k][a] = c.lo[a];
                n.hi[k][a] = c.hi[a];
            }
            lx_pad_box(n.lo[k], n.hi[k]);
            n.child[k] = collapse(bn, kids[k], W, out, leaf_ref, depth + 1, max_depth, sah, root_area);
        } else {
            n.child[k] = LX_CHILD_EMPTY;
        }
    }
    out[me] = n;
    return me;
}

double lx_wide_sah(const std::vector<LxWideNode>& nodes, uint32_t root) {
    if (nodes.empty() || (root & LX_LEAF_BIT)) return 0.0;
    auto node_box = [&](uint32_t i, float* lo, float* hi) {
        const LxWideNode& w = nodes[i];
        for (int a = 0; a < 3; a++) {
            lo[a] = FLT_MAX;
            hi[a] = -FLT_MAX;
        }
        for (uint32_t k = 0; k < 8; k++) {
            if (w.child[k] == LX_CHILD_EMPTY) continue;
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], w.lo[k][a]);
                hi[a] = std::max(hi[a], w.hi[k][a]);
            }
        }
    };
    float rlo[3], rhi[3];
    node_box(root, rlo, rhi);
    double ra = (double)area(rlo, rhi);
    if (!(ra > 0.0)) return 0.0;
    double sah = 0.0;
    std::vector<uint32_t> todo{root};
    while (!todo.empty()) {
        uint32_t i = todo.back();
        todo.pop_back();
        float lo[3], hi[3];
        node_box(i, lo, hi);
        sah += (double)area(lo, hi) / ra;
        const LxWideNode& w = nodes[i];
        for (uint32_t k = 0; k < 8; k++) {
            uint32_t c = w.child[k];
            if (c == LX_CHILD_EMPTY) continue;
            if (c & LX_LEAF_BIT) sah += (double)area(w.lo[k], w.hi[k]) / ra * (double)(((c >> 28) & 7u) + 1u);
            else todo.push_back(c);
        }
    }
    return sah;
}

void lx_wide_dfs_order(std::vector<LxWideNode>& nodes, uint32_t& root) {
    if (nodes.empty() || (root & LX_LEAF_BIT)) return;
    std::vector<uint32_t> order;
    order.reserve(nodes.size());
    std::vector<uint32_t> todo{root};
    while (!todo.empty()) {
        uint32_t i = todo.back();
        todo.pop_back();
        order.push_back(i);
        const LxWideNode& w = nodes[i];
        for (int k = 7; k >= 0; k--) { /* slot 0 is visited first */
            uint32_t c = w.child[k];
            if (c != LX_CHILD_EMPTY && !(c & LX_LEAF_BIT)) todo.push_back(c);
        }
    }
    std::vector<uint32_t> newid(nodes.size(), LX_CHILD_EMPTY);
    for (uint32_t k = 0; k < (uint32_t)order.size(); k++) newid[order[k]] = k;
    std::vector<LxWideNode> out(order.size());
    for (uint32_t k = 0; k < (uint32_t)order.size(); k++) {
        LxWideNode w = nodes[order[k]];
        for (int c = 0; c < 8; c++)
            if (w.child[c] != LX_CHILD_EMPTY && !(w.child[c] & LX_LEAF_BIT)) w.child[c] = newid[w.child[c]];
        out[k] = w;
    }
    nodes.swap(out);
    root = 0;
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
    out.root = collapse(B.nodes, root, width == 8 ? 8 : 4, out.nodes, make_leaf_ref, 0, md, sah, ra);
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
    out.root = collapse(B.nodes, root, width == 8 ? 8 : 4, out.nodes, ref, 0, md, sah, ra);
    out.depth = md;
    out.sah = sah;
}
