/*
 * lx_trace_dev.h -- software two-level BVH traversal for gfx950.
 *
 * Replaces OptiX's hardware traversal (optixTrace in
 * src/optixshaders/optix_triangle.cu:68-135 of the reference) with:
 *  - a TLAS of BVH4/BVH8 nodes over instance world boxes, one instance per
 *    leaf; entering a leaf maps the ray to object space once and pushes a
 *    sentinel that restores the world ray on the way back;
 *  - per-mesh BLAS trees with up to 8 triangles per leaf, triangles stored in
 *    leaf order as (v0 | prim, e1, e2) so one leaf is a run of coalescable
 *    16-byte loads;
 *  - a per-lane short stack in LDS (column per lane: conflict-free for a
 *    wave64), spilling to a per-lane global region past its depth;
 *  - robust slab tests: the near and far planes of each axis are picked by
 *    the sign of the ray direction when they are loaded (per-lane byte
 *    offsets into the node), t = fma(plane, 1/d, -o/d) with o/d rounded
 *    outward, the far distance inflated by 2^-21 relative, and culling
 *    against the best hit only beyond a 2^-16 relative margin, so culling
 *    never changes the result of the triangle test; the closest hit is the
 *    lexicographic minimum of (t, instance, primitive) over all accepted
 *    candidates and does not depend on traversal order or tree shape.
 * The triangle test is the reference's software Moller-Trumbore
 * (src/intersections/mesh_isect.cu:77-137) computed in object space.
 */
#ifndef LX_TRACE_DEV_H
#define LX_TRACE_DEV_H

#include "lx_shade_dev.h"

#include "lx_kernels.h"

struct LxHitD {
    float t;
    uint32_t inst, prim;
    float u, v;
};

/* hash_u32, optix_triangle.cu:41-49 */
LXD uint32_t lx_hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* __anyhit__ms_all, optix_triangle.cu:317-366.  Keyed on the pixel index
 * (deterministic) instead of the compacted launch index. */
LXD bool lx_any_hit_accept(const LxDevScene& S, const LX_GLOBAL LxInstance& I, uint32_t inst, uint32_t prim, float bu,
                           float bv, f3 wdir, uint32_t key, uint32_t iteration) {
    const lx_material_desc& mat = S.mats[I.material];
    if (mat.alpha_mode == LX_ALPHA_OPAQUE) return true;
    float alpha = mat.alpha;
    if (mat.albedo_tex >= 0) {
        f2 uv = mk2(0.0f, 0.0f);
        if (I.has_n_uv & 2u) {
            const LX_GLOBAL uint32_t* ix = lx_gptr(S.idx) + I.idx_off + 3u * prim;
            const LX_GLOBAL float* UV = lx_gptr(S.uv) + 2u * (size_t)I.vtx_off;
            f2 A = mk2(UV[2 * ix[0]], UV[2 * ix[0] + 1]);
            f2 B = mk2(UV[2 * ix[1]], UV[2 * ix[1] + 1]);
            f2 C = mk2(UV[2 * ix[2]], UV[2 * ix[2] + 1]);
            float bA = 1.0f - bu - bv;
            uv = mk2((bA * A.x + bu * B.x) + bv * C.x, (bA * A.y + bu * B.y) + bv * C.y);
        }
        uv = mk2(uv.x * mat.albedo_xf.scale[0], uv.y * mat.albedo_xf.scale[1]);
        if (mat.albedo_xf.rotation != 0.0f) {
            float c = lxm_cos(mat.albedo_xf.rotation), s = lxm_sin(mat.albedo_xf.rotation);
            uv = mk2(c * uv.x - s * uv.y, s * uv.x + c * uv.y);
        }
        uv = mk2(uv.x + mat.albedo_xf.offset[0], uv.y + mat.albedo_xf.offset[1]);
        float t[4];
        lx_tex_sample(S.texs[mat.albedo_tex], uv.x, uv.y, t);
        alpha *= t[3];
    }
    if (mat.alpha_mode == LX_ALPHA_MASK) return !(alpha < mat.alpha_cutoff);
    uint32_t h = lx_hash_u32(key ^ lx_hash_u32(iteration));
    h = lx_hash_u32(h ^ lxm_f2u(wdir.x));
    h = lx_hash_u32(h ^ lxm_f2u(wdir.y));
    h = lx_hash_u32(h ^ lxm_f2u(wdir.z));
    h = lx_hash_u32(h ^ inst);
    h = lx_hash_u32(h ^ prim);
    float rnd = (h >> 8) * (1.0f / 16777216.0f);
    return !(rnd >= alpha);
}

/* Moller-Trumbore with the reference's epsilon bounds (mesh_isect.cu:77-137). */
LXD bool lx_mt(f3 o, f3 d, const float4& A, const float4& E1, const float4& E2, float tmax, float& t, float& u,
               float& v) {
    const float eps = (float)(1.1920929E-7F);
    f3 e1 = mk3(E1.x, E1.y, E1.z), e2 = mk3(E2.x, E2.y, E2.z);
    f3 c = cross(d, e2);
    float det = dot(e1, c);
    if (det > -eps && det < eps) return false;
    float inv_det = 1.0f / det;
    f3 a = o - mk3(A.x, A.y, A.z);
    u = inv_det * dot(a, c);
    if (u < -eps || u > 1.0f + eps) return false;
    c = cross(a, e1);
    v = inv_det * dot(d, c);
    if ((v < 0 && lxm_abs(v) > eps) || (u + v > 1 && lxm_abs(u + v - 1) > eps)) return false;
    t = inv_det * dot(e2, c);
    if (!(t > eps)) return false;
    if (!(t < tmax)) return false;
    return true;
}

LXD bool lx_better(float t, uint32_t inst, uint32_t prim, const LxHitD& h) {
    if (t < h.t) return true;
    if (t == h.t) {
        if (inst < h.inst) return true;
        if (inst == h.inst && prim < h.prim) return true;
    }
    return false;
}

#define LX_FAR_INFLATE 1.00000048f  /* 1 + 2^-21 */
#define LX_CULL_MARGIN 1.00001526f  /* 1 + 2^-16 */

struct LxStack {
    LX_LDS uint32_t* lds;     /* this lane's column: entry k at lds[k * stride] */
    uint32_t stride;
    LX_GLOBAL uint32_t* spill; /* this lane's spill column: entry k at spill[k * sstride] */
    uint32_t sstride;
    int sp;
    LXD void push(uint32_t v) {
        if (sp < LX_STACK) lds[sp * stride] = v;
        else spill[(sp - LX_STACK) * sstride] = v;
        sp++;
    }
    LXD uint32_t pop() {
        sp--;
        return sp < LX_STACK ? lds[sp * stride] : spill[(sp - LX_STACK) * sstride];
    }
};

template <int W>
struct LxNodeT;
template <>
struct LxNodeT<4> {
    typedef LxNode4 type;
};
template <>
struct LxNodeT<8> {
    typedef LxNode8 type;
};

/* Slab-test constants of a ray (o, d): ix = 1/d per axis, clamped to
 * +-2^100 (+-2^60 on an axis where |o| >= 2^26, so that o * ix stays
 * finite), and o * ix rounded outward by 2^-22 of its magnitude: pn up, pf
 * down.  Then fma(plane, ix, -pn) <= (plane - o) * ix <= fma(plane, ix, -pf)
 * up to the fma's own rounding (2^-24 relative, covered by the far
 * inflation of 2^-21).  A clamped axis (|d| < 2^-100, d = 0 included) tests
 * as a direction of 2^-100: its slab differs from the exact one only when
 * the origin lies within 2^-100 t of a face, i.e. inside the box's padding
 * (at least 2^-60), where no triangle of the box is accepted. */
struct LxSlab {
    f3 ix, pn, pf;
};

LXD float lx_slab_clamp(float ix, float o) {
    const float lim = lxm_abs(o) < 67108864.0f ? 0x1p100f : 0x1p60f;
    return __builtin_fminf(__builtin_fmaxf(ix, -lim), lim);
}

/* Reciprocal direction for the slab test: the hardware reciprocal (1 ulp;
 * +-inf for a zero component).  Its error scales an axis's entry and exit
 * distances alike and differs between axes by at most 2^-22 relative, which
 * the far inflation of 2^-21 covers with the fma's rounding. */
LXD f3 lx_rcp3(f3 d) {
    return mk3(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z));
}

/* iv = lx_rcp3(d) */
LXD LxSlab lx_slab(f3 o, f3 iv) {
    LxSlab s;
    s.ix = mk3(lx_slab_clamp(iv.x, o.x), lx_slab_clamp(iv.y, o.y), lx_slab_clamp(iv.z, o.z));
    const float px = o.x * s.ix.x, py = o.y * s.ix.y, pz = o.z * s.ix.z;
    s.pn = mk3(__builtin_fmaf(lxm_abs(px), 0x1p-22f, px), __builtin_fmaf(lxm_abs(py), 0x1p-22f, py),
               __builtin_fmaf(lxm_abs(pz), 0x1p-22f, pz));
    s.pf = mk3(__builtin_fmaf(-lxm_abs(px), 0x1p-22f, px), __builtin_fmaf(-lxm_abs(py), 0x1p-22f, py),
               __builtin_fmaf(-lxm_abs(pz), 0x1p-22f, pz));
    return s;
}

/* Slab test of the W children of node `node`; writes entry distances and
 * child refs and returns a W-bit hit mask.  A node holds, per axis, the
 * children's lo planes then their hi planes (W/4 float4 each); the near
 * planes of an axis are its lo planes when the direction is positive, so
 * each lane loads its near and far planes at byte offsets picked by the
 * signs of its ray.  Empty slots carry inverted boxes (lo = +inf,
 * hi = -inf), which fail the test whatever the direction. */
template <int W>
LXD uint32_t lx_boxw(const void* nodes, uint32_t node, const LxSlab& sl, float tcut, float* tn, uint32_t* cs) {
    constexpr uint32_t G = W / 4;   /* float4 groups per plane set */
    constexpr uint32_t HB = 16 * G; /* bytes from an axis's lo planes to its hi planes */
    const LX_GLOBAL char* nb = (const LX_GLOBAL char*)lx_gptr((const char*)nodes);
    const uint32_t base = node * (uint32_t)sizeof(typename LxNodeT<W>::type);
    const uint32_t nx = base | ((lxm_f2u(sl.ix.x) >> 31) * HB);
    const uint32_t ny = base | ((lxm_f2u(sl.ix.y) >> 31) * HB);
    const uint32_t nz = base | ((lxm_f2u(sl.ix.z) >> 31) * HB);
    float nxs[W], fxs[W], nys[W], fys[W], nzs[W], fzs[W];
#pragma unroll
    for (int g = 0; g < (int)G; g++) {
        lx_v4f a = *(const LX_GLOBAL lx_v4f*)(nb + nx + 16 * g);
        lx_v4f b = *(const LX_GLOBAL lx_v4f*)(nb + (nx ^ HB) + 16 * g);
        lx_v4f c = *(const LX_GLOBAL lx_v4f*)(nb + ny + 2 * HB + 16 * g);
        lx_v4f d = *(const LX_GLOBAL lx_v4f*)(nb + (ny ^ HB) + 2 * HB + 16 * g);
        lx_v4f e = *(const LX_GLOBAL lx_v4f*)(nb + nz + 4 * HB + 16 * g);
        lx_v4f f = *(const LX_GLOBAL lx_v4f*)(nb + (nz ^ HB) + 4 * HB + 16 * g);
        lx_v4i ch = *(const LX_GLOBAL lx_v4i*)(nb + base + 6 * HB + 16 * g);
        nxs[4 * g + 0] = a.x, nxs[4 * g + 1] = a.y, nxs[4 * g + 2] = a.z, nxs[4 * g + 3] = a.w;
        fxs[4 * g + 0] = b.x, fxs[4 * g + 1] = b.y, fxs[4 * g + 2] = b.z, fxs[4 * g + 3] = b.w;
        nys[4 * g + 0] = c.x, nys[4 * g + 1] = c.y, nys[4 * g + 2] = c.z, nys[4 * g + 3] = c.w;
        fys[4 * g + 0] = d.x, fys[4 * g + 1] = d.y, fys[4 * g + 2] = d.z, fys[4 * g + 3] = d.w;
        nzs[4 * g + 0] = e.x, nzs[4 * g + 1] = e.y, nzs[4 * g + 2] = e.z, nzs[4 * g + 3] = e.w;
        fzs[4 * g + 0] = f.x, fzs[4 * g + 1] = f.y, fzs[4 * g + 2] = f.z, fzs[4 * g + 3] = f.w;
        cs[4 * g + 0] = (uint32_t)ch.x, cs[4 * g + 1] = (uint32_t)ch.y;
        cs[4 * g + 2] = (uint32_t)ch.z, cs[4 * g + 3] = (uint32_t)ch.w;
    }
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        float ax = __builtin_fmaf(nxs[k], sl.ix.x, -sl.pn.x), bx = __builtin_fmaf(fxs[k], sl.ix.x, -sl.pf.x);
        float ay = __builtin_fmaf(nys[k], sl.ix.y, -sl.pn.y), by = __builtin_fmaf(fys[k], sl.ix.y, -sl.pf.y);
        float az = __builtin_fmaf(nzs[k], sl.ix.z, -sl.pn.z), bz = __builtin_fmaf(fzs[k], sl.ix.z, -sl.pf.z);
        float tnear = __builtin_fmaxf(__builtin_fmaxf(ax, ay), __builtin_fmaxf(az, 0.0f));
        float tfar = __builtin_fminf(__builtin_fminf(bx, by), __builtin_fminf(bz, tcut));
        bool h = tnear <= tfar * LX_FAR_INFLATE;
        tn[k] = tnear;
        mask |= h ? (1u << k) : 0u;
    }
    /* the child refs are used only when a child is hit; consuming them here
     * keeps their load in flight with the planes' instead of issuing it
     * after the test */
#pragma unroll
    for (int k = 0; k < W; k++) __asm__ volatile("" ::"v"(cs[k]));
    return mask;
}

/* After a node test with hit mask m != 0: returns the nearest hit child (the
 * next node) and pushes the other hit children far to near, so they pop
 * near to far.  BVH4 orders by a 5-exchange sorting network (branch-free);
 * wider nodes by repeated selection.  The order only decides which boxes
 * are culled by a closer hit found first, never which hit is returned. */
template <int W>
LXD uint32_t lx_order_children(const float* tn, const uint32_t* cs, uint32_t m, LxStack& st, uint32_t* overflow) {
    if (W == 4) {
        const float inf = __builtin_inff();
        float t0 = (m & 1u) ? tn[0] : inf, t1 = (m & 2u) ? tn[1] : inf;
        float t2 = (m & 4u) ? tn[2] : inf, t3 = (m & 8u) ? tn[3] : inf;
        uint32_t c0 = cs[0], c1 = cs[1], c2 = cs[2], c3 = cs[3];
#define LX_CX(a, b)                                                                                                    \
    {                                                                                                                  \
        bool sw = t##b < t##a;                                                                                         \
        float tt = sw ? t##b : t##a;                                                                                   \
        t##b = sw ? t##a : t##b;                                                                                       \
        t##a = tt;                                                                                                     \
        uint32_t cc = sw ? c##b : c##a;                                                                                \
        c##b = sw ? c##a : c##b;                                                                                       \
        c##a = cc;                                                                                                     \
    }
        LX_CX(0, 1) LX_CX(2, 3) LX_CX(0, 2) LX_CX(1, 3) LX_CX(1, 2)
#undef LX_CX
        /* hit children have finite entry distances and sort first */
        bool h1 = t1 != inf, h2 = t2 != inf, h3 = t3 != inf;
        int n = (int)h1 + (int)h2 + (int)h3;
        int sp = st.sp;
        if (sp + n > LX_STACK + LX_SPILL) {
            atomicOr(overflow, 1u);
            return c0;
        }
        if (sp + 3 <= LX_STACK) {
            /* the n hit children far to near at sp.., the others above the new top */
            st.lds[(sp + (h1 ? n - 1 : 0)) * st.stride] = c1;
            st.lds[(sp + (h2 ? n - 2 : 1)) * st.stride] = c2;
            st.lds[(sp + (h3 ? 0 : 2)) * st.stride] = c3;
            st.sp = sp + n;
        } else {
            if (h3) st.push(c3);
            if (h2) st.push(c2);
            if (h1) st.push(c1);
        }
        return c0;
    } else {
        float bt = __builtin_inff();
        int bk = -1;
#pragma unroll
        for (int k = 0; k < W; k++)
            if ((m >> k) & 1u)
                if (tn[k] < bt) {
                    bt = tn[k];
                    bk = k;
                }
        m &= ~(1u << bk);
        while (m) {
            float ft = -1.0f;
            int fk = -1;
#pragma unroll
            for (int k = 0; k < W; k++)
                if ((m >> k) & 1u)
                    if (tn[k] > ft) {
                        ft = tn[k];
                        fk = k;
                    }
            m &= ~(1u << fk);
            if (st.sp >= LX_STACK + LX_SPILL) {
                atomicOr(overflow, 1u);
                continue;
            }
            st.push(cs[fk]);
        }
        return cs[bk];
    }
}

/* Closest (ANY = false) or any (ANY = true) accepted hit along the ray. */
/* Per-ray work counters of the counting replay (bytes model). */
struct LxWork {
    uint32_t nodes, tris, insts;
};

/* Next entry of the traversal stack; a sentinel (leaving an instance)
 * restores the world-space ray and is consumed in the same step, so an
 * iteration of the traversal loops never spends itself on a sentinel.
 * Returns false when the stack is empty (the ray is done). */
template <bool EARLY = false>
LXD bool lx_pop_next(LxStack& st, uint32_t& node, bool& in_blas, f3& o, f3& d, LxSlab& sl, f3 wo, f3 wd, f3 iwd) {
    if (st.sp == 0) return false;
    node = st.pop();
    if (node == LX_SENTINEL) {
        in_blas = false;
        /* EARLY: with nothing left above the instance the ray is done, and
         * the world ray need not be restored */
        if (EARLY && st.sp == 0) return false;
        o = wo;
        d = wd;
        sl = lx_slab(wo, iwd);
        if (!EARLY && st.sp == 0) return false;
        node = st.pop();
    }
    return true;
}

/* Enters the group of TLAS leaf `node` (one instance, or instances sharing
 * a transform and a tagged BLAS): ray to object space, sentinel pushed, node
 * = the BLAS root (tested in the same iteration); cur_inst = the instance,
 * or LX_GROUP_TAGGED when each triangle names its own.  Returns false if the
 * stack is full (the group is skipped and flagged). */
template <bool CNT>
LXD bool lx_enter_instance(const LxDevScene& S, uint32_t& node, uint32_t& cur_inst, uint32_t& inst_flags,
                           bool& in_blas, f3& o, f3& d, LxSlab& sl, f3 wo, f3 wd, uint32_t alpha_mask, LxStack& st,
                           uint32_t* overflow, LxWork* work) {
    if (st.sp >= LX_STACK + LX_SPILL) {
        atomicOr(overflow, 1u);
        return false;
    }
    const LX_GLOBAL LxGroup& G = lx_gptr(S.groups)[node & LX_LEAF_FIRST_MASK];
    o = xf_point((const float*)G.inv, wo);
    d = xf_vector((const float*)G.inv, wd);
    sl = lx_slab(o, lx_rcp3(d));
    inst_flags = G.flags & alpha_mask;
    cur_inst = G.inst;
    in_blas = true;
    st.push(LX_SENTINEL);
    if (CNT) work->insts++;
    node = G.blas_root;
    return true;
}

/* One triangle of a leaf: the Moller-Trumbore test, for closest-hit rays
 * the lexicographic order against the best hit so far, and the alpha
 * any-hit program; returns true if the triangle is accepted (and recorded
 * in best). */
LXD bool lx_leaf_tri(const LxDevScene& S, bool any, const float4& A, const float4& E1, const float4& E2,
                     uint32_t cur_inst, uint32_t inst_flags, f3 o, f3 d, f3 wd, float tmax, uint32_t key,
                     uint32_t iteration, LxHitD& best) {
    float t, u, v;
    if (!lx_mt(o, d, A, E1, E2, tmax, t, u, v)) return false;
    uint32_t prim = lxm_f2u(A.w);
    uint32_t inst = cur_inst == LX_GROUP_TAGGED ? lxm_f2u(E1.w) : cur_inst;
    if (!any && !lx_better(t, inst, prim, best)) return false;
    if (inst_flags & LX_INST_ALPHA) {
        const LX_GLOBAL LxInstance& I = lx_gptr(S.inst)[inst];
        if ((I.flags & LX_INST_ALPHA) && !lx_any_hit_accept(S, I, inst, prim, u, v, wd, key, iteration)) return false;
    }
    best.t = t;
    best.inst = inst;
    best.prim = prim;
    best.u = u;
    best.v = v;
    return true;
}

/* Triangle leaf of the current group, for an any-hit (any) or closest-hit
 * ray (one copy of the code serves both, so a wave whose lanes trace both
 * kinds runs it once); returns true if an any-hit ray found its hit. */
template <bool CNT>
LXD bool lx_leaf(const LxDevScene& S, bool any, uint32_t node, uint32_t cur_inst, uint32_t inst_flags, f3 o, f3 d,
                 f3 wd, float tmax, uint32_t key, uint32_t iteration, LxHitD& best, LxWork* work) {
    const uint32_t first = node & LX_LEAF_FIRST_MASK;
    const uint32_t cnt = ((node >> 28) & 7u) + 1u;
    const LX_GLOBAL lx_v4f* tb = (const LX_GLOBAL lx_v4f*)lx_gptr(S.tris) + 3u * (size_t)first;
    for (uint32_t k = 0; k < cnt; k++) {
        const LX_GLOBAL lx_v4f* p0 = tb + 3u * k;
        float4 A0 = lx_ld4(p0), B0 = lx_ld4(p0 + 1), C0 = lx_ld4(p0 + 2);
        if (CNT) work->tris++;
        if (lx_leaf_tri(S, any, A0, B0, C0, cur_inst, inst_flags, o, d, wd, tmax, key, iteration, best) && any)
            return true;
    }
    return false;
}

/* Closest (ANY = false) or any (ANY = true) accepted hit along the ray.
 * One loop iteration = (instance entry, then) one triangle leaf or one
 * node test. */
template <bool ANY, bool CNT = false, int W = 4>
LXD LxHitD lx_trace(const LxDevScene& S, f3 wo, f3 wd, float tmax, uint32_t key, uint32_t iteration,
                    uint32_t alpha_mask, LxStack& st, uint32_t* overflow, LxWork* work = nullptr) {
    LxHitD best;
    best.t = __builtin_inff();
    best.inst = 0xffffffffu;
    best.prim = 0xffffffffu;
    best.u = best.v = 0.0f;
    f3 iwd = lx_rcp3(wd);
    f3 o = wo, d = wd;
    LxSlab sl = lx_slab(wo, iwd);
    bool in_blas = false;
    uint32_t cur_inst = 0;
    uint32_t inst_flags = 0;
    st.sp = 0;
    uint32_t node = S.tlas_root;
    while (true) {
        /* one step: (an instance entry, then) one triangle leaf or one node
         * test; every step that ends its subtree pops in one place */
        bool pop = false;
        if ((node & LX_LEAF_BIT) && !in_blas)
            pop = !lx_enter_instance<CNT>(S, node, cur_inst, inst_flags, in_blas, o, d, sl, wo, wd, alpha_mask, st,
                                          overflow, work);
        if (!pop) {
            if (node & LX_LEAF_BIT) {
                /* a TLAS leaf here (popped after a skipped instance) waits for the next step */
                if (in_blas) {
                    if (lx_leaf<CNT>(S, ANY, node, cur_inst, inst_flags, o, d, wd, tmax, key, iteration, best, work))
                        return best;
                    pop = true;
                }
            } else {
                float tcut = (best.inst == 0xffffffffu ? tmax : best.t) * LX_CULL_MARGIN;
                float tn[W];
                uint32_t cs[W];
                uint32_t m = lx_boxw<W>(S.nodes, node, sl, tcut, tn, cs);
                if (CNT) work->nodes++;
                if (m == 0u) pop = true;
                else node = lx_order_children<W>(tn, cs, m, st, overflow);
            }
        }
        if (pop && !lx_pop_next(st, node, in_blas, o, d, sl, wo, wd, iwd)) break;
    }
    return best;
}

/* Direct-light shadow rays.  sampleDirectLight's ray is a closest-hit ray,
 * but shadePath uses its hit only through the hit material's emission
 * (shade_path.cu:246-289), which is exactly zero on a material whose
 * emission colour x strength is zero in every channel: such a closest hit
 * adds +0 to a colour that is never -0, i.e. nothing, as long as the term's
 * other factors are finite (where one overflows, the closest-hit form's
 * 0 x inf would make the pixel NaN; this form skips the term).  So the ray is
 * answered by two cheaper queries with the same result for the colour:
 *   DARK  (lx_trace_q<1>): any accepted candidate on a non-emitting instance
 *         that comes before `bound` in the lexicographic (t, instance,
 *         primitive) order -- an occlusion query that stops at the first one;
 *   LIGHT (lx_trace_q<2>): the closest accepted candidate on an emitting
 *         instance before `bound`, over the emitter TLAS.
 * Candidates are the same Moller-Trumbore and alpha any-hit acceptances as
 * lx_trace's.  Returns whether a candidate was found (in best). */
template <int MODE, bool CNT, int W>
LXD bool lx_trace_q(const LxDevScene& S, uint32_t root, f3 wo, f3 wd, float tmax, const LxHitD& bound, uint32_t key,
                    uint32_t iteration, uint32_t alpha_mask, LxStack& st, uint32_t* overflow, LxWork* work,
                    LxHitD& best) {
    best = bound;
    f3 iwd = lx_rcp3(wd);
    f3 o = wo, d = wd;
    LxSlab sl = lx_slab(wo, iwd);
    bool in_blas = false;
    uint32_t cur_inst = 0, inst_flags = 0, emis = 0;
    st.sp = 0;
    uint32_t node = root;
    bool found = false;
    while (true) {
        bool pop = false;
        if ((node & LX_LEAF_BIT) && !in_blas) {
            const uint32_t gf = lx_gptr(S.groups)[node & LX_LEAF_FIRST_MASK].flags;
            /* an occlusion query skips groups that only emit */
            if (MODE == 1 && (gf & LX_GROUP_EMIS_ALL)) pop = true;
            else {
                emis = gf & (LX_GROUP_EMIS_ANY | LX_GROUP_EMIS_ALL);
                pop = !lx_enter_instance<CNT>(S, node, cur_inst, inst_flags, in_blas, o, d, sl, wo, wd, alpha_mask,
                                              st, overflow, work);
            }
        }
        if (!pop) {
            if (node & LX_LEAF_BIT) {
                if (in_blas) {
                    const uint32_t first = node & LX_LEAF_FIRST_MASK;
                    const uint32_t cnt = ((node >> 28) & 7u) + 1u;
                    const LX_GLOBAL lx_v4f* tb = (const LX_GLOBAL lx_v4f*)lx_gptr(S.tris) + 3u * (size_t)first;
                    for (uint32_t k = 0; k < cnt; k++) {
                        const LX_GLOBAL lx_v4f* p0 = tb + 3u * k;
                        float4 A = lx_ld4(p0), B = lx_ld4(p0 + 1), Cc = lx_ld4(p0 + 2);
                        if (CNT) work->tris++;
                        float t, u, v;
                        if (!lx_mt(o, d, A, B, Cc, tmax, t, u, v)) continue;
                        const uint32_t prim = lxm_f2u(A.w);
                        const uint32_t inst = cur_inst == LX_GROUP_TAGGED ? lxm_f2u(B.w) : cur_inst;
                        if (!lx_better(t, inst, prim, best)) continue;
                        bool em;
                        if (emis == (LX_GROUP_EMIS_ANY | LX_GROUP_EMIS_ALL)) em = true;
                        else if (emis == 0u) em = false;
                        else em = (lx_gptr(S.inst)[inst].flags & LX_INST_EMISSIVE) != 0u;
                        if (MODE == 1 ? em : !em) continue;
                        if (inst_flags & LX_INST_ALPHA) {
                            const LX_GLOBAL LxInstance& I = lx_gptr(S.inst)[inst];
                            if ((I.flags & LX_INST_ALPHA) && !lx_any_hit_accept(S, I, inst, prim, u, v, wd, key, iteration))
                                continue;
                        }
                        best.t = t;
                        best.inst = inst;
                        best.prim = prim;
                        best.u = u;
                        best.v = v;
                        found = true;
                        if (MODE == 1) return true;
                    }
                    pop = true;
                }
            } else {
                float tcut = (best.inst == 0xffffffffu ? tmax : best.t) * LX_CULL_MARGIN;
                float tn[W];
                uint32_t cs[W];
                uint32_t m = lx_boxw<W>(S.nodes, node, sl, tcut, tn, cs);
                if (CNT) work->nodes++;
                if (m == 0u) pop = true;
                else node = lx_order_children<W>(tn, cs, m, st, overflow);
            }
        }
        if (pop && !lx_pop_next(st, node, in_blas, o, d, sl, wo, wd, iwd)) break;
    }
    return found;
}

/* The direct-light shadow ray (so, ld, tmax) as far as shading can tell:
 * its closest hit if that emits, otherwise a miss. */
template <bool CNT, int W>
LXD LxHitD lx_trace_dl(const LxDevScene& S, f3 so, f3 ld, float tmax, uint32_t key, uint32_t iteration,
                       LxStack& st, uint32_t* overflow, LxWork* work) {
    LxHitD none;
    none.t = __builtin_inff();
    none.inst = 0xffffffffu;
    none.prim = 0xffffffffu;
    none.u = none.v = 0.0f;
    LxHitD dark, light;
    const bool occ = lx_trace_q<1, CNT, W>(S, S.tlas_root, so, ld, tmax, none, key, iteration, LX_INST_ALPHA, st,
                                           overflow, work, dark);
    if (S.emis_tlas_root == LX_CHILD_EMPTY ||
        !lx_trace_q<2, CNT, W>(S, S.emis_tlas_root, so, ld, tmax, occ ? dark : none, key, iteration, LX_INST_ALPHA,
                               st, overflow, work, light))
        return none; /* no emitting candidate before the first occluder found */
    /* an emitting candidate before it: the closest hit is that one unless a
     * non-emitting candidate comes first */
    if (occ && lx_trace_q<1, CNT, W>(S, S.tlas_root, so, ld, tmax, light, key, iteration, LX_INST_ALPHA, st, overflow,
                                     work, dark))
        return none;
    return light;
}

/* Two rays of one lane in one traversal loop: first (when has0) an any-hit
 * ray (o0, d0, tmax0), then (when has1) a closest-hit ray (o1, d1, tmax1);
 * at least one of them must be present.  A lane that
 * finishes its first ray starts the second at once instead of idling until
 * every lane of its wave has finished its first ray.  Each ray's result is
 * what lx_trace returns for it alone (the candidate set does not depend on
 * traversal order). */
template <bool CNT, int W>
LXD LxHitD lx_trace_pair(const LxDevScene& S, bool has0, f3 o0, f3 d0, float tmax0, bool has1, f3 o1, f3 d1,
                         float tmax1, uint32_t key, uint32_t iteration, uint32_t alpha_mask, LxStack& st,
                         uint32_t* overflow, bool& occluded0, LxWork* wclosest, LxWork* wany) {
    LxHitD best;
    bool any = has0;
    f3 wo = any ? o0 : o1, wd = any ? d0 : d1;
    float tmax = any ? tmax0 : tmax1;
    f3 iwd, o, d;
    LxSlab sl;
    bool in_blas;
    uint32_t cur_inst = 0, inst_flags = 0, node;
    occluded0 = false;
    /* (re)start the lane on its current ray */
    auto begin = [&]() {
        best.t = __builtin_inff();
        best.inst = 0xffffffffu;
        best.prim = 0xffffffffu;
        best.u = best.v = 0.0f;
        iwd = lx_rcp3(wd);
        o = wo;
        d = wd;
        sl = lx_slab(wo, iwd);
        in_blas = false;
        st.sp = 0;
        node = S.tlas_root;
    };
    begin();
    while (true) {
        LxWork* work = any ? wany : wclosest;
        bool pop = false, found_any = false;
        if ((node & LX_LEAF_BIT) && !in_blas)
            pop = !lx_enter_instance<CNT>(S, node, cur_inst, inst_flags, in_blas, o, d, sl, wo, wd, alpha_mask, st,
                                          overflow, work);
        if (!pop) {
            if (node & LX_LEAF_BIT) {
                /* a TLAS leaf here (popped after a skipped instance) waits for the next step */
                if (in_blas) {
                    found_any = lx_leaf<CNT>(S, any, node, cur_inst, inst_flags, o, d, wd, tmax, key, iteration, best,
                                             work);
                    pop = !found_any;
                }
            } else {
                float tcut = (best.inst == 0xffffffffu ? tmax : best.t) * LX_CULL_MARGIN;
                float tn[W];
                uint32_t cs[W];
                uint32_t m = lx_boxw<W>(S.nodes, node, sl, tcut, tn, cs);
                if (CNT) work->nodes++;
                if (m == 0u) pop = true;
                else node = lx_order_children<W>(tn, cs, m, st, overflow);
            }
        }
        bool more = true;
        if (pop) more = lx_pop_next<true>(st, node, in_blas, o, d, sl, wo, wd, iwd);
        if (found_any || !more) {
            if (!any) break;
            /* first ray finished: record it and start the second */
            occluded0 = best.inst != 0xffffffffu;
            if (!has1) break;
            any = false;
            wo = o1;
            wd = d1;
            tmax = tmax1;
            begin();
        }
    }
    return best;
}

/* __closesthit__ch, optix_triangle.cu:156-243: interpolated normal (not
 * flipped toward the ray), uv, uv-derivative tangent, object -> world. */
LXD void lx_hit_attrs(const LxDevScene& S, const LxHitD& h, bool want_tangent, f3& N, f3& Tg, f2& uv, int& mat) {
    const LX_GLOBAL LxInstance& I = lx_gptr(S.inst)[h.inst];
    const LX_GLOBAL uint32_t* ix = lx_gptr(S.idx) + I.idx_off + 3u * h.prim;
    uint32_t i0 = ix[0], i1 = ix[1], i2 = ix[2];
    const LX_GLOBAL float* P = lx_gptr(S.pos) + 3u * (size_t)I.vtx_off;
    f3 A = mk3(P[3 * i0], P[3 * i0 + 1], P[3 * i0 + 2]);
    f3 B = mk3(P[3 * i1], P[3 * i1 + 1], P[3 * i1 + 2]);
    f3 C = mk3(P[3 * i2], P[3 * i2 + 1], P[3 * i2 + 2]);
    f3 e1 = B - A, e2 = C - A;
    f3 normal;
    float bA = 1.0f - h.u - h.v;
    if (I.has_n_uv & 1u) {
        const LX_GLOBAL float* Nn = lx_gptr(S.nrm) + 3u * (size_t)I.vtx_off;
        f3 nA = mk3(Nn[3 * i0], Nn[3 * i0 + 1], Nn[3 * i0 + 2]);
        f3 nB = mk3(Nn[3 * i1], Nn[3 * i1 + 1], Nn[3 * i1 + 2]);
        f3 nC = mk3(Nn[3 * i2], Nn[3 * i2 + 1], Nn[3 * i2 + 2]);
        normal = normalize(((bA * nA) + (h.u * nB)) + (h.v * nC));
    } else {
        normal = normalize(cross(e1, e2));
    }
    f3 tangent = mk3(1.0f, 0.0f, 0.0f);
    uv = mk2(0.0f, 0.0f);
    if (I.has_n_uv & 2u) {
        const LX_GLOBAL float* U = lx_gptr(S.uv) + 2u * (size_t)I.vtx_off;
        f2 a = mk2(U[2 * i0], U[2 * i0 + 1]), b = mk2(U[2 * i1], U[2 * i1 + 1]), c = mk2(U[2 * i2], U[2 * i2 + 1]);
        uv = mk2((bA * a.x + h.u * b.x) + h.v * c.x, (bA * a.y + h.u * b.y) + h.v * c.y);
        if (want_tangent) {
            f2 d1 = mk2(b.x - a.x, b.y - a.y), d2 = mk2(c.x - a.x, c.y - a.y);
            float det = (d1.x * d2.y - d2.x * d1.y);
            float f = (lxm_abs(det) < 1e-10f) ? 0.0f : 1.0f / det;
            if (f != 0.0f)
                tangent = mk3(f * (d2.y * e1.x - d1.y * e2.x), f * (d2.y * e1.y - d1.y * e2.y),
                              f * (d2.y * e1.z - d1.y * e2.z));
        }
    }
    normal = normalize(xf_normal((const float*)I.inv, normal));
    if (want_tangent) {
        tangent = normalize(xf_vector((const float*)I.m, tangent));
        f3 T = tangent - dot(tangent, normal) * normal;
        float len2 = dot(T, T);
        if (len2 > 1e-10f) {
            tangent = T * (1.0f / lxm_sqrt(len2));
        } else {
            tangent = (lxm_abs(normal.x) > 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
            tangent = normalize(cross(tangent, normal));
        }
    }
    N = normal;
    Tg = tangent;
    mat = (int)I.material;
}

/* __closesthit__ch_direct_light, optix_triangle.cu:259-305 */
LXD void lx_dl_attrs(const LxDevScene& S, const LxHitD& h, f3& N, f2& uv, int& mat) {
    const LX_GLOBAL LxInstance& I = lx_gptr(S.inst)[h.inst];
    const LX_GLOBAL uint32_t* ix = lx_gptr(S.idx) + I.idx_off + 3u * h.prim;
    uint32_t i0 = ix[0], i1 = ix[1], i2 = ix[2];
    float bA = 1.0f - h.u - h.v;
    f3 normal;
    if (I.has_n_uv & 1u) {
        const LX_GLOBAL float* Nn = lx_gptr(S.nrm) + 3u * (size_t)I.vtx_off;
        f3 nA = mk3(Nn[3 * i0], Nn[3 * i0 + 1], Nn[3 * i0 + 2]);
        f3 nB = mk3(Nn[3 * i1], Nn[3 * i1 + 1], Nn[3 * i1 + 2]);
        f3 nC = mk3(Nn[3 * i2], Nn[3 * i2 + 1], Nn[3 * i2 + 2]);
        normal = normalize(((bA * nA) + (h.u * nB)) + (h.v * nC));
    } else {
        const LX_GLOBAL float* P = lx_gptr(S.pos) + 3u * (size_t)I.vtx_off;
        f3 A = mk3(P[3 * i0], P[3 * i0 + 1], P[3 * i0 + 2]);
        f3 B = mk3(P[3 * i1], P[3 * i1 + 1], P[3 * i1 + 2]);
        f3 C = mk3(P[3 * i2], P[3 * i2 + 1], P[3 * i2 + 2]);
        normal = normalize(cross(B - A, C - A));
    }
    uv = mk2(0.0f, 0.0f);
    if (I.has_n_uv & 2u) {
        const LX_GLOBAL float* U = lx_gptr(S.uv) + 2u * (size_t)I.vtx_off;
        f2 a = mk2(U[2 * i0], U[2 * i0 + 1]), b = mk2(U[2 * i1], U[2 * i1 + 1]), c = mk2(U[2 * i2], U[2 * i2 + 1]);
        uv = mk2((bA * a.x + h.u * b.x) + h.v * c.x, (bA * a.y + h.u * b.y) + h.v * c.y);
    }
    N = normalize(xf_normal((const float*)I.inv, normal));
    mat = (int)I.material;
}

#endif
