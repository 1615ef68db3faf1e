/*
 * lxo_oracle.cpp -- CPU oracle: a from-scratch restatement of the reference
 * renderer's hot path, used ONLY by the test suite, the smoke check and the
 * benchmark's CPU-baseline leg (never by the product library).
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Floating-point expressions keep the
 * reference's operand order and association (glm / sutil semantics), so that
 * with contraction off the result is reproducible bit for bit.
 *
 * Documented deviations from the reference (each also used by the GPU path):
 *  - transcendentals come from include/luxion_detmath.h instead of the CUDA
 *    device libm (closed, not reproducible);
 *  - closest-hit ties are broken by the lexicographic minimum of
 *    (t, instance, primitive) instead of traversal order;
 *  - leaves are never culled by their (uninitialised) entry distance
 *    (mesh_isect.cu:56-65 reads an uninitialised tmin for leaves), and
 *    inner nodes are pruned only when their entry distance exceeds the best
 *    t by a 2^-16 relative margin, so pruning never changes the result;
 *  - instance inverse transforms are computed in double and rounded once;
 *  - texture filtering is the exact-weight bilinear wrap sampler below
 *    (tex2D's 8-bit fixed-point weights are hardware-specific);
 *  - the blend any-hit hash is keyed on the pixel index, not the compacted
 *    queue slot (optix_triangle.cu:355), which is nondeterministic.
 *  - the triangle test is the software Moller-Trumbore test (OptiX's
 *    watertight hardware test is closed).
 *
 * Parity unpinned (restated, with no closed form, known answer or brute
 * force to check it against; the GPU renders these bit for bit as the
 * oracle does): the diffuse / specular lobe split of sampleCookTorrance when
 * the transmission lobe is off (only the mixture's E[1/pdf] is pinned), the
 * reflection sampler's zero-direction rejections, shadePath's origin offsets
 * and epsilons and its 3.14159 environment lookups, the raygen's shadow-ray
 * origin and tmax, and the legacy glass shader's lobe choice beyond its
 * binomial check.  Every other function is pinned by tests/test_oracle_cpu.py,
 * tests/test_oracle_pins_cpu.py, tests/test_fp64_oracle_cpu.py,
 * tests/test_legacy_debug_cpu.py or tests/test_gltf_cpu.py.
 */
#include "lxo_oracle.h"
#include "../include/luxion_detmath.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>
#include <vector>
#include <limits>
#include <type_traits>

/* Arithmetic type of the restatement: float (the oracle proper, bit-exact
 * with the reference's fp32 sequences) or, built with -DLXO_FP64, double: the
 * same algorithm, random streams and constants in double precision, with
 * libm's double transcendentals, as the FP64 reference for converged-image
 * error.  The random draws stay float (the reference's u01). */
#ifdef LXO_FP64
typedef double lxr;
#define LXO_D2(name, expr)                                                                                             \
    template <class A, class B>                                                                                        \
    static inline typename std::enable_if<std::is_same<A, double>::value || std::is_same<B, double>::value,            \
                                          double>::type                                                                \
    name(A a_, B b_) {                                                                                                 \
        double x = (double)a_, y = (double)b_;                                                                         \
        return expr;                                                                                                   \
    }
LXO_D2(lxm_min, x < y ? x : y)
LXO_D2(lxm_max, x > y ? x : y)
LXO_D2(lxm_atan2, std::atan2(x, y))
LXO_D2(lxm_pow, std::pow(x, y))
#undef LXO_D2
template <class A, class B, class C>
static inline typename std::enable_if<std::is_same<A, double>::value || std::is_same<B, double>::value ||
                                          std::is_same<C, double>::value,
                                      double>::type
lxm_clamp(A x, B lo, C hi) {
    return lxm_min(lxm_max((double)x, (double)lo), (double)hi);
}
static inline double lxm_sqrt(double x) { return std::sqrt(x); }
static inline double lxm_abs(double x) { return std::fabs(x); }
static inline double lxm_floor(double x) { return std::floor(x); }
static inline int lxm_isnan(double x) { return x != x; }
static inline int lxm_isfinite(double x) { return std::isfinite(x) ? 1 : 0; }
static inline double lxm_sin(double x) { return std::sin(x); }
static inline double lxm_cos(double x) { return std::cos(x); }
static inline double lxm_atan(double x) { return std::atan(x); }
static inline double lxm_acos(double x) { return std::acos(x); }
static inline double lxm_log2(double x) { return std::log2(x); }
static inline double lxm_exp2(double x) { return std::exp2(x); }
#else
typedef float lxr;
#endif

namespace {

/* ------------------------------------------------------------ vec3 (glm) */
struct V3 {
    lxr x, y, z;
};
struct V2 {
    lxr x, y;
};
static inline V3 v3(lxr x, lxr y, lxr z) { return V3{x, y, z}; }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 muls(V3 a, lxr s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 smul(lxr s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V3 divs(V3 a, lxr s) { return v3(a.x / s, a.y / s, a.z / s); } /* glm vec / scalar */
static inline V3 divv(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
/* glm compute_dot<tvec3>: tmp = x * y; tmp.x + tmp.y + tmp.z */
static inline lxr dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 x, V3 y) {
    return v3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
/* glm normalize: x * inversesqrt(dot(x, x)), inversesqrt = 1 / sqrt */
static inline V3 normalize(V3 a) { return muls(a, 1.0f / lxm_sqrt(dot(a, a))); }
static inline lxr length(V3 a) { return lxm_sqrt(dot(a, a)); }
static inline V3 vmax0(V3 a) { return v3(lxm_max(a.x, 0.0f), lxm_max(a.y, 0.0f), lxm_max(a.z, 0.0f)); }

/* 3x4 row-major affine map; point = (m0 x + m1 y) + (m2 z + m3) (glm mat4 *
 * vec4 pairwise order, glm/detail/type_mat4x4.inl operator*). */
static inline V3 xform_point(const lxr* m, V3 p) {
    return v3((m[0] * p.x + m[1] * p.y) + (m[2] * p.z + m[3]), (m[4] * p.x + m[5] * p.y) + (m[6] * p.z + m[7]),
              (m[8] * p.x + m[9] * p.y) + (m[10] * p.z + m[11]));
}
static inline V3 xform_vector(const lxr* m, V3 v) {
    return v3((m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[4] * v.x + m[5] * v.y) + m[6] * v.z,
              (m[8] * v.x + m[9] * v.y) + m[10] * v.z);
}
/* normal object->world by the transpose of the world->object map */
static inline V3 xform_normal(const lxr* inv, V3 n) {
    return v3((inv[0] * n.x + inv[4] * n.y) + inv[8] * n.z, (inv[1] * n.x + inv[5] * n.y) + inv[9] * n.z,
              (inv[2] * n.x + inv[6] * n.y) + inv[10] * n.z);
}

/* Affine inverse in double, rounded once to the arithmetic type. */
static void affine_inverse(const lxr* m, lxr* out) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double inv_det = 1.0 / det;
    double r[9] = {A * inv_det,
                   -(b * i - c * h) * inv_det,
                   (b * f - c * e) * inv_det,
                   B * inv_det,
                   (a * i - c * g) * inv_det,
                   -(a * f - c * d) * inv_det,
                   C * inv_det,
                   -(a * h - b * g) * inv_det,
                   (a * e - b * d) * inv_det};
    double tx = m[3], ty = m[7], tz = m[11];
    for (int row = 0; row < 3; row++) {
        out[row * 4 + 0] = (lxr)r[row * 3 + 0];
        out[row * 4 + 1] = (lxr)r[row * 3 + 1];
        out[row * 4 + 2] = (lxr)r[row * 3 + 2];
        out[row * 4 + 3] = (lxr)(-(r[row * 3 + 0] * tx + r[row * 3 + 1] * ty + r[row * 3 + 2] * tz));
    }
}

/* ------------------------------------------------------------------- RNG */
/* utilhash, include/intersections.h:13-22 */
static inline uint32_t utilhash(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

/* thrust::default_random_engine = minstd_rand (a = 48271, m = 2^31 - 1),
 * seeded per include/common.h:38-43; seed rule: state = s mod m, 0 -> 1. */
struct Rng {
    uint32_t x;
    float u01() {
        x = (uint32_t)(((uint64_t)x * 48271u) % 2147483647u);
        /* uniform_real_distribution<float>(0, 1): float(x - min) /
         * (1 + float(max - min)), min = 1, max = 2^31 - 2 */
        float r = (float)(x - 1u);
        r = r / (1.0f + (float)(2147483646u - 1u));
        return r * (1.0f - 0.0f) + 0.0f;
    }
};

static inline Rng make_rng(int iter, int index, int depth, uint32_t seed) {
    uint32_t h = utilhash((uint32_t)(1u << 31) | ((uint32_t)depth << 22) | (uint32_t)iter) ^ utilhash((uint32_t)index);
    if (seed != 0u) h ^= utilhash(seed * 0x9E3779B9u + 0x7F4A7C15u);
    Rng r;
    r.x = h % 2147483647u;
    if (r.x == 0u) r.x = 1u;
    return r;
}

/* ------------------------------------------------------------- constants */
static const lxr PI_F = 3.1415926535897932384626422832795028841971f; /* utilities.h:13 */
static const lxr TWO_PI_F = 6.2831853071795864769252867665590057683943f;
static const lxr SQRT_OF_ONE_THIRD = 0.5773502691896257645091487805019574556476f;
static const lxr INV_PI_CT = 0.3183098f;                             /* cook_torrance.h:14 */
static const lxr GLM_PI_F = (lxr)3.14159265358979323846264338327950288; /* glm::pi<float>() */
static const lxr CT_EPS = 1e-30f;                                     /* cook_torrance.cu:3 */
static const lxr PDF_CLAMP = 1e-5f;                                   /* cook_torrance.cu:4 */
static const double SHADE_M_PI = 3.14159;                               /* shade_path.cu:23 */

/* -------------------------------------------------------------- textures */
struct Tex {
    int w = 0, h = 0;
    std::vector<lxr> rgba; /* decoded texels */
};

/* Exact-weight bilinear filter, wrap addressing, normalised coordinates: the
 * documented semantics of tex2D<float4> with cudaFilterModeLinear /
 * cudaAddressModeWrap (texture.cu:52-56), without the hardware's 8-bit weight
 * quantisation. */
static inline void tex_sample(const Tex& t, lxr u, lxr v, lxr* out) {
    if (!lxm_isfinite(u)) u = 0.0f;
    if (!lxm_isfinite(v)) v = 0.0f;
    u = u - lxm_floor(u);
    v = v - lxm_floor(v);
    lxr x = u * (lxr)t.w - 0.5f;
    lxr y = v * (lxr)t.h - 0.5f;
    lxr fx = lxm_floor(x), fy = lxm_floor(y);
    lxr ax = x - fx, ay = y - fy;
    int i0 = (int)fx, j0 = (int)fy;
    if (i0 < 0) i0 += t.w;
    if (i0 >= t.w) i0 -= t.w;
    if (j0 < 0) j0 += t.h;
    if (j0 >= t.h) j0 -= t.h;
    int i1 = i0 + 1 == t.w ? 0 : i0 + 1;
    int j1 = j0 + 1 == t.h ? 0 : j0 + 1;
    const lxr* t00 = &t.rgba[((size_t)j0 * t.w + i0) * 4];
    const lxr* t10 = &t.rgba[((size_t)j0 * t.w + i1) * 4];
    const lxr* t01 = &t.rgba[((size_t)j1 * t.w + i0) * 4];
    const lxr* t11 = &t.rgba[((size_t)j1 * t.w + i1) * 4];
    /* nested lerps: exact for constant textures and at texel centres */
    for (int c = 0; c < 4; c++) {
        lxr a = t00[c] + ax * (t10[c] - t00[c]);
        lxr b = t01[c] + ax * (t11[c] - t01[c]);
        out[c] = a + ay * (b - a);
    }
}

static void decode_texture(const lx_texture_desc& d, Tex& t) {
    t.w = (int)d.width;
    t.h = (int)d.height;
    size_t n = (size_t)d.width * d.height * 4;
    t.rgba.resize(n);
    if (d.format == LX_TEX_RGBA8) {
        const uint8_t* s = (const uint8_t*)d.data;
        for (size_t i = 0; i < n; i++) t.rgba[i] = ((lxr)s[i]) / 255.0f;
    } else if (d.format == LX_TEX_RGBA16) {
        const uint16_t* s = (const uint16_t*)d.data;
        for (size_t i = 0; i < n; i++) t.rgba[i] = ((lxr)s[i]) / 65535.0f;
    } else {
        const float* s = (const float*)d.data;
        for (size_t i = 0; i < n; i++) t.rgba[i] = s[i];
    }
}

/* ------------------------------------------------------------ BVH (ref) */
/* BoundingBox, include/mesh/bounding_box.h:10-62 (EPSILON = 1e-5f) */
struct BBox {
    V3 mn, mx;
    BBox() : mn(v3(FLT_MAX, FLT_MAX, FLT_MAX)), mx(v3(-FLT_MAX, -FLT_MAX, -FLT_MAX)) {}
    static BBox of_points(const V3* p, int n) {
        V3 a = p[0], b = p[0];
        for (int i = 1; i < n; i++) {
            a.x = lxm_min(a.x, p[i].x);
            a.y = lxm_min(a.y, p[i].y);
            a.z = lxm_min(a.z, p[i].z);
            b.x = lxm_max(b.x, p[i].x);
            b.y = lxm_max(b.y, p[i].y);
            b.z = lxm_max(b.z, p[i].z);
        }
        BBox r;
        r.mn = sub(a, v3(1e-5f, 1e-5f, 1e-5f));
        r.mx = add(b, v3(1e-5f, 1e-5f, 1e-5f));
        return r;
    }
    /* expandBox, bounding_box.h:37-45 -- y and z read the other box's .x */
    void expand_box_ref(const BBox& o) {
        mn.x = lxm_min(mn.x, o.mn.x);
        mn.y = lxm_min(mn.y, o.mn.x);
        mn.z = lxm_min(mn.z, o.mn.x);
        mx.x = lxm_max(mx.x, o.mx.x);
        mx.y = lxm_max(mx.y, o.mx.x);
        mx.z = lxm_max(mx.z, o.mx.x);
    }
    void expand_point(V3 p) {
        mn.x = lxm_min(mn.x, p.x);
        mn.y = lxm_min(mn.y, p.y);
        mn.z = lxm_min(mn.z, p.z);
        mx.x = lxm_max(mx.x, p.x);
        mx.y = lxm_max(mx.y, p.y);
        mx.z = lxm_max(mx.z, p.z);
    }
    lxr surface_area() const {
        lxr dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    bool leaf;
    BBox box;
    int tri;
    int left;
};

struct Mesh {
    std::vector<V3> pos, nrm;
    std::vector<V2> uv;
    std::vector<uint32_t> idx;
    bool has_n = false, has_uv = false;
    std::vector<BNode> bvh;
    int bvh_depth = 0;
};

/* Triangle::centroid, sceneStructs.h:48-55 */
static inline V3 centroid(const Mesh& m, int t) {
    V3 c = v3(0.0f, 0.0f, 0.0f);
    c = add(c, m.pos[m.idx[3 * t + 0]]);
    c = add(c, m.pos[m.idx[3 * t + 1]]);
    c = add(c, m.pos[m.idx[3 * t + 2]]);
    return divs(c, 3.0f);
}

/* fill_bvh, src/mesh/bvh_build.cu:15-318: binary binned SAH, 16 bins x 3
 * axes, one triangle per leaf, 15 % median fallback.  Uses the same
 * std::partition / std::nth_element calls so libstdc++ yields the same
 * topology as a g++ build of the reference. */
static BBox fill_bvh(int idx, int start, int end, std::vector<BNode>& h, const Mesh& m, std::vector<int>& ti,
                     int& alloc, int depth, int& max_depth) {
    const int BIN_COUNT = 16;
    if (depth > max_depth) max_depth = depth;
    if (start == end) {
        BNode n;
        n.leaf = true;
        n.tri = ti[start];
        n.left = -1;
        V3 p[3] = {m.pos[m.idx[3 * ti[start] + 0]], m.pos[m.idx[3 * ti[start] + 1]], m.pos[m.idx[3 * ti[start] + 2]]};
        n.box = BBox::of_points(p, 3);
        h[idx] = n;
        return n.box;
    }
    int cnt = end - start + 1;
    std::vector<V3> cen(cnt);
    for (int i = start; i <= end; i++) cen[i - start] = centroid(m, ti[i]);
    lxr mnx = FLT_MAX, mxx = -FLT_MAX, mny = FLT_MAX, mxy = -FLT_MAX, mnz = FLT_MAX, mxz = -FLT_MAX;
    for (auto& c : cen) {
        mnx = lxm_min(mnx, c.x);
        mxx = lxm_max(mxx, c.x);
        mny = lxm_min(mny, c.y);
        mxy = lxm_max(mxy, c.y);
        mnz = lxm_min(mnz, c.z);
        mxz = lxm_max(mxz, c.z);
    }
    lxr dx = mxx - mnx, dy = mxy - mny, dz = mxz - mnz;
    lxr wx = (dx > 0) ? dx / BIN_COUNT : 1.0f;
    lxr wy = (dy > 0) ? dy / BIN_COUNT : 1.0f;
    lxr wz = (dz > 0) ? dz / BIN_COUNT : 1.0f;
    auto bin_of = [&](lxr c, lxr mn, lxr w) {
        int b = (int)((c - mn) / w);
        return b < 0 ? 0 : (b > BIN_COUNT - 1 ? BIN_COUNT - 1 : b);
    };
    int cx[16] = {0}, cy[16] = {0}, cz[16] = {0};
    BBox bx[16], by[16], bz[16];
    for (int i = 0; i < cnt; i++) {
        int ix = bin_of(cen[i].x, mnx, wx), iy = bin_of(cen[i].y, mny, wy), iz = bin_of(cen[i].z, mnz, wz);
        cx[ix]++;
        cy[iy]++;
        cz[iz]++;
        int t = ti[i + start];
        for (int k = 0; k < 3; k++) {
            V3 p = m.pos[m.idx[3 * t + k]];
            bx[ix].expand_point(p);
        }
        for (int k = 0; k < 3; k++) {
            V3 p = m.pos[m.idx[3 * t + k]];
            by[iy].expand_point(p);
        }
        for (int k = 0; k < 3; k++) {
            V3 p = m.pos[m.idx[3 * t + k]];
            bz[iz].expand_point(p);
        }
    }
    auto best_split = [&](const int* c, const BBox* b, int& best_bin) {
        int lc[16], rc[16];
        BBox lb[16], rb[16];
        lc[0] = c[0];
        lb[0] = b[0];
        for (int i = 1; i < 16; i++) {
            lc[i] = lc[i - 1] + c[i];
            lb[i] = lb[i - 1];
            lb[i].expand_box_ref(b[i]);
        }
        rc[15] = c[15];
        rb[15] = b[15];
        for (int i = 14; i >= 0; i--) {
            rc[i] = rc[i + 1] + c[i];
            rb[i] = rb[i + 1];
            rb[i].expand_box_ref(b[i]);
        }
        lxr min_cost = FLT_MAX;
        best_bin = -1;
        for (int i = 0; i <= 14; i++) {
            lxr cost = 1 + (lxr(lc[i]) / lxr(cnt)) * lb[i].surface_area() +
                         (lxr(rc[i + 1]) / lxr(cnt)) * rb[i + 1].surface_area();
            if (cost < min_cost) {
                min_cost = cost;
                best_bin = i;
            }
        }
        return min_cost;
    };
    int bbx, bby, bbz;
    lxr cxm = best_split(cx, bx, bbx), cym = best_split(cy, by, bby), czm = best_split(cz, bz, bbz);
    int axis, best_bin;
    if (cxm <= cym && cxm <= czm) {
        axis = 0;
        best_bin = bbx;
    } else if (cym <= cxm && cym <= czm) {
        axis = 1;
        best_bin = bby;
    } else {
        axis = 2;
        best_bin = bbz;
    }
    auto pivot = std::partition(ti.begin() + start, ti.begin() + end + 1, [&](int t) {
        V3 c = centroid(m, t);
        int b;
        if (axis == 0) b = bin_of(c.x, mnx, wx);
        else if (axis == 1) b = bin_of(c.y, mny, wy);
        else b = bin_of(c.z, mnz, wz);
        return b <= best_bin;
    });
    int mid = (int)std::distance(ti.begin(), pivot);
    lxr lower = lxr(mid - start) / lxr(cnt);
    lxr upper = lxr(end - mid + 1) / lxr(cnt);
    if (lower < 0.15f || upper < 0.15f) {
        int new_mid = (start + end + 1) / 2;
        std::nth_element(ti.begin() + start, ti.begin() + new_mid, ti.begin() + end + 1, [&](int a, int b) {
            V3 ca = centroid(m, a), cb = centroid(m, b);
            lxr va = axis == 0 ? ca.x : (axis == 1 ? ca.y : ca.z);
            lxr vb = axis == 0 ? cb.x : (axis == 1 ? cb.y : cb.z);
            return va < vb;
        });
        mid = new_mid;
    }
    int left = alloc, right = alloc + 1;
    alloc += 2;
    BBox a = fill_bvh(left, start, mid - 1, h, m, ti, alloc, depth + 1, max_depth);
    BBox b = fill_bvh(right, mid, end, h, m, ti, alloc, depth + 1, max_depth);
    V3 pts[4] = {a.mx, a.mn, b.mx, b.mn};
    BNode n;
    n.leaf = false;
    n.box = BBox::of_points(pts, 4);
    n.tri = -1;
    n.left = left;
    h[idx] = n;
    return n.box;
}

/* --------------------------------------------------------------- scene */
struct Mat {
    V3 color;
    int albedo_tex;
    lx_tex_transform albedo_xf;
    lxr roughness, metallic;
    int mr_tex;
    lx_tex_transform mr_xf;
    int normal_tex;
    lx_tex_transform normal_xf;
    V3 emission_color;
    lxr emission_strength;
    int emissive_tex;
    lx_tex_transform emissive_xf;
    lxr ior, transmission, alpha;
    int alpha_mode;
    lxr alpha_cutoff;
    int type; /* LX_MATERIAL_*: legacy shading class */
};

struct Inst {
    uint32_t mesh, mat;
    lxr m[12], inv[12];
    BBox wbox; /* conservative world box, used only to skip instances */
};

} // namespace

struct lxo_scene {
    std::vector<Mesh> meshes;
    std::vector<Inst> insts;
    std::vector<Mat> mats;
    std::vector<Tex> texs;
    bool has_env = false;
    Tex env;
    lx_camera_desc cam;
    int trace_depth = 8;
    /* precompute_emissive_mesh_area, scene_precompute.cpp:22-74 */
    std::vector<uint32_t> emissive;         /* instance ids */
    std::vector<lxr> emissive_prefix;     /* normalised */
    std::vector<std::vector<lxr>> tri_cdf; /* per emissive instance */
    lxr total_area = 0.0f;
    /* precompute_hdri_emission, scene_precompute.cpp:76-144 */
    std::vector<lxr> marginal, conditional;
    double build_ms = 0;
};

namespace {

static lxr emission_len(const Mat& m) { return length(m.emission_color); }

/* scene_precompute.cpp:22-74; fp64 mode accumulates the prefix sums in
 * double and rounds the normalised values once. */
static void precompute_emitters(lxo_scene& s, int cdf_mode) {
    s.emissive.clear();
    s.emissive_prefix.clear();
    s.tri_cdf.clear();
    lxr area_f = 0.0f;
    double area_d = 0.0;
    std::vector<double> prefix_d;
    for (uint32_t gi = 0; gi < s.insts.size(); gi++) {
        const Inst& g = s.insts[gi];
        const Mat& mat = s.mats[g.mat];
        if (!(emission_len(mat) * mat.emission_strength > 1e-5f)) continue;
        const Mesh& m = s.meshes[g.mesh];
        size_t nt = m.idx.size() / 3;
        std::vector<lxr> cdf(nt);
        lxr cum_f = 0.0f;
        double cum_d = 0.0;
        std::vector<double> cd(nt);
        for (size_t t = 0; t < nt; t++) {
            V3 A = xform_point(g.m, m.pos[m.idx[3 * t + 0]]);
            V3 B = xform_point(g.m, m.pos[m.idx[3 * t + 1]]);
            V3 C = xform_point(g.m, m.pos[m.idx[3 * t + 2]]);
            lxr a = 0.5f * length(cross(sub(B, A), sub(C, A)));
            cum_f += a;
            cum_d += (double)a;
            cdf[t] = cum_f;
            cd[t] = cum_d;
        }
        if (cdf_mode == LX_CDF_FP32_REFERENCE) {
            for (size_t t = 0; t < nt; t++) cdf[t] /= cum_f;
            area_f += cum_f;
            s.emissive_prefix.push_back(area_f);
        } else {
            for (size_t t = 0; t < nt; t++) cdf[t] = (lxr)(cd[t] / cum_d);
            area_d += cum_d;
            prefix_d.push_back(area_d);
        }
        s.emissive.push_back(gi);
        s.tri_cdf.push_back(std::move(cdf));
    }
    if (cdf_mode == LX_CDF_FP32_REFERENCE) {
        for (auto& p : s.emissive_prefix) p /= area_f;
        s.total_area = area_f;
    } else {
        for (double p : prefix_d) s.emissive_prefix.push_back((lxr)(p / area_d));
        s.total_area = (lxr)area_d;
    }
}

/* scene_precompute.cpp:76-144.  Weight = length(vec4 texel) * sin(theta),
 * theta = PI * (y + 0.5) / H (alpha included, Appendix quirk). */
static void precompute_env(lxo_scene& s, int cdf_mode) {
    int W = s.env.w, H = s.env.h;
    s.conditional.assign((size_t)W * H, 0.0f);
    s.marginal.assign(H, 0.0f);
    lxr total_f = 0.0f;
    double total_d = 0.0;
    std::vector<double> marg_d(H);
    std::vector<double> row_d(W);
    for (int y = 0; y < H; y++) {
        lxr row_f = 0.0f;
        double rowsum_d = 0.0;
        lxr theta = PI_F * (y + 0.5f) / H;
        lxr st = lxm_sin(theta);
        for (int x = 0; x < W; x++) {
            const lxr* t = &s.env.rgba[((size_t)y * W + x) * 4];
            lxr e = lxm_sqrt(((t[0] * t[0] + t[1] * t[1]) + (t[2] * t[2] + t[3] * t[3])));
            lxr w = e * st;
            row_f += w;
            rowsum_d += (double)w;
            s.conditional[(size_t)y * W + x] = row_f;
            row_d[x] = rowsum_d;
        }
        if (cdf_mode == LX_CDF_FP32_REFERENCE) {
            if (row_f > 0.0f) {
                for (int x = 0; x < W; x++) s.conditional[(size_t)y * W + x] /= row_f;
            } else {
                for (int x = 0; x < W; x++) s.conditional[(size_t)y * W + x] = (lxr)(x + 1) / (lxr)W;
            }
            s.conditional[(size_t)y * W + (W - 1)] = 1.0f;
            total_f += row_f;
            s.marginal[y] = total_f;
        } else {
            if (rowsum_d > 0.0) {
                for (int x = 0; x < W; x++) s.conditional[(size_t)y * W + x] = (lxr)(row_d[x] / rowsum_d);
            } else {
                for (int x = 0; x < W; x++) s.conditional[(size_t)y * W + x] = (lxr)(x + 1) / (lxr)W;
            }
            s.conditional[(size_t)y * W + (W - 1)] = 1.0f;
            total_d += rowsum_d;
            marg_d[y] = total_d;
        }
    }
    if (cdf_mode == LX_CDF_FP32_REFERENCE) {
        for (int y = 0; y < H; y++) s.marginal[y] /= total_f;
    } else {
        for (int y = 0; y < H; y++) s.marginal[y] = (lxr)(marg_d[y] / total_d);
    }
    s.marginal[H - 1] = 1.0f;
}

/* ------------------------------------------------------ triangle tests */
struct Ray {
    V3 o, d;
    lxr tmax;
};

struct Hit {
    lxr t;
    uint32_t inst, prim;
    lxr u, v;
};

static inline bool better(lxr t, uint32_t inst, uint32_t prim, const Hit& h) {
    if (t < h.t) return true;
    if (t == h.t) {
        if (inst < h.inst) return true;
        if (inst == h.inst && prim < h.prim) return true;
    }
    return false;
}

/* Moller-Trumbore, src/intersections/mesh_isect.cu:77-137 (epsilon
 * 1.1920929e-7, inflated barycentric bounds, t > epsilon), plus t < tmax. */
static inline bool mt_test(const Mesh& m, uint32_t tri, V3 o, V3 d, lxr tmax, lxr& t, lxr& u, lxr& v) {
    const lxr eps = (lxr)(1.1920929E-7F);
    V3 a = m.pos[m.idx[3 * tri + 0]];
    V3 e1 = sub(m.pos[m.idx[3 * tri + 1]], a);
    V3 e2 = sub(m.pos[m.idx[3 * tri + 2]], a);
    V3 c = cross(d, e2);
    lxr det = dot(e1, c);
    if (det > -eps && det < eps) return false;
    lxr inv_det = 1.0f / det;
    a = sub(o, a);
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

static uint32_t hash_u32(uint32_t x) {
    /* optix_triangle.cu:41-49 */
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static inline V2 v2(lxr x, lxr y) { return V2{x, y}; }

/* __anyhit__ms_all, optix_triangle.cu:317-366: true if the candidate is
 * accepted.  dir is the world-space ray direction. */
static bool any_hit_accept(const lxo_scene& s, uint32_t inst, uint32_t prim, lxr bu, lxr bv, V3 wdir,
                           uint32_t key, uint32_t iteration) {
    const Inst& g = s.insts[inst];
    const Mat& mat = s.mats[g.mat];
    if (mat.alpha_mode == LX_ALPHA_OPAQUE) return true;
    lxr alpha = mat.alpha;
    if (mat.albedo_tex >= 0) {
        const Mesh& m = s.meshes[g.mesh];
        V2 uv = v2(0.0f, 0.0f);
        if (m.has_uv) {
            V2 A = m.uv[m.idx[3 * prim + 0]], B = m.uv[m.idx[3 * prim + 1]], C = m.uv[m.idx[3 * prim + 2]];
            lxr bA = 1.0f - bu - bv;
            uv = v2((bA * A.x + bu * B.x) + bv * C.x, (bA * A.y + bu * B.y) + bv * C.y);
        }
        uv = v2(uv.x * mat.albedo_xf.scale[0], uv.y * mat.albedo_xf.scale[1]);
        if (mat.albedo_xf.rotation != 0.0f) {
            lxr c = lxm_cos(mat.albedo_xf.rotation), sn = lxm_sin(mat.albedo_xf.rotation);
            uv = v2(c * uv.x - sn * uv.y, sn * uv.x + c * uv.y);
        }
        uv = v2(uv.x + mat.albedo_xf.offset[0], uv.y + mat.albedo_xf.offset[1]);
        lxr tex[4];
        tex_sample(s.texs[mat.albedo_tex], uv.x, uv.y, tex);
        alpha *= tex[3];
    }
    if (mat.alpha_mode == LX_ALPHA_MASK) return !(alpha < mat.alpha_cutoff);
    uint32_t h = hash_u32(key ^ hash_u32(iteration));
    h = hash_u32(h ^ lxm_f2u(wdir.x));
    h = hash_u32(h ^ lxm_f2u(wdir.y));
    h = hash_u32(h ^ lxm_f2u(wdir.z));
    h = hash_u32(h ^ inst);
    h = hash_u32(h ^ prim);
    lxr rnd = (h >> 8) * (1.0f / 16777216.0f);
    return !(rnd >= alpha);
}

/* RayBoxInterection, src/mesh/bounding_box.cu:7-40 */
static inline bool ray_box(const BBox& b, V3 o, V3 invd, const int* sign, lxr& thit) {
    V3 bounds[2] = {b.mn, b.mx};
    lxr t_min = (bounds[sign[0]].x - o.x) * invd.x;
    lxr t_max = (bounds[1 - sign[0]].x - o.x) * invd.x;
    lxr t_ymin = (bounds[sign[1]].y - o.y) * invd.y;
    lxr t_ymax = (bounds[1 - sign[1]].y - o.y) * invd.y;
    if (t_min > t_ymax || t_ymin > t_max) return false;
    t_min = lxm_max(t_min, t_ymin);
    t_max = lxm_min(t_max, t_ymax);
    lxr t_zmin = (bounds[sign[2]].z - o.z) * invd.z;
    lxr t_zmax = (bounds[1 - sign[2]].z - o.z) * invd.z;
    if (t_min > t_zmax || t_zmin > t_max) return false;
    t_min = lxm_max(t_min, t_zmin);
    if (t_max < 0.0f) return false;
    thit = t_min;
    return t_min < t_max;
}

struct TraceCtx {
    int alpha;
    uint32_t iteration;
    uint32_t key;
    bool any; /* stop at the first accepted hit */
};

/* Candidate handling shared by BVH and brute-force traversal. */
static inline bool consider(const lxo_scene& s, uint32_t inst, const Mesh& m, uint32_t tri, V3 o, V3 d, V3 wdir,
                            lxr tmax, Hit& best, const TraceCtx& tc) {
    lxr t, u, v;
    if (!mt_test(m, tri, o, d, tmax, t, u, v)) return false;
    if (!tc.any && !better(t, inst, tri, best)) return false;
    if (tc.alpha && !any_hit_accept(s, inst, tri, u, v, wdir, tc.key, tc.iteration)) return false;
    best.t = t;
    best.inst = inst;
    best.prim = tri;
    best.u = u;
    best.v = v;
    return true;
}

/* Per-instance DFS, src/intersections/mesh_isect.cu:6-149 (stack of node
 * indices, nearer child first, prune when the entry distance exceeds the
 * best t), with the deviations listed at the top of the file. */
static bool trace_instance_bvh(const lxo_scene& s, uint32_t gi, V3 wo, V3 wd, lxr tmax, Hit& best,
                               const TraceCtx& tc) {
    const Inst& g = s.insts[gi];
    const Mesh& m = s.meshes[g.mesh];
    V3 o = xform_point(g.inv, wo);
    V3 d = xform_vector(g.inv, wd);
    V3 invd = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int sign[3] = {invd.x < 0.0f ? 1 : 0, invd.y < 0.0f ? 1 : 0, invd.z < 0.0f ? 1 : 0};
    const std::vector<BNode>& bvh = m.bvh;
    std::vector<int> stack;
    stack.reserve(64);
    stack.push_back(0);
    bool found = false;
    while (!stack.empty()) {
        int ni = stack.back();
        stack.pop_back();
        const BNode& n = bvh[ni];
        if (!n.leaf) {
            int l = n.left, r = n.left + 1;
            lxr tl = -FLT_MAX, tr = -FLT_MAX;
            bool hl = bvh[l].leaf || ray_box(bvh[l].box, o, invd, sign, tl);
            bool hr = bvh[r].leaf || ray_box(bvh[r].box, o, invd, sign, tr);
            /* prune only beyond a 2^-16 relative margin so that box rounding
             * and triangle-test rounding can never change the closest hit */
            lxr cut = best.inst == 0xffffffffu ? FLT_MAX : best.t * 1.00001526f;
            if (hl && tl > cut) hl = false;
            if (hr && tr > cut) hr = false;
            if (hl && hr) {
                stack.push_back(tl <= tr ? r : l);
                stack.push_back(tl <= tr ? l : r);
            } else if (hl) {
                stack.push_back(l);
            } else if (hr) {
                stack.push_back(r);
            }
        } else {
            if (consider(s, gi, m, (uint32_t)n.tri, o, d, wd, tmax, best, tc)) {
                found = true;
                if (tc.any) return true;
            }
        }
    }
    return found;
}

static bool trace_instance_brute(const lxo_scene& s, uint32_t gi, V3 wo, V3 wd, lxr tmax, Hit& best,
                                 const TraceCtx& tc) {
    const Inst& g = s.insts[gi];
    const Mesh& m = s.meshes[g.mesh];
    V3 o = xform_point(g.inv, wo);
    V3 d = xform_vector(g.inv, wd);
    bool found = false;
    uint32_t nt = (uint32_t)(m.idx.size() / 3);
    for (uint32_t t = 0; t < nt; t++) {
        if (consider(s, gi, m, t, o, d, wd, tmax, best, tc)) {
            found = true;
            if (tc.any) return true;
        }
    }
    return found;
}

/* Conservative slab test used only to skip whole instances. */
static inline bool world_box_hit(const BBox& b, V3 o, V3 d, lxr tmax) {
    lxr t0 = 0.0f, t1 = tmax;
    const lxr* bo = &o.x;
    const lxr* bd = &d.x;
    const lxr* mn = &b.mn.x;
    const lxr* mx = &b.mx.x;
    for (int a = 0; a < 3; a++) {
        if (bd[a] == 0.0f) {
            if (bo[a] < mn[a] || bo[a] > mx[a]) return false;
            continue;
        }
        double inv = 1.0 / (double)bd[a];
        double ta = ((double)mn[a] - bo[a]) * inv, tb = ((double)mx[a] - bo[a]) * inv;
        if (ta > tb) std::swap(ta, tb);
        if (ta > t0) t0 = (lxr)ta;
        if (tb < t1) t1 = (lxr)(tb * (1.0 + 1e-6) + 1e-30);
        if (t0 > t1 * 1.000001f + 1e-6f) return false;
    }
    return true;
}

static Hit trace_scene(const lxo_scene& s, V3 o, V3 d, lxr tmax, int mode, const TraceCtx& tc) {
    Hit best;
    best.t = -1.0f;
    best.inst = 0xffffffffu;
    best.prim = 0xffffffffu;
    best.u = best.v = 0.0f;
    Hit cur = best;
    cur.t = FLT_MAX;
    for (uint32_t gi = 0; gi < s.insts.size(); gi++) {
        if (!world_box_hit(s.insts[gi].wbox, o, d, tmax)) continue;
        bool f = mode == LXO_TRAVERSE_BRUTE ? trace_instance_brute(s, gi, o, d, tmax, cur, tc)
                                            : trace_instance_bvh(s, gi, o, d, tmax, cur, tc);
        if (f && tc.any) return cur;
    }
    if (cur.inst != 0xffffffffu) return cur;
    return best;
}

/* ---------------------------------------------------- hit attributes */
struct Rec {
    lxr t;
    V3 N, T;
    int mat;
    V2 uv;
};

/* __closesthit__ch, optix_triangle.cu:156-243 */
static Rec closest_hit_attrs(const lxo_scene& s, const Hit& h) {
    const Inst& g = s.insts[h.inst];
    const Mesh& m = s.meshes[g.mesh];
    uint32_t i0 = m.idx[3 * h.prim + 0], i1 = m.idx[3 * h.prim + 1], i2 = m.idx[3 * h.prim + 2];
    V3 A = m.pos[i0], B = m.pos[i1], C = m.pos[i2];
    V3 e1 = sub(B, A), e2 = sub(C, A);
    V3 normal = normalize(cross(e1, e2));
    lxr bA = 1.0f - h.u - h.v;
    if (m.has_n) {
        V3 nA = m.nrm[i0], nB = m.nrm[i1], nC = m.nrm[i2];
        normal = normalize(add(add(smul(bA, nA), smul(h.u, nB)), smul(h.v, nC)));
    }
    V3 tangent = v3(1.0f, 0.0f, 0.0f);
    V2 uv = v2(0.0f, 0.0f);
    if (m.has_uv) {
        V2 a = m.uv[i0], b = m.uv[i1], c = m.uv[i2];
        uv = v2((bA * a.x + h.u * b.x) + h.v * c.x, (bA * a.y + h.u * b.y) + h.v * c.y);
        V2 d1 = v2(b.x - a.x, b.y - a.y), d2 = v2(c.x - a.x, c.y - a.y);
        lxr det = (d1.x * d2.y - d2.x * d1.y);
        lxr f = (lxm_abs(det) < 1e-10f) ? 0.0f : 1.0f / det;
        if (f == 0.0f) {
            tangent = v3(1.0f, 0.0f, 0.0f);
        } else {
            tangent = v3(f * (d2.y * e1.x - d1.y * e2.x), f * (d2.y * e1.y - d1.y * e2.y),
                         f * (d2.y * e1.z - d1.y * e2.z));
        }
    }
    normal = normalize(xform_normal(g.inv, normal));
    tangent = normalize(xform_vector(g.m, tangent));
    V3 T = sub(tangent, smul(dot(tangent, normal), normal));
    lxr len2 = dot(T, T);
    if (len2 > 1e-10f) {
        tangent = muls(T, 1.0f / lxm_sqrt(len2));
    } else {
        tangent = (lxm_abs(normal.x) > 0.9f) ? v3(0, 1, 0) : v3(1, 0, 0);
        tangent = normalize(cross(tangent, normal));
    }
    Rec r;
    r.t = h.t;
    r.N = normal;
    r.T = tangent;
    r.mat = (int)g.mat;
    r.uv = uv;
    return r;
}

/* __closesthit__ch_direct_light, optix_triangle.cu:259-305 */
static Rec dl_hit_attrs(const lxo_scene& s, const Hit& h) {
    const Inst& g = s.insts[h.inst];
    const Mesh& m = s.meshes[g.mesh];
    uint32_t i0 = m.idx[3 * h.prim + 0], i1 = m.idx[3 * h.prim + 1], i2 = m.idx[3 * h.prim + 2];
    lxr bA = 1.0f - h.u - h.v;
    V3 normal;
    if (m.has_n) {
        normal = normalize(add(add(smul(bA, m.nrm[i0]), smul(h.u, m.nrm[i1])), smul(h.v, m.nrm[i2])));
    } else {
        V3 A = m.pos[i0], B = m.pos[i1], C = m.pos[i2];
        normal = normalize(cross(sub(B, A), sub(C, A)));
    }
    V2 uv = v2(0.0f, 0.0f);
    if (m.has_uv) {
        V2 a = m.uv[i0], b = m.uv[i1], c = m.uv[i2];
        uv = v2((bA * a.x + h.u * b.x) + h.v * c.x, (bA * a.y + h.u * b.y) + h.v * c.y);
    }
    normal = normalize(xform_normal(g.inv, normal));
    Rec r;
    r.t = h.t;
    r.N = normal;
    r.T = v3(0, 0, 0);
    r.mat = (int)g.mat;
    r.uv = uv;
    return r;
}

/* ------------------------------------------------------- materials */
/* apply_texture_transform, src/material_queries.cu:3-17 */
static V2 apply_xf(V2 uv, const lx_tex_transform& x) {
    uv = v2(uv.x * x.scale[0], uv.y * x.scale[1]);
    if (x.rotation != 0.0f) {
        lxr c = lxm_cos(x.rotation), s = lxm_sin(x.rotation);
        uv = v2(c * uv.x - s * uv.y, s * uv.x + c * uv.y);
    }
    return v2(uv.x + x.offset[0], uv.y + x.offset[1]);
}

/* get_albedo, material_queries.cu:19-30 */
static V3 get_albedo(const lxo_scene& s, const Mat& m, V2 uv) {
    if (m.albedo_tex >= 0) {
        uv = apply_xf(uv, m.albedo_xf);
        lxr t[4];
        tex_sample(s.texs[m.albedo_tex], uv.x, uv.y, t);
        return v3(lxm_pow(t[0], 2.2f), lxm_pow(t[1], 2.2f), lxm_pow(t[2], 2.2f));
    }
    return m.color;
}

/* get_normal, material_queries.cu:33-89 (TBN = mat3(T, B, N); glm mat3 *
 * vec3 sums column-wise left to right) */
static V3 get_normal(const lxo_scene& s, const Mat& m, const Rec& r) {
    if (m.normal_tex < 0) return r.N;
    V2 uv = apply_xf(r.uv, m.normal_xf);
    lxr t[4];
    tex_sample(s.texs[m.normal_tex], uv.x, uv.y, t);
    V3 ln = v3(t[0] * 2.0f - 1.0f, t[1] * 2.0f - 1.0f, t[2] * 2.0f - 1.0f);
    V3 N = r.N;
    lxr l2 = dot(ln, ln);
    if (!(l2 > 1e-12f)) return N;
    V3 T = sub(r.T, smul(dot(r.T, N), N));
    lxr t2 = dot(T, T);
    if (!(t2 > 1e-12f)) return N;
    T = muls(T, 1.0f / lxm_sqrt(t2));
    V3 B = cross(N, T);
    V3 mapped = v3(T.x * ln.x + B.x * ln.y + N.x * ln.z, T.y * ln.x + B.y * ln.y + N.y * ln.z,
                   T.z * ln.x + B.z * ln.y + N.z * ln.z);
    lxr m2 = dot(mapped, mapped);
    if (!(m2 > 1e-12f)) return N;
    mapped = muls(mapped, 1.0f / lxm_sqrt(m2));
    if (!(dot(mapped, N) > 1e-4f)) return N;
    return mapped;
}

/* get_metallic_roughness, material_queries.cu:92-106 */
static void get_mr(const lxo_scene& s, const Mat& m, V2 uv, lxr& rough, lxr& metal) {
    rough = m.roughness;
    metal = m.metallic;
    if (m.mr_tex >= 0) {
        uv = apply_xf(uv, m.mr_xf);
        lxr t[4];
        tex_sample(s.texs[m.mr_tex], uv.x, uv.y, t);
        rough *= t[1];
        metal *= t[2];
    }
}

/* get_emission, material_queries.cu:108-120 */
static V3 get_emission(const lxo_scene& s, const Mat& m, V2 uv) {
    V3 e = muls(m.emission_color, m.emission_strength);
    if (m.emissive_tex >= 0) {
        uv = apply_xf(uv, m.emissive_xf);
        lxr t[4];
        tex_sample(s.texs[m.emissive_tex], uv.x, uv.y, t);
        e = mul(e, v3(lxm_pow(t[0], 2.2f), lxm_pow(t[1], 2.2f), lxm_pow(t[2], 2.2f)));
    }
    return e;
}

/* get_alpha, material_queries.cu:119-137 */
static lxr get_alpha(const lxo_scene& s, const Mat& m, V2 uv) {
    if (m.alpha_mode == LX_ALPHA_OPAQUE) return 1.0f;
    lxr alpha = m.alpha;
    if (m.albedo_tex >= 0) {
        V2 tuv = apply_xf(uv, m.albedo_xf);
        lxr t[4];
        tex_sample(s.texs[m.albedo_tex], tuv.x, tuv.y, t);
        alpha = alpha * t[3];
    }
    if (m.alpha_mode == LX_ALPHA_MASK) return alpha >= m.alpha_cutoff ? 1.0f : 0.0f;
    return alpha;
}

/* render_material_debug_mode, src/material_debug_render.cu:3-26, with the
 * normal-map texel get_normal writes to out_normal_map
 * (material_queries.cu:36-39 and 83). */
static V3 debug_view(const lxo_scene& s, const Mat& m, int mode, V3 albedo, V3 normal, V2 uv, lxr rough,
                     lxr metal) {
    if (mode == LX_DEBUG_ALBEDO) return albedo;
    if (mode == LX_DEBUG_WORLD_NORMAL) return add(muls(normal, 0.5f), v3(0.5f, 0.5f, 0.5f));
    if (mode == LX_DEBUG_NORMAL_MAP) {
        if (m.normal_tex < 0) return v3(0.5f, 0.5f, 1.0f);
        V2 tuv = apply_xf(uv, m.normal_xf);
        lxr t[4];
        tex_sample(s.texs[m.normal_tex], tuv.x, tuv.y, t);
        return v3(t[0], t[1], t[2]);
    }
    if (mode == LX_DEBUG_ROUGHNESS) return v3(rough, rough, rough);
    if (mode == LX_DEBUG_METALLIC) return v3(metal, metal, metal);
    if (mode == LX_DEBUG_ALPHA) {
        lxr a = get_alpha(s, m, uv);
        return v3(a, a, a);
    }
    return v3(0.0f, 0.0f, 0.0f);
}

/* ------------------------------------------------------ Cook-Torrance */
/* cook_torrance.cu:8-25 */
static lxr D_TR(V3 h, V3 n, lxr alpha) {
    lxr ndh = dot(n, h);
    if (ndh <= 0.0f) return 0.0f;
    lxr a2 = alpha * alpha;
    lxr ndh2 = ndh * ndh;
    lxr dc = (ndh2 * a2) + (1.0f - ndh2);
    lxr den = dc * dc * GLM_PI_F;
    return a2 / lxm_max(den, CT_EPS);
}
/* cook_torrance.cu:27-31 */
static V3 F_schlick(lxr vdh, V3 f0) {
    lxr m = 1.0f - vdh;
    lxr m2 = m * m;
    return add(f0, muls(sub(v3(1, 1, 1), f0), m2 * m2 * m));
}
/* cook_torrance.cu:34-50 */
static lxr F_diel(lxr cosI, lxr etaI, lxr etaT) {
    cosI = lxm_clamp(cosI, 0.0f, 1.0f);
    lxr sinI = lxm_sqrt(lxm_max(0.0f, 1.0f - cosI * cosI));
    lxr sinT = (etaI / etaT) * sinI;
    if (sinT >= 1.0f) return 1.0f;
    lxr cosT = lxm_sqrt(lxm_max(0.0f, 1.0f - sinT * sinT));
    lxr rpar = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    lxr rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return (rpar * rpar + rperp * rperp) / 2.0f;
}
/* cook_torrance.cu:56-66 */
static lxr G1(V3 w, V3 n, lxr alpha) {
    lxr a2 = alpha * alpha;
    lxr ndw = lxm_max(dot(n, w), 0.0f);
    lxr num = 2.0f * ndw;
    lxr under = a2 + (1.0f - a2) * (ndw * ndw);
    lxr den = ndw + lxm_sqrt(under);
    return num / lxm_max(den, CT_EPS);
}
/* cook_torrance.cu:72-77 (glm::mix(x, y, a) = x + a * (y - x)) */
static V3 Fresnel(lxr c, V3 albedo, lxr metallic, lxr etaI, lxr etaT) {
    lxr fd = F_diel(c, etaI, etaT);
    V3 diel = v3(fd, fd, fd);
    V3 cond = F_schlick(c, albedo);
    return add(diel, smul(metallic, sub(cond, diel)));
}
/* cook_torrance.cu:79-110 */
static V3 BRDF(V3 v, V3 n, V3 l, V3 albedo, lxr roughness, lxr metallic, lxr ior, lxr transmission) {
    ior = lxm_max(ior, 1.0001f);
    bool enter = dot(v, n) > 0.0f;
    V3 nf = enter ? n : neg(n);
    lxr etaI = enter ? 1.0f : ior, etaT = enter ? ior : 1.0f;
    V3 h = normalize(add(v, l));
    roughness = lxm_clamp(roughness, 0.0001f, 1.0f);
    lxr alpha = roughness * roughness;
    lxr D = D_TR(h, nf, alpha);
    V3 F = Fresnel(lxm_max(dot(v, h), 0.0f), albedo, metallic, etaI, etaT);
    lxr G = G1(v, nf, alpha) * G1(l, nf, alpha);
    V3 num = muls(smul(D, F), G);
    lxr den = 4 * lxm_max(dot(nf, v), 0.0f) * lxm_max(dot(nf, l), 0.0f);
    V3 spec = divs(num, lxm_max(den, CT_EPS));
    V3 nonspec = sub(v3(1, 1, 1), F);
    V3 kD = muls(smul(1.0f - metallic, nonspec), 1.0f - transmission);
    V3 diffuse = muls(mul(kD, albedo), INV_PI_CT);
    return add(diffuse, spec);
}
/* cook_torrance.cu:112-142 */
static V3 BTDF(V3 wo, V3 n, V3 wi, V3 albedo, lxr roughness, lxr metallic, lxr ior, lxr transmission) {
    lxr cosI = dot(wo, n);
    bool enter = cosI > 0.f;
    lxr etaI = enter ? 1.0f : ior, etaT = enter ? ior : 1.0f;
    V3 ht = normalize(neg(add(smul(etaI, wo), smul(etaT, wi))));
    lxr jd = (etaT * dot(wi, ht) + etaI * dot(wo, ht));
    lxr jac = (etaI * etaI) / lxm_max(jd * jd, CT_EPS);
    lxr ctn = lxm_abs(dot(wi, ht)) * lxm_abs(dot(wo, ht));
    lxr ctd = lxm_abs(dot(wi, n)) * lxm_abs(dot(wo, n));
    roughness = lxm_clamp(roughness, 0.0001f, 1.0f);
    lxr alpha = roughness * roughness;
    lxr D = D_TR(ht, n, alpha);
    lxr F = F_diel(lxm_abs(dot(wo, ht)), etaI, etaT);
    bool gv = dot(wi, ht) * dot(wi, n) > 0.0f && dot(wo, ht) * dot(wo, n) > 0.0f;
    lxr G = gv ? G1(dot(wi, n) > 0.0f ? wi : neg(wi), n, alpha) * G1(dot(wo, n) > 0.0f ? wo : neg(wo), n, alpha)
                 : 0.0f;
    lxr DFG = D * G * (1.0f - F);
    return muls(albedo, transmission * (1.0f - metallic) * (ctn / ctd) * jac * DFG);
}

static void lobe_probs(V3 wo, V3 n, V3 color, lxr metallic, lxr ior, lxr transmission, lxr& pRefr,
                       lxr& pDiff, lxr& pSpec) {
    bool enter = dot(wo, n) > 0.0f;
    lxr etaI = enter ? 1.0f : ior, etaT = enter ? ior : 1.0f;
    lxr cm = lxm_abs(dot(wo, n));
    V3 Fa = Fresnel(cm, color, metallic, etaI, etaT);
    lxr sr = lxm_clamp((Fa.x + Fa.y + Fa.z) / 3.0f, 0.01f, 0.99f);
    lxr Fm = F_diel(cm, etaI, etaT);
    pRefr = transmission * (1.0f - metallic) * (1.0f - Fm);
    pDiff = (1.0f - pRefr) * (1.0f - sr) * (1.0f - metallic);
    pSpec = 1.0f - pRefr - pDiff;
}

/* calculateRandomDirectionInHemisphere, src/interactions.cu:7-45 */
static V3 cosine_hemisphere(V3 n, Rng& rng) {
    lxr up = lxm_sqrt(rng.u01());
    lxr over = lxm_sqrt(1 - up * up);
    lxr around = rng.u01() * TWO_PI_F;
    V3 dnn;
    if (lxm_abs(n.x) < SQRT_OF_ONE_THIRD) dnn = v3(1, 0, 0);
    else if (lxm_abs(n.y) < SQRT_OF_ONE_THIRD) dnn = v3(0, 1, 0);
    else dnn = v3(0, 0, 1);
    V3 p1 = normalize(cross(n, dnn));
    V3 p2 = normalize(cross(n, p1));
    return add(add(smul(up, n), smul(lxm_cos(around) * over, p1)), smul(lxm_sin(around) * over, p2));
}

/* sampleCookTorrance, cook_torrance.cu:156-224 */
static V3 sample_ct(V3 wo, V3 n, lxr roughness, lxr metallic, Rng& rng, V3 color, lxr ior, lxr transmission) {
    lxr r = rng.u01();
    ior = lxm_max(ior, 1.0001f);
    roughness = lxm_clamp(roughness, 0.0001f, 1.0f);
    lxr alpha = roughness * roughness;
    bool enter = dot(wo, n) > 0.0f;
    V3 nf = enter ? n : neg(n);
    lxr etaI = enter ? 1.0f : ior, etaT = enter ? ior : 1.0f;
    lxr pRefr, pDiff, pSpec;
    lobe_probs(wo, n, color, metallic, ior, transmission, pRefr, pDiff, pSpec);
    if (r > pSpec && r <= pSpec + pDiff) return cosine_hemisphere(nf, rng);
    lxr x1 = rng.u01();
    lxr x2 = rng.u01();
    lxr th = lxm_atan((alpha * lxm_sqrt(x1)) / lxm_sqrt(1.0f - x1));
    lxr ph = 2.0f * PI_F * x2;
    /* sphericalToCartesian(vec3(1, theta, phi)), cook_torrance.cu:144-154 */
    lxr st = lxm_sin(th);
    V3 hs = v3(1.0f * st * lxm_cos(ph), 1.0f * st * lxm_sin(ph), 1.0f * lxm_cos(th));
    V3 h = normalize(hs);
    V3 t = normalize(lxm_abs(nf.z) < 0.999f ? cross(v3(0, 0, 1), nf) : cross(v3(1, 0, 0), nf));
    V3 b = cross(nf, t);
    V3 hw = add(add(smul(h.x, t), smul(h.y, b)), smul(h.z, nf));
    V3 I = neg(wo);
    if (r <= pSpec) {
        /* glm::reflect: I - N * dot(N, I) * 2 */
        V3 wi = sub(I, muls(muls(hw, dot(hw, I)), 2.0f));
        if (dot(wi, nf) <= 0.0f) return v3(0, 0, 0);
        return normalize(wi);
    }
    /* glm::refract(I, N, eta) */
    lxr eta = etaI / etaT;
    lxr dv = dot(hw, I);
    lxr k = 1.0f - eta * eta * (1.0f - dv * dv);
    if (!(k >= 0.0f)) return v3(0, 0, 0); /* glm yields NaN * 0 -> rejected by isnan */
    V3 wi = sub(smul(eta, I), smul(eta * dv + lxm_sqrt(k), hw));
    if (dot(wi, wi) < CT_EPS || lxm_isnan(wi.x) || lxm_isnan(wi.y) || lxm_isnan(wi.z) || dot(wi, nf) >= 0.0f)
        return v3(0, 0, 0);
    return normalize(wi);
}

/* PDF_GGX, cook_torrance.cu:226-237 */
static lxr pdf_ggx(V3 wo, V3 wi, V3 n, lxr roughness) {
    V3 h = normalize(add(wo, wi));
    roughness = lxm_clamp(roughness, 0.0001f, 1.0f);
    lxr alpha = roughness * roughness;
    lxr ph = D_TR(h, n, alpha) * dot(n, h);
    return ph / lxm_max((4 * lxm_abs(dot(wo, h))), CT_EPS);
}
/* PDF_GGX_Transmission, cook_torrance.cu:239-258 */
static lxr pdf_ggx_t(V3 wo, V3 wi, V3 n, lxr roughness, lxr ior) {
    lxr cosI = dot(wo, n);
    bool enter = cosI > 0.f;
    lxr etaI = enter ? 1.0f : ior, etaT = enter ? ior : 1.0f;
    V3 ht = normalize(neg(add(smul(etaI, wo), smul(etaT, wi))));
    roughness = lxm_clamp(roughness, 0.0001f, 1.0f);
    lxr alpha = roughness * roughness;
    lxr den = etaT * dot(wi, ht) + etaI * dot(wo, ht);
    lxr ph = D_TR(ht, n, alpha) * lxm_abs(dot(n, ht));
    return ph * (etaT * etaT * lxm_abs(dot(wi, ht))) / lxm_max(den * den, CT_EPS);
}
/* PDF, cook_torrance.cu:260-295 */
static lxr PDF(V3 wo, V3 wi, V3 n, lxr roughness, lxr metallic, V3 color, lxr ior, lxr transmission) {
    ior = lxm_max(ior, 1.0001f);
    V3 nf = dot(wo, n) < 0.0f ? neg(n) : n;
    lxr dwn = dot(wi, nf);
    lxr pd = (dwn > 0.0f ? dwn : 0.0f) * INV_PI_CT;
    bool refract = (dot(wi, n) * dot(wo, n)) < 0.0f;
    lxr pRefr, pDiff, pSpec;
    lobe_probs(wo, n, color, metallic, ior, transmission, pRefr, pDiff, pSpec);
    lxr pdf;
    if (refract) {
        pdf = pRefr * pdf_ggx_t(wo, wi, n, roughness, ior);
    } else {
        pdf = pDiff * pd + pSpec * pdf_ggx(wo, wi, nf, roughness);
    }
    return lxm_max(pdf, PDF_CLAMP);
}
/* shadePathCookTorrance, cook_torrance.cu:298-331 */
static V3 ct_eval_cos(V3 wo, V3 normal, V3 wi, V3 albedo, lxr roughness, lxr metallic, lxr ior,
                      lxr transmission) {
    ior = lxm_max(ior, 1.0001f);
    bool refract = (dot(wi, normal) * dot(wo, normal)) < 0.0f;
    V3 f = refract ? BTDF(wo, normal, wi, albedo, roughness, metallic, ior, transmission)
                   : BRDF(wo, normal, wi, albedo, roughness, metallic, ior, transmission);
    return muls(f, lxm_abs(dot(wi, normal)));
}

/* ---------------------------------------- legacy shaders (UBER_SHADER 0) */
/* glm::reflect: I - N * dot(N, I) * 2 */
static V3 glm_reflect(V3 I, V3 N) { return sub(I, muls(muls(N, dot(N, I)), 2.0f)); }

/* Lambert::PDF, src/shaders/lambert.cu:9-12: max(1e-6, max(0, cos) / pi) */
static lxr lambert_pdf(V3 wi, V3 n) {
    lxr c = dot(wi, n);
    if (!(c > 0.0f)) c = 0.0f;
    lxr p = c * INV_PI_CT;
    return p > 1e-6f ? p : 1e-6f;
}

/* TransmissiveGlass::FresnelDielectricEval, src/shaders/glass.cu:27-55 */
static lxr glass_fresnel(lxr cos_i, lxr ior) {
    lxr eta_i = 1.0f, eta_t = ior;
    cos_i = lxm_clamp(cos_i, -1.0f, 1.0f);
    if (!(cos_i > 0.0f)) {
        eta_i = ior;
        eta_t = 1.0f;
        cos_i = lxm_abs(cos_i);
    }
    lxr sin_i = lxm_sqrt(lxm_max(0.0f, 1.0f - cos_i * cos_i));
    lxr sin_t = (eta_i / eta_t) * sin_i;
    if (sin_t >= 1.0f) return 1.0f; /* total internal reflection */
    lxr cos_t = lxm_sqrt(lxm_max(0.0f, 1.0f - sin_t * sin_t));
    lxr par = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    lxr per = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return (par * par + per * per) / 2.0f;
}

/* glm::refract(I, N, eta) as the vector overload computes it: NaN when k < 0
 * (sqrt of a negative times 0). */
static V3 glm_refract(V3 I, V3 N, lxr eta) {
    lxr dv = dot(N, I);
    lxr k = 1.0f - eta * eta * (1.0f - dv * dv);
    lxr mask = k >= 0.0f ? 1.0f : 0.0f;
    lxr root = k >= 0.0f ? lxm_sqrt(k) : std::numeric_limits<lxr>::quiet_NaN();
    return muls(sub(smul(eta, I), smul(eta * dv + root, N)), mask);
}

/* sampleGlass (glass.cu:57-75) with sampleSpecularRefl / sampleSpecularTrans
 * (glass.cu:5-25) */
static V3 sample_glass(V3 dir, V3 n, lxr ior, Rng& rng) {
    V3 wo = neg(dir);
    lxr r = rng.u01();
    lxr F = lxm_clamp(glass_fresnel(dot(wo, n), ior), 0.0f, 1.0f);
    bool enter = dot(wo, n) > 0.0f;
    V3 loc = enter ? n : neg(n);
    if (r < F) return glm_reflect(neg(wo), loc);
    V3 wi = glm_refract(neg(wo), loc, enter ? (1.0f / ior) : (ior / 1.0f));
    if (dot(wi, wi) < 1e-6f || lxm_isnan(wi.x) || lxm_isnan(wi.y) || lxm_isnan(wi.z)) return glm_reflect(neg(wo), loc);
    return normalize(wi);
}

/* sample_materials, src/sample_materials.cu:24-37, non-microfacet classes */
static V3 legacy_sample(int type, V3 dir, V3 n, lxr ior, Rng& rng) {
    switch (type) {
    case LX_MATERIAL_DIFFUSE:
    case LX_MATERIAL_EMISSIVE: return cosine_hemisphere(n, rng); /* Lambert::sampleHemisphere */
    case LX_MATERIAL_SPECULAR: return glm_reflect(neg(neg(dir)), n); /* PerfectSpecular::sampleMirror */
    default: return sample_glass(dir, n, ior, rng);
    }
}

/* update_throughput_materials, src/update_throughput_materials.cu:36-62,
 * non-microfacet classes; returns the new last_pdf. */
static lxr legacy_update(int type, V3 dir, V3 n, V3 wi, V3 color, lxr ior, lxr transmission, V3& thr,
                           lxr last_pdf) {
    if (type == LX_MATERIAL_DIFFUSE) {
        lxr c = dot(wi, n);
        V3 f = muls(muls(color, INV_PI_CT), c > 0.0f ? c : 0.0f);
        lxr pdf = lambert_pdf(wi, n);
        thr = mul(thr, divs(f, pdf));
        return pdf;
    }
    if (type == LX_MATERIAL_SPECULAR) {
        thr = mul(thr, color);
        return 1.0f;
    }
    if (type == LX_MATERIAL_GLASS) {
        V3 wo = neg(dir);
        bool enter = dot(wo, n) > 0.f;
        lxr eta_i = enter ? 1.0f : ior;
        lxr eta_t = enter ? ior : 1.0f;
        bool refracted = dot(wo, n) * dot(wi, n) < 0.0f;
        lxr att = refracted ? (eta_i / eta_t) : 1.0f;
        att = att * att;
        V3 one = v3(1.0f, 1.0f, 1.0f);
        V3 tint = refracted ? add(one, smul(transmission, sub(color, one))) : one; /* glm::mix(1, color, T) */
        thr = mul(thr, smul(att, tint));
        return 1.0f;
    }
    return last_pdf;
}

/* ------------------------------------------------------- CDF sampling */
/* cudaUtils::select_from_cdf, src/common.cu:3-15 */
static int select_from_cdf(const lxr* cdf, int size, lxr xi) {
    int lo = 0, hi = size - 1;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (cdf[mid] < xi) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* envmap_pdf, shade_path.cu:27-42 */
static lxr envmap_pdf(const lxo_scene& s, V3 d) {
    int W = s.env.w, H = s.env.h;
    lxr phi = lxm_atan2(d.z, d.x);
    lxr theta = lxm_acos(lxm_clamp(d.y, -1.0f, 1.0f));
    lxr u = (phi + PI_F) / (2.0f * PI_F);
    lxr v = theta / PI_F;
    int col = (int)(u * W);
    col = col < 0 ? 0 : (col > W - 1 ? W - 1 : col);
    int row = (int)(v * H);
    row = row < 0 ? 0 : (row > H - 1 ? H - 1 : row);
    lxr pm = (s.marginal[row] - (row == 0 ? 0.0f : s.marginal[row - 1])) * H;
    lxr pc = (s.conditional[(size_t)row * W + col] - (col == 0 ? 0.0f : s.conditional[(size_t)row * W + col - 1])) * W;
    lxr st = lxm_sin(theta);
    return (st > 1e-6f) ? (pm * pc) / (2.0f * PI_F * PI_F * st) : 0.0f;
}

/* Environment texel lookup of shade_path.cu:205-215 / 309-318: note the
 * double-precision M_PI = 3.14159 of shade_path.cu:23. */
static V3 env_lookup(const lxo_scene& s, V3 d, lxr intensity) {
    lxr phi = lxm_atan2(d.z, d.x);
    lxr theta = lxm_acos(lxm_clamp(d.y, -1.0f, 1.0f));
    lxr u = (lxr)(((double)phi + SHADE_M_PI) * (double)(1.0f / (2.0f * SHADE_M_PI)));
    lxr v = (lxr)((double)theta * (double)(1.0f / SHADE_M_PI));
    lxr t[4];
    tex_sample(s.env, u, v, t);
    return muls(v3(t[0], t[1], t[2]), intensity);
}

/* sampleDirectLight, src/kernels/sample_direct_light.cu:32-66 (emitter) */
static V3 sample_light_point(const lxo_scene& s, Rng& rng) {
    lxr rnd = rng.u01();
    int sel = select_from_cdf(s.emissive_prefix.data(), (int)s.emissive.size(), rnd);
    const Inst& g = s.insts[s.emissive[sel]];
    const Mesh& m = s.meshes[g.mesh];
    lxr lower = (sel == 0) ? 0.0f : s.emissive_prefix[sel - 1];
    lxr upper = s.emissive_prefix[sel];
    lxr den = upper - lower;
    lxr to = (den > 1e-10f) ? (rnd - lower) / den : 0.0f;
    to = lxm_clamp(to, 0.0f, 1.0f);
    const std::vector<lxr>& cdf = s.tri_cdf[sel];
    int tri = select_from_cdf(cdf.data(), (int)cdf.size(), to);
    V3 v0 = m.pos[m.idx[3 * tri + 0]], v1 = m.pos[m.idx[3 * tri + 1]], v2 = m.pos[m.idx[3 * tri + 2]];
    lxr r1 = lxm_sqrt(rng.u01());
    lxr r2 = rng.u01();
    lxr u = 1.0f - r1;
    lxr v = r2 * r1;
    V3 local = add(add(smul(u, v0), smul(v, v1)), smul(1.0f - u - v, v2));
    return xform_point(g.m, local);
}

/* sampleDirectLight, sample_direct_light.cu:68-100 (environment) */
static V3 sample_env_dir(const lxo_scene& s, Rng& rng) {
    int W = s.env.w, H = s.env.h;
    lxr rnd = rng.u01();
    int mg = select_from_cdf(s.marginal.data(), H, rnd);
    lxr lower = (mg == 0) ? 0.0f : s.marginal[mg - 1];
    lxr upper = s.marginal[mg];
    lxr den = upper - lower;
    lxr off = (den > 1e-10f) ? (rnd - lower) / den : 0.0f;
    off = lxm_clamp(off, 0.0f, 1.0f);
    int cd = select_from_cdf(s.conditional.data() + (size_t)mg * W, W, off);
    lxr u = (cd + 0.5f) / W;
    lxr v = (mg + 0.5f) / H;
    lxr phi = (lxr)((double)u * 2.0 * (double)PI_F - (double)PI_F);
    lxr theta = v * PI_F;
    lxr st = lxm_sin(theta);
    return v3(st * lxm_cos(phi), lxm_cos(theta), st * lxm_sin(phi));
}

/* --------------------------------------------------------- render loop */
struct Path {
    V3 o, d;
    V3 color, thr;
    V3 sample_dir, dl_sample, env_sample;
    uint32_t pixel;
    bool kill, spec;
    lxr last_pdf;
};

struct BounceRecords {
    Rec isect, dl;
    lxr env_t;
};

struct RenderCfg {
    int W, H, depth;
    bool dmis, emis;
    int debug_mode; /* LX_DEBUG_* */
    bool legacy;    /* LX_SHADING_LEGACY */
    lxr intensity;
    uint32_t seed;
    lx_camera_desc cam;
    int mode;
};

/* generateRayFromCamera, src/kernels/generate_ray_from_camera.cu:8-39 */
static void camera_ray(const lx_camera_desc& c, int iter, int x, int y, int trace_depth, uint32_t seed, V3& dir) {
    int index = x + y * c.width;
    Rng rng = make_rng(iter, index, trace_depth, seed);
    lxr x1 = rng.u01() - 0.5f;
    lxr x2 = rng.u01() - 0.5f;
    lxr pX = (lxr(x) + x1 + 0.5f) - (lxr)c.width * 0.5f;
    lxr pY = (lxr(y) + x2 + 0.5f) - (lxr)c.height * 0.5f;
    V3 view = v3(c.view[0], c.view[1], c.view[2]);
    V3 right = v3(c.right[0], c.right[1], c.right[2]);
    V3 up = v3(c.up[0], c.up[1], c.up[2]);
    dir = normalize(sub(sub(view, muls(muls(right, c.pixel_length[0]), pX)), muls(muls(up, c.pixel_length[1]), pY)));
}

struct Counters {
    uint64_t rad = 0, dl = 0, env = 0, bounces = 0;
};

/* shadePath, src/kernels/shade_path.cu:94-364, for one path; returns true if
 * the path is pushed to the next queue. */
static bool shade_one(const lxo_scene& s, const RenderCfg& cfg, Path& p, const BounceRecords& br, int iter, int depth,
                      bool compact) {
    Rng rng = make_rng(iter, (int)p.pixel, depth, cfg.seed);
    const Rec& is = br.isect;
    bool has_exr = s.has_env;
    V3 dir = p.d;
    if (is.t > 0.0f) {
        const Mat& mat = s.mats[is.mat];
        /* UBER_SHADER 1 treats every material as microfacet; the legacy build
         * dispatches on the class (shade_path.cu:44-92) */
        int type = cfg.legacy ? mat.type : LX_MATERIAL_MICROFACET;
        V3 color = get_albedo(s, mat, is.uv);
        if (type == LX_MATERIAL_EMISSIVE && mat.albedo_tex < 0) color = mat.emission_color; /* gltf.cpp:211-213 */
        V3 normal = get_normal(s, mat, is);
        lxr rough, metal;
        get_mr(s, mat, is.uv, rough, metal);
        V3 emission = get_emission(s, mat, is.uv);
        bool is_specular = cfg.legacy ? (type == LX_MATERIAL_SPECULAR || type == LX_MATERIAL_GLASS) : rough < 0.05f;
        V3 nf = dot(neg(dir), normal) < 0.0f ? neg(normal) : normal;
        if (cfg.debug_mode != 0) {
            /* render_material_debug_mode: the view replaces the colour, the path ends */
            p.color = debug_view(s, mat, cfg.debug_mode, color, normal, is.uv, rough, metal);
            p.kill = true;
        } else {
        bool micro = type == LX_MATERIAL_MICROFACET;
        if (micro) p.sample_dir = sample_ct(neg(dir), normal, rough, metal, rng, color, mat.ior, mat.transmission);
        else p.sample_dir = legacy_sample(type, dir, normal, mat.ior, rng);
        bool emissive = cfg.legacy ? type == LX_MATERIAL_EMISSIVE : length(emission) > 0.00001f;
        /* get_pdf_bsdf / get_brdf, shade_path.cu:60-92 */
        auto pdf_bsdf = [&](V3 wi) {
            return micro ? PDF(neg(dir), wi, normal, rough, metal, color, mat.ior, mat.transmission)
                         : (type == LX_MATERIAL_DIFFUSE ? lambert_pdf(wi, normal) : 0.0f);
        };
        auto brdf_of = [&](V3 wi) {
            return micro ? BRDF(neg(dir), normal, wi, color, rough, metal, mat.ior, mat.transmission)
                         : (type == LX_MATERIAL_DIFFUSE ? divs(color, PI_F) : v3(0.0f, 0.0f, 0.0f));
        };
        if (cfg.emis) {
            if (emissive) {
                lxr w = 1.0f;
                if (depth > 1 && !p.spec) {
                    lxr pb = p.last_pdf;
                    lxr pe = envmap_pdf(s, dir);
                    w = (pb * pb) / (pb * pb + pe * pe);
                }
                p.color = add(p.color, mul(p.thr, smul(w, emission)));
                p.kill = true;
            } else if (br.env_t < 0.0f && !is_specular && !p.kill && has_exr) {
                V3 ed = p.env_sample;
                lxr pe = envmap_pdf(s, ed);
                lxr pb = pdf_bsdf(ed);
                lxr w = (pe * pe) / (pe * pe + pb * pb);
                V3 brdf = brdf_of(ed);
                V3 L = env_lookup(s, normalize(ed), cfg.intensity);
                lxr cs = dot(nf, ed);
                if (cs > 0.0f && pe > 1e-8f) {
                    V3 contrib = divs(muls(mul(L, brdf), cs), pe);
                    p.color = add(p.color, muls(mul(p.thr, contrib), w));
                }
            }
        } else if (cfg.dmis) {
            if (emissive) {
                lxr w = 1.0f;
                if (depth > 1 && !p.spec) {
                    lxr pb = p.last_pdf;
                    lxr dist_sq = is.t * is.t;
                    lxr cl = dot(is.N, neg(dir));
                    if (cl > 0.0001f) {
                        lxr pa = 1.0f / s.total_area;
                        lxr psa = pa * (dist_sq / cl);
                        w = (pb * pb) / (pb * pb + psa * psa);
                    } else {
                        w = 0.0f;
                    }
                }
                p.color = add(p.color, mul(p.thr, smul(w, emission)));
                p.kill = true;
            } else if (br.dl.t > 0.0f && !is_specular && !p.kill) {
                V3 sp = add(smul(is.t, dir), p.o);
                V3 lv = sub(p.dl_sample, sp);
                lxr dsq = dot(lv, lv);
                lxr dist = lxm_sqrt(dsq);
                V3 ld = divs(lv, dist);
                lxr cs = dot(nf, ld);
                lxr cl = dot(br.dl.N, neg(ld));
                if (cs > 0.0f && cl > 0.0f) {
                    lxr pa = 1.0f / s.total_area;
                    lxr psa = pa * (dsq / cl);
                    lxr pb = pdf_bsdf(ld);
                    lxr w = (psa * psa) / (psa * psa + pb * pb);
                    V3 brdf = brdf_of(ld);
                    V3 Le = get_emission(s, s.mats[br.dl.mat], br.dl.uv);
                    V3 contrib = divs(muls(mul(Le, brdf), cs), psa);
                    p.color = add(p.color, muls(mul(p.thr, contrib), w));
                }
            }
        } else if (emissive) {
            p.color = add(p.color, mul(p.thr, emission));
            p.kill = true;
        }
        if (!p.kill) {
            /* update_throughput_materials, update_throughput_materials.cu:8-67 */
            if (micro) {
                bool transmitted = !cfg.legacy && dot(neg(dir), normal) * dot(p.sample_dir, normal) < 0.0f;
                bool spec = is_specular || transmitted;
                V3 ct = ct_eval_cos(neg(dir), normal, p.sample_dir, color, rough, metal, mat.ior, mat.transmission);
                lxr pdf = PDF(neg(dir), p.sample_dir, normal, rough, metal, color, mat.ior, mat.transmission);
                p.thr = mul(p.thr, divs(ct, pdf));
                p.last_pdf = pdf;
                p.spec = spec;
            } else {
                p.last_pdf = legacy_update(type, dir, normal, p.sample_dir, color, mat.ior, mat.transmission, p.thr,
                                           p.last_pdf);
                p.spec = is_specular;
            }
            if (depth >= 3) {
                lxr mt = lxm_max(p.thr.x, lxm_max(p.thr.y, p.thr.z));
                lxr sv = lxm_clamp(mt, 0.0f, 1.0f);
                if (rng.u01() >= sv) p.kill = true;
                else p.thr = divs(p.thr, sv);
            }
        }
        }
    } else if (!p.kill && has_exr) {
        V3 d = normalize(dir);
        V3 L = env_lookup(s, d, cfg.intensity);
        lxr w = 1.0f;
        if (cfg.emis && depth > 1 && !p.spec) {
            lxr pb = p.last_pdf;
            lxr pe = envmap_pdf(s, d);
            w = (pb * pb) / (pb * pb + pe * pe);
        }
        p.color = add(p.color, muls(mul(p.thr, L), w));
    }
    if (is.t == -1.0f) {
        p.kill = true;
    } else {
        /* getPointOnRay, include/intersections.h:29-32 */
        V3 hp = add(p.o, smul(is.t - .0001f, normalize(dir)));
        V3 n = is.N;
        lxr l2 = dot(p.sample_dir, p.sample_dir);
        if (!(l2 > 1e-12f) || !lxm_isfinite(l2)) {
            p.kill = true;
            return false;
        }
        p.d = p.sample_dir;
        if (dot(p.d, n) > 0.0f) p.o = add(hp, muls(n, 0.01f));
        else p.o = sub(hp, muls(n, 0.01f));
    }
    return compact && !p.kill;
}

/* One iteration of pathtrace(), src/pathtrace.cu:97-367, over a list of
 * pixels. */
static void render_iteration(const lxo_scene& s, const RenderCfg& cfg, int iter, const uint32_t* pixels, uint32_t n,
                             lxr* rgb_sum, Counters& cnt, std::vector<std::vector<uint32_t>>* survivors,
                             std::vector<std::vector<lxr>>* records) {
    std::vector<Path> paths(n);
    std::vector<uint32_t> queue(n), next;
    next.reserve(n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t px = pixels ? pixels[i] : i;
        Path& p = paths[i];
        int x = (int)(px % (uint32_t)cfg.W), y = (int)(px / (uint32_t)cfg.W);
        p.o = v3(cfg.cam.position[0], cfg.cam.position[1], cfg.cam.position[2]);
        camera_ray(cfg.cam, iter, x, y, cfg.depth, cfg.seed, p.d);
        p.color = v3(0, 0, 0);
        p.thr = v3(1, 1, 1);
        p.kill = false;
        p.spec = false;
        p.last_pdf = -1.0f;
        p.pixel = px;
        queue[i] = i;
    }
    std::vector<BounceRecords> recs(n);
    int depth = 0;
    TraceCtx tc;
    tc.iteration = (uint32_t)iter;
    while (true) {
        /* sampleDirectLight, pre-increment depth */
        if (cfg.dmis || cfg.emis) {
            for (uint32_t q : queue) {
                Path& p = paths[q];
                Rng rng = make_rng(iter, (int)p.pixel, depth, cfg.seed);
                if (cfg.dmis) p.dl_sample = sample_light_point(s, rng);
                if (cfg.emis) p.env_sample = sample_env_dir(s, rng);
            }
        }
        /* __raygen__rg, optix_triangle.cu:52-143 */
        for (uint32_t q : queue) {
            Path& p = paths[q];
            BounceRecords& br = recs[q];
            tc.key = p.pixel;
            tc.alpha = 1;
            tc.any = false;
            Hit h = trace_scene(s, p.o, p.d, 1e16f, cfg.mode, tc);
            cnt.rad++;
            if (h.inst != 0xffffffffu) {
                br.isect = closest_hit_attrs(s, h);
                V3 hp = add(p.o, smul(br.isect.t, p.d));
                V3 so = add(hp, muls(br.isect.N, 1e-4f));
                if (cfg.dmis) {
                    V3 ld = sub(p.dl_sample, so);
                    lxr dist = length(ld);
                    ld = muls(ld, 1.0f / dist);
                    Hit hd = trace_scene(s, so, ld, dist * 1.0001f + 1e-4f, cfg.mode, tc);
                    cnt.dl++;
                    if (hd.inst != 0xffffffffu) br.dl = dl_hit_attrs(s, hd);
                    else br.dl.t = -1.0f;
                } else {
                    br.dl.t = -1.0f;
                }
                if (cfg.emis) {
                    TraceCtx ta = tc;
                    ta.any = true;
                    Hit he = trace_scene(s, so, p.env_sample, 1e16f, cfg.mode, ta);
                    cnt.env++;
                    br.env_t = he.inst != 0xffffffffu ? 1.0f : -1.0f;
                } else {
                    br.env_t = 1.0f;
                }
            } else {
                br.isect.t = -1.0f;
                br.dl.t = -1.0f;
                br.env_t = -0.5f;
            }
        }
        depth++;
        bool compact = depth < cfg.depth;
        next.clear();
        for (uint32_t q : queue) {
            cnt.bounces++;
            if (records) {
                const Rec& r = recs[q].isect;
                std::vector<lxr>& rv = (*records)[depth - 1];
                lxr t = r.t > 0.0f ? r.t : -1.0f;
                rv.push_back(t);
                rv.push_back(r.t > 0.0f ? r.N.x : 0.0f);
                rv.push_back(r.t > 0.0f ? r.N.y : 0.0f);
                rv.push_back(r.t > 0.0f ? r.N.z : 0.0f);
                rv.push_back(r.t > 0.0f ? r.uv.x : 0.0f);
                rv.push_back(r.t > 0.0f ? r.uv.y : 0.0f);
                rv.push_back(r.t > 0.0f ? (lxr)r.mat : -1.0f);
                rv.push_back((lxr)paths[q].pixel);
            }
            if (shade_one(s, cfg, paths[q], recs[q], iter, depth, compact)) next.push_back(q);
        }
        if (survivors && compact) {
            std::vector<uint32_t>& sv = (*survivors)[depth - 1];
            for (uint32_t q : next) sv.push_back(paths[q].pixel);
        }
        if (compact) queue.swap(next);
        if (depth == cfg.depth || queue.empty()) break;
    }
    /* finalGather, src/kernels/image_kernels.cu:49-60 */
    for (uint32_t i = 0; i < n; i++) {
        rgb_sum[3 * i + 0] += paths[i].color.x;
        rgb_sum[3 * i + 1] += paths[i].color.y;
        rgb_sum[3 * i + 2] += paths[i].color.z;
    }
}

static RenderCfg make_cfg(const lxo_scene& s, const lxo_render_opts* o) {
    RenderCfg c;
    c.cam = o->camera;
    c.W = o->width > 0 ? o->width : c.cam.width;
    c.H = o->height > 0 ? o->height : c.cam.height;
    c.cam.width = c.W;
    c.cam.height = c.H;
    c.depth = o->max_depth > 0 ? o->max_depth : s.trace_depth;
    c.debug_mode = o->material_debug_mode;
    if (c.debug_mode != 0) c.depth = 1; /* pathtrace.cu:107 */
    c.legacy = o->shading_model == LX_SHADING_LEGACY;
    c.dmis = o->dmis != 0;
    c.emis = o->emis != 0;
    c.intensity = o->envmap_intensity;
    c.seed = o->seed;
    c.mode = o->traverse_mode;
    return c;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

/* ===================================================================== C API */
extern "C" {

lxo_scene* lxo_scene_create(const lx_scene_desc* d, int cdf_mode, int build_bvh) {
    lxo_scene* s = new lxo_scene();
    s->meshes.resize(d->num_meshes);
    for (uint32_t i = 0; i < d->num_meshes; i++) {
        const lx_mesh_desc& md = d->meshes[i];
        Mesh& m = s->meshes[i];
        m.pos.resize(md.num_verts);
        for (uint32_t v = 0; v < md.num_verts; v++)
            m.pos[v] = v3(md.positions[3 * v], md.positions[3 * v + 1], md.positions[3 * v + 2]);
        m.has_n = md.normals != nullptr;
        if (m.has_n) {
            m.nrm.resize(md.num_verts);
            for (uint32_t v = 0; v < md.num_verts; v++)
                m.nrm[v] = v3(md.normals[3 * v], md.normals[3 * v + 1], md.normals[3 * v + 2]);
        }
        m.has_uv = md.uvs != nullptr;
        if (m.has_uv) {
            m.uv.resize(md.num_verts);
            for (uint32_t v = 0; v < md.num_verts; v++) m.uv[v] = V2{md.uvs[2 * v], md.uvs[2 * v + 1]};
        }
        m.idx.assign(md.indices, md.indices + 3 * (size_t)md.num_tris);
    }
    double t0 = now_ms();
    if (build_bvh) {
        /* one tree per mesh (BVH::make_bvh per Geom, bvh.cu:18-38); the meshes
         * are independent, so they are built on up to 16 threads */
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k = next++; k < s->meshes.size(); k = next++) {
                Mesh& m = s->meshes[k];
                int nt = (int)(m.idx.size() / 3);
                if (nt == 0) continue;
                m.bvh.resize(2 * nt - 1);
                std::vector<int> ti(nt);
                std::iota(ti.begin(), ti.end(), 0);
                int alloc = 1, md = 0;
                fill_bvh(0, 0, nt - 1, m.bvh, m, ti, alloc, 0, md);
                m.bvh_depth = md;
            }
        };
        unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        nthr = (unsigned)std::min<size_t>(nthr, std::max<size_t>(1, s->meshes.size()));
        std::vector<std::thread> builders;
        for (unsigned i = 1; i < nthr; i++) builders.emplace_back(work);
        work();
        for (auto& t : builders) t.join();
    }
    s->build_ms = now_ms() - t0;
    s->mats.resize(d->num_materials);
    for (uint32_t i = 0; i < d->num_materials; i++) {
        const lx_material_desc& a = d->materials[i];
        Mat& m = s->mats[i];
        m.color = v3(a.base_color[0], a.base_color[1], a.base_color[2]);
        m.albedo_tex = a.albedo_tex;
        m.albedo_xf = a.albedo_xf;
        m.roughness = a.roughness;
        m.metallic = a.metallic;
        m.mr_tex = a.metallic_roughness_tex;
        m.mr_xf = a.metallic_roughness_xf;
        m.normal_tex = a.normal_tex;
        m.normal_xf = a.normal_xf;
        m.emission_color = v3(a.emission_color[0], a.emission_color[1], a.emission_color[2]);
        m.emission_strength = a.emission_strength;
        m.emissive_tex = a.emissive_tex;
        m.emissive_xf = a.emissive_xf;
        m.ior = a.ior;
        m.transmission = a.transmission;
        m.alpha = a.alpha;
        m.alpha_mode = a.alpha_mode;
        m.alpha_cutoff = a.alpha_cutoff;
        m.type = a.material_type;
    }
    s->texs.resize(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; i++) decode_texture(d->textures[i], s->texs[i]);
    s->insts.resize(d->num_instances);
    for (uint32_t i = 0; i < d->num_instances; i++) {
        Inst& g = s->insts[i];
        g.mesh = d->instances[i].mesh;
        g.mat = d->instances[i].material;
        for (int k = 0; k < 12; k++) g.m[k] = d->instances[i].transform[k];
        affine_inverse(g.m, g.inv);
        const Mesh& m = s->meshes[g.mesh];
        BBox b;
        for (const V3& p : m.pos) {
            /* world box in double then padded: only used to skip instances */
            double w[3];
            for (int r = 0; r < 3; r++)
                w[r] = (double)g.m[r * 4] * p.x + (double)g.m[r * 4 + 1] * p.y + (double)g.m[r * 4 + 2] * p.z +
                       (double)g.m[r * 4 + 3];
            b.expand_point(v3((float)w[0], (float)w[1], (float)w[2]));
        }
        float pad = 1e-4f * (1.0f + lxm_max(lxm_max(lxm_abs(b.mn.x), lxm_abs(b.mx.x)),
                                             lxm_max(lxm_max(lxm_abs(b.mn.y), lxm_abs(b.mx.y)),
                                                     lxm_max(lxm_abs(b.mn.z), lxm_abs(b.mx.z)))));
        b.mn = sub(b.mn, v3(pad, pad, pad));
        b.mx = add(b.mx, v3(pad, pad, pad));
        g.wbox = b;
    }
    s->cam = d->camera;
    s->trace_depth = d->trace_depth > 0 ? d->trace_depth : 8;
    s->has_env = d->envmap != nullptr;
    if (s->has_env) {
        decode_texture(*d->envmap, s->env);
        precompute_env(*s, cdf_mode);
    }
    precompute_emitters(*s, cdf_mode);
    return s;
}

void lxo_scene_destroy(lxo_scene* s) { delete s; }
double lxo_scene_bvh_build_ms(const lxo_scene* s) { return s->build_ms; }
uint32_t lxo_scene_bvh_nodes(const lxo_scene* s, uint32_t mesh) { return (uint32_t)s->meshes[mesh].bvh.size(); }
int lxo_scene_bvh_depth(const lxo_scene* s, uint32_t mesh) { return s->meshes[mesh].bvh_depth; }

int lxo_read_env_cdfs(const lxo_scene* s, float* marginal, float* conditional) {
    if (!s->has_env) return -1;
    if (marginal) std::copy(s->marginal.begin(), s->marginal.end(), marginal);
    if (conditional) std::copy(s->conditional.begin(), s->conditional.end(), conditional);
    return 0;
}

int lxo_read_emitter_cdfs(const lxo_scene* s, uint32_t* inst, float* prefix, uint32_t* num, float* total) {
    if (num) *num = (uint32_t)s->emissive.size();
    if (inst) std::memcpy(inst, s->emissive.data(), s->emissive.size() * 4);
    if (prefix) std::copy(s->emissive_prefix.begin(), s->emissive_prefix.end(), prefix);
    if (total) *total = s->total_area;
    return 0;
}

int lxo_read_instance_tri_cdf(const lxo_scene* s, uint32_t k, float* tri_prefix, uint32_t* n) {
    if (k >= s->tri_cdf.size()) return -1;
    if (n) *n = (uint32_t)s->tri_cdf[k].size();
    if (tri_prefix) std::copy(s->tri_cdf[k].begin(), s->tri_cdf[k].end(), tri_prefix);
    return 0;
}

int lxo_trace(const lxo_scene* s, const lx_ray* rays, uint32_t n, int mode, int alpha, uint32_t iteration,
              const uint32_t* pixel, lx_hit* hits, uint8_t* occluded) {
    for (uint32_t i = 0; i < n; i++) {
        TraceCtx tc;
        tc.alpha = alpha;
        tc.iteration = iteration;
        tc.key = pixel ? pixel[i] : i;
        V3 o = v3(rays[i].origin[0], rays[i].origin[1], rays[i].origin[2]);
        V3 d = v3(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]);
        if (hits) {
            tc.any = false;
            Hit h = trace_scene(*s, o, d, rays[i].tmax, mode, tc);
            hits[i].t = h.inst == 0xffffffffu ? -1.0f : h.t;
            hits[i].instance = h.inst;
            hits[i].prim = h.prim;
            hits[i].u = h.u;
            hits[i].v = h.v;
        }
        if (occluded) {
            tc.any = true;
            Hit h = trace_scene(*s, o, d, rays[i].tmax, mode, tc);
            occluded[i] = h.inst != 0xffffffffu ? 1 : 0;
        }
    }
    return 0;
}

int lxo_render(const lxo_scene* s, const lxo_render_opts* o, uint32_t first_iter, uint32_t n_iters,
               const uint32_t* pixels, uint32_t n_pixels, float* rgb_sum, lxo_stats* stats) {
    RenderCfg cfg = make_cfg(*s, o);
    uint32_t n = pixels ? n_pixels : (uint32_t)(cfg.W * cfg.H);
    std::vector<lxr> acc(3 * (size_t)n, (lxr)0);
    int threads = o->threads > 0 ? o->threads : 1;
    if ((uint32_t)threads > n) threads = n > 0 ? (int)n : 1;
    std::vector<uint32_t> all;
    if (!pixels) {
        all.resize(n);
        std::iota(all.begin(), all.end(), 0u);
        pixels = all.data();
    }
    std::vector<Counters> cnt(threads);
    double t0 = now_ms();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            uint32_t b = (uint32_t)((uint64_t)n * t / threads), e = (uint32_t)((uint64_t)n * (t + 1) / threads);
            for (uint32_t it = 0; it < n_iters; it++)
                render_iteration(*s, cfg, (int)(first_iter + it), pixels + b, e - b, acc.data() + 3 * (size_t)b,
                                 cnt[t], nullptr, nullptr);
        });
    }
    for (auto& th : pool) th.join();
    std::copy(acc.begin(), acc.end(), rgb_sum);
    double t1 = now_ms();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (auto& c : cnt) {
            stats->radiance_rays += c.rad;
            stats->dl_shadow_rays += c.dl;
            stats->env_shadow_rays += c.env;
            stats->path_bounces += c.bounces;
        }
        stats->wall_ms = t1 - t0;
    }
    return 0;
}

#ifndef LXO_FP64
/* Survivor sets, hit records and unit functions: the fp32 oracle's checks
 * only (the FP64 build exports scenes, tracing and rendering). */
int lxo_debug_bounce(const lxo_scene* s, const lxo_render_opts* o, uint32_t iter, const uint32_t* pixels,
                     uint32_t n_pixels, uint32_t* survivors, uint32_t* survivor_counts, float* records,
                     uint32_t* record_counts, float* rgb_sum) {
    RenderCfg cfg = make_cfg(*s, o);
    uint32_t n = pixels ? n_pixels : (uint32_t)(cfg.W * cfg.H);
    std::vector<std::vector<uint32_t>> sv(cfg.depth);
    std::vector<std::vector<float>> rv(cfg.depth);
    Counters c;
    std::vector<float> sum(3 * (size_t)n, 0.0f);
    render_iteration(*s, cfg, (int)iter, pixels, n, sum.data(), c, &sv, &rv);
    for (int d = 0; d < cfg.depth; d++) {
        std::sort(sv[d].begin(), sv[d].end());
        survivor_counts[d] = (uint32_t)sv[d].size();
        if (survivors) std::memcpy(survivors + (size_t)d * n, sv[d].data(), sv[d].size() * 4);
        size_t nr = rv[d].size() / 8;
        std::vector<size_t> order(nr);
        std::iota(order.begin(), order.end(), (size_t)0);
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return rv[d][8 * a + 7] < rv[d][8 * b + 7]; });
        record_counts[d] = (uint32_t)nr;
        if (records)
            for (size_t k = 0; k < nr; k++)
                std::memcpy(records + ((size_t)d * n + k) * 8, &rv[d][8 * order[k]], 32);
    }
    if (rgb_sum) std::memcpy(rgb_sum, sum.data(), sum.size() * 4);
    return 0;
}

uint32_t lxo_utilhash(uint32_t a) { return utilhash(a); }

void lxo_rng_draws(int iter, int index, int depth, uint32_t seed, int n, float* out, uint32_t* states) {
    Rng r = make_rng(iter, index, depth, seed);
    if (states) states[0] = r.x;
    for (int i = 0; i < n; i++) {
        out[i] = r.u01();
        if (states) states[i + 1] = r.x;
    }
}

void lxo_camera_ray(const lx_camera_desc* cam, int iter, int x, int y, int trace_depth, uint32_t seed, float* dir3) {
    V3 d;
    camera_ray(*cam, iter, x, y, trace_depth, seed, d);
    dir3[0] = d.x;
    dir3[1] = d.y;
    dir3[2] = d.z;
}

static V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

float lxo_ct_pdf(const float* wo, const float* wi, const float* n, float r, float m, const float* c, float ior,
                 float tr) {
    return PDF(ld3(wo), ld3(wi), ld3(n), r, m, ld3(c), ior, tr);
}
void lxo_ct_brdf(const float* v, const float* n, const float* l, const float* a, float r, float m, float ior, float tr,
                 float* out3) {
    V3 f = BRDF(ld3(v), ld3(n), ld3(l), ld3(a), r, m, ior, tr);
    out3[0] = f.x;
    out3[1] = f.y;
    out3[2] = f.z;
}
void lxo_ct_btdf(const float* wo, const float* n, const float* wi, const float* a, float r, float m, float ior,
                 float tr, float* out3) {
    V3 f = BTDF(ld3(wo), ld3(n), ld3(wi), ld3(a), r, m, lxm_max(ior, 1.0001f), tr);
    out3[0] = f.x;
    out3[1] = f.y;
    out3[2] = f.z;
}
float lxo_ct_fresnel_dielectric(float c, float ei, float et) { return F_diel(c, ei, et); }
float lxo_ct_d(const float* h, const float* n, float alpha) { return D_TR(ld3(h), ld3(n), alpha); }
float lxo_ct_smith_g1(const float* w, const float* n, float alpha) { return G1(ld3(w), ld3(n), alpha); }

float lxo_glass_fresnel(float cos_i, float ior) { return glass_fresnel(cos_i, ior); }
float lxo_lambert_pdf(const float* wi, const float* n) { return lambert_pdf(ld3(wi), ld3(n)); }

int lxo_legacy_sample(int type, const float* dir, const float* n, float ior, int iter, int index, int depth,
                      uint32_t seed, float* dir3) {
    Rng rng = make_rng(iter, index, depth, seed);
    Rng probe = rng;
    V3 d = legacy_sample(type, ld3(dir), ld3(n), ior, rng);
    dir3[0] = d.x;
    dir3[1] = d.y;
    dir3[2] = d.z;
    int k = 0;
    while (probe.x != rng.x && k < 16) {
        probe.u01();
        k++;
    }
    return k;
}

float lxo_legacy_update(int type, const float* dir, const float* n, const float* wi, const float* color, float ior,
                        float transmission, float* thr3) {
    V3 thr = ld3(thr3);
    float pdf = legacy_update(type, ld3(dir), ld3(n), ld3(wi), ld3(color), ior, transmission, thr, -1.0f);
    thr3[0] = thr.x;
    thr3[1] = thr.y;
    thr3[2] = thr.z;
    return pdf;
}

int lxo_ct_sample(const float* wo, const float* n, float r, float m, const float* c, float ior, float tr, int iter,
                  int index, int depth, uint32_t seed, float* dir3) {
    Rng rng = make_rng(iter, index, depth, seed);
    Rng probe = rng;
    V3 d = sample_ct(ld3(wo), ld3(n), r, m, rng, ld3(c), ior, tr);
    dir3[0] = d.x;
    dir3[1] = d.y;
    dir3[2] = d.z;
    int k = 0;
    while (probe.x != rng.x && k < 16) {
        probe.u01();
        k++;
    }
    return k;
}

int lxo_select_from_cdf(const float* cdf, int size, float xi) { return select_from_cdf(cdf, size, xi); }

void lxo_texture_sample(const lx_texture_desc* tex, float u, float v, float* out4) {
    Tex t;
    decode_texture(*tex, t);
    tex_sample(t, u, v, out4);
}

float lxo_envmap_pdf(const lxo_scene* s, const float* d) { return envmap_pdf(*s, ld3(d)); }

void lxo_sample_env(const lxo_scene* s, int iter, int index, int depth, uint32_t seed, int dmis, float* dir3) {
    Rng rng = make_rng(iter, index, depth, seed);
    if (dmis) sample_light_point(*s, rng);
    V3 d = sample_env_dir(*s, rng);
    dir3[0] = d.x;
    dir3[1] = d.y;
    dir3[2] = d.z;
}

void lxo_sample_light(const lxo_scene* s, int iter, int index, int depth, uint32_t seed, float* p3) {
    Rng rng = make_rng(iter, index, depth, seed);
    V3 p = sample_light_point(*s, rng);
    p3[0] = p.x;
    p3[1] = p.y;
    p3[2] = p.z;
}

/* Batched forms of the building blocks above, for the Monte-Carlo and
 * quadrature pins (one call per array instead of one per sample). */
void lxo_ct_eval_batch(int what, const float* wo, const float* wi, uint32_t n_dirs, const float* n, float r, float m,
                       const float* c, float ior, float tr, float* out) {
    for (uint32_t i = 0; i < n_dirs; i++) {
        const float* w = wi + 3 * (size_t)i;
        if (what == 0) {
            out[i] = PDF(ld3(wo), ld3(w), ld3(n), r, m, ld3(c), ior, tr);
        } else {
            V3 f = what == 1 ? BTDF(ld3(wo), ld3(n), ld3(w), ld3(c), r, m, lxm_max(ior, 1.0001f), tr)
                             : BRDF(ld3(wo), ld3(n), ld3(w), ld3(c), r, m, ior, tr);
            out[3 * (size_t)i + 0] = f.x;
            out[3 * (size_t)i + 1] = f.y;
            out[3 * (size_t)i + 2] = f.z;
        }
    }
}
void lxo_ct_sample_batch(const float* wo, const float* n, float r, float m, const float* c, float ior, float tr,
                         int iter, int depth, uint32_t seed, int index0, uint32_t count, float* dirs) {
    for (uint32_t i = 0; i < count; i++) {
        Rng rng = make_rng(iter, index0 + (int)i, depth, seed);
        V3 d = sample_ct(ld3(wo), ld3(n), r, m, rng, ld3(c), ior, tr);
        dirs[3 * (size_t)i + 0] = d.x;
        dirs[3 * (size_t)i + 1] = d.y;
        dirs[3 * (size_t)i + 2] = d.z;
    }
}
void lxo_envmap_pdf_batch(const lxo_scene* s, const float* dirs, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; i++) out[i] = envmap_pdf(*s, ld3(dirs + 3 * (size_t)i));
}
void lxo_sample_env_batch(const lxo_scene* s, int iter, int depth, uint32_t seed, int index0, uint32_t count,
                          float* dirs) {
    for (uint32_t i = 0; i < count; i++) {
        Rng rng = make_rng(iter, index0 + (int)i, depth, seed);
        V3 d = sample_env_dir(*s, rng);
        dirs[3 * (size_t)i + 0] = d.x;
        dirs[3 * (size_t)i + 1] = d.y;
        dirs[3 * (size_t)i + 2] = d.z;
    }
}
void lxo_sample_light_batch(const lxo_scene* s, int iter, int depth, uint32_t seed, int index0, uint32_t count,
                            float* pts) {
    for (uint32_t i = 0; i < count; i++) {
        Rng rng = make_rng(iter, index0 + (int)i, depth, seed);
        V3 p = sample_light_point(*s, rng);
        pts[3 * (size_t)i + 0] = p.x;
        pts[3 * (size_t)i + 1] = p.y;
        pts[3 * (size_t)i + 2] = p.z;
    }
}
/* get_normal of material `mat` of the scene for a hit with geometric /
 * interpolated normal N, tangent T and uv */
void lxo_get_normal(const lxo_scene* s, uint32_t mat, const float* N, const float* T, const float* uv, float* out3) {
    Rec r;
    r.t = 1.0f;
    r.N = ld3(N);
    r.T = ld3(T);
    r.uv = v2(uv[0], uv[1]);
    r.mat = (int)mat;
    V3 o = get_normal(*s, s->mats[mat], r);
    out3[0] = o.x;
    out3[1] = o.y;
    out3[2] = o.z;
}

void lxo_math_eval(int fn, const float* x, const float* y, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
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
}

#endif /* LXO_FP64 */

} /* extern "C" */
