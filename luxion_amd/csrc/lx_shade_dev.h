/*
 * lx_shade_dev.h -- gfx950 device library of the path tracer: RNG, texture
 * sampling, material queries, Cook-Torrance uber-shader, light / environment
 * sampling.  Pure FP32 VALU work.
 *
 * Floating-point expressions keep the operand order and association of the
 * reference renderer's glm / sutil code (cited per function, paths relative
 * to the reference repository root); the build compiles with contraction
 * off and correctly rounded fp32 divide / sqrt, so each value is bitwise
 * reproducible on a host CPU.
 */
#ifndef LX_SHADE_DEV_H
#define LX_SHADE_DEV_H

#include "lx_types.h"
#include "../../include/luxion_detmath.h"

#define LXD __device__ __forceinline__

/* Explicit address spaces: scene data read through pointers cast to the
 * global address space compiles to global (not flat) memory instructions,
 * which neither wait on the LDS counter nor carry 64-bit addresses per lane;
 * the traversal stack column is typed LDS. */
#define LX_GLOBAL __attribute__((address_space(1)))
#define LX_LDS __attribute__((address_space(3)))
template <class T>
LXD const LX_GLOBAL T* lx_gptr(const T* p) {
    return (const LX_GLOBAL T*)p;
}
/* clang vector types: HIP's float4 / int4 cannot be copied from an
 * address-space-qualified reference */
typedef float lx_v4f __attribute__((ext_vector_type(4)));
typedef int lx_v4i __attribute__((ext_vector_type(4)));
LXD float4 lx_ld4(const LX_GLOBAL lx_v4f* p) {
    lx_v4f v = *p;
    return make_float4(v.x, v.y, v.z, v.w);
}

struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};
LXD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
LXD f2 mk2(float x, float y) { return f2{x, y}; }
LXD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
LXD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
LXD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
LXD f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
LXD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
LXD f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
LXD f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); } /* glm: per component */
LXD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LXD f3 cross(f3 x, f3 y) { return mk3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y); }
LXD f3 normalize(f3 a) { return a * (1.0f / lxm_sqrt(dot(a, a))); }
LXD float length(f3 a) { return lxm_sqrt(dot(a, a)); }

/* 3x4 row-major affine maps (glm mat4 * vec4 pairwise order). */
LXD f3 xf_point(const float* m, f3 p) {
    return mk3((m[0] * p.x + m[1] * p.y) + (m[2] * p.z + m[3]), (m[4] * p.x + m[5] * p.y) + (m[6] * p.z + m[7]),
               (m[8] * p.x + m[9] * p.y) + (m[10] * p.z + m[11]));
}
LXD f3 xf_vector(const float* m, f3 v) {
    return mk3((m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[4] * v.x + m[5] * v.y) + m[6] * v.z,
               (m[8] * v.x + m[9] * v.y) + m[10] * v.z);
}
LXD f3 xf_normal(const float* inv, f3 n) {
    return mk3((inv[0] * n.x + inv[4] * n.y) + inv[8] * n.z, (inv[1] * n.x + inv[5] * n.y) + inv[9] * n.z,
               (inv[2] * n.x + inv[6] * n.y) + inv[10] * n.z);
}

/* ---------------------------------------------------------------- RNG */
/* utilhash (include/intersections.h:13-22) + minstd_rand seeded as in
 * include/common.h:38-43; u01 per thrust uniform_real_distribution<float>. */
LXD uint32_t lx_utilhash(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

struct LxRng {
    uint32_t x;
    LXD float u01() {
        /* 48271 * x < 2^47: exact in 64 bits; mod by 2^31 - 1 via folding */
        uint64_t p = (uint64_t)x * 48271u;
        uint32_t r = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
        if (r >= 2147483647u) r -= 2147483647u;
        x = r;
        float f = (float)(r - 1u);
        f = f / (1.0f + (float)(2147483646u - 1u));
        return f * (1.0f - 0.0f) + 0.0f;
    }
};

LXD LxRng lx_make_rng(int iter, int index, int depth, uint32_t seed) {
    uint32_t h = lx_utilhash(0x80000000u | ((uint32_t)depth << 22) | (uint32_t)iter) ^ lx_utilhash((uint32_t)index);
    if (seed != 0u) h ^= lx_utilhash(seed * 0x9E3779B9u + 0x7F4A7C15u);
    LxRng r;
    r.x = h % 2147483647u;
    if (r.x == 0u) r.x = 1u;
    return r;
}

/* -------------------------------------------------------------- textures */
LXD void lx_texel(const LxTexDesc& t, int i, int j, float* o) {
    size_t k = ((size_t)j * (size_t)t.w + (size_t)i);
    if (t.fmt == LX_TEX_RGBA8) {
        uint32_t p = lx_gptr((const uint32_t*)t.data)[k];
        o[0] = ((float)(p & 0xffu)) / 255.0f;
        o[1] = ((float)((p >> 8) & 0xffu)) / 255.0f;
        o[2] = ((float)((p >> 16) & 0xffu)) / 255.0f;
        o[3] = ((float)(p >> 24)) / 255.0f;
    } else if (t.fmt == LX_TEX_RGBA16) {
        const LX_GLOBAL uint16_t* s = lx_gptr((const uint16_t*)t.data) + 4 * k;
        o[0] = ((float)s[0]) / 65535.0f;
        o[1] = ((float)s[1]) / 65535.0f;
        o[2] = ((float)s[2]) / 65535.0f;
        o[3] = ((float)s[3]) / 65535.0f;
    } else {
        float4 v = lx_ld4((const LX_GLOBAL lx_v4f*)t.data + k);
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
}

/* Exact-weight bilinear filter with wrap addressing on normalised
 * coordinates: the documented semantics of the reference's tex2D<float4>
 * objects (src/texture.cu:52-56; cudaFilterModeLinear, cudaAddressModeWrap). */
LXD void lx_tex_sample(const LxTexDesc& t, float u, float v, float* out) {
    if (!lxm_isfinite(u)) u = 0.0f;
    if (!lxm_isfinite(v)) v = 0.0f;
    u = u - lxm_floor(u);
    v = v - lxm_floor(v);
    float x = u * (float)t.w - 0.5f;
    float y = v * (float)t.h - 0.5f;
    float fx = lxm_floor(x), fy = lxm_floor(y);
    float ax = x - fx, ay = y - fy;
    int i0 = (int)fx, j0 = (int)fy;
    if (i0 < 0) i0 += t.w;
    if (i0 >= t.w) i0 -= t.w;
    if (j0 < 0) j0 += t.h;
    if (j0 >= t.h) j0 -= t.h;
    int i1 = i0 + 1 == t.w ? 0 : i0 + 1;
    int j1 = j0 + 1 == t.h ? 0 : j0 + 1;
    float t00[4], t10[4], t01[4], t11[4];
    lx_texel(t, i0, j0, t00);
    lx_texel(t, i1, j0, t10);
    lx_texel(t, i0, j1, t01);
    lx_texel(t, i1, j1, t11);
    /* nested lerps: exact for constant textures and at texel centres */
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float a = t00[c] + ax * (t10[c] - t00[c]);
        float b = t01[c] + ax * (t11[c] - t01[c]);
        out[c] = a + ay * (b - a);
    }
}

/* ------------------------------------------------------------ materials */
/* apply_texture_transform, src/material_queries.cu:3-17 */
LXD f2 lx_apply_xf(f2 uv, const lx_tex_transform& x) {
    uv = mk2(uv.x * x.scale[0], uv.y * x.scale[1]);
    if (x.rotation != 0.0f) {
        float c = lxm_cos(x.rotation), s = lxm_sin(x.rotation);
        uv = mk2(c * uv.x - s * uv.y, s * uv.x + c * uv.y);
    }
    return mk2(uv.x + x.offset[0], uv.y + x.offset[1]);
}

/* get_albedo, material_queries.cu:19-30 */
LXD f3 lx_get_albedo(const LxDevScene& S, const lx_material_desc& m, f2 uv) {
    if (m.albedo_tex >= 0) {
        uv = lx_apply_xf(uv, m.albedo_xf);
        float t[4];
        lx_tex_sample(S.texs[m.albedo_tex], uv.x, uv.y, t);
        return mk3(lxm_pow(t[0], 2.2f), lxm_pow(t[1], 2.2f), lxm_pow(t[2], 2.2f));
    }
    return mk3(m.base_color[0], m.base_color[1], m.base_color[2]);
}

/* get_normal, material_queries.cu:33-89 */
LXD f3 lx_get_normal(const LxDevScene& S, const lx_material_desc& m, f3 N, f3 Tg, f2 uv0) {
    if (m.normal_tex < 0) return N;
    f2 uv = lx_apply_xf(uv0, m.normal_xf);
    float t[4];
    lx_tex_sample(S.texs[m.normal_tex], uv.x, uv.y, t);
    f3 ln = mk3(t[0] * 2.0f - 1.0f, t[1] * 2.0f - 1.0f, t[2] * 2.0f - 1.0f);
    float l2 = dot(ln, ln);
    if (!(l2 > 1e-12f)) return N;
    f3 T = Tg - dot(Tg, N) * N;
    float t2 = dot(T, T);
    if (!(t2 > 1e-12f)) return N;
    T = T * (1.0f / lxm_sqrt(t2));
    f3 B = cross(N, T);
    f3 mp = mk3(T.x * ln.x + B.x * ln.y + N.x * ln.z, T.y * ln.x + B.y * ln.y + N.y * ln.z,
                T.z * ln.x + B.z * ln.y + N.z * ln.z);
    float m2 = dot(mp, mp);
    if (!(m2 > 1e-12f)) return N;
    mp = mp * (1.0f / lxm_sqrt(m2));
    if (!(dot(mp, N) > 1e-4f)) return N;
    return mp;
}

/* get_metallic_roughness, material_queries.cu:92-106 */
LXD void lx_get_mr(const LxDevScene& S, const lx_material_desc& m, f2 uv, float& rough, float& metal) {
    rough = m.roughness;
    metal = m.metallic;
    if (m.metallic_roughness_tex >= 0) {
        uv = lx_apply_xf(uv, m.metallic_roughness_xf);
        float t[4];
        lx_tex_sample(S.texs[m.metallic_roughness_tex], uv.x, uv.y, t);
        rough *= t[1];
        metal *= t[2];
    }
}

/* get_emission, material_queries.cu:108-120 */
LXD f3 lx_get_emission(const LxDevScene& S, const lx_material_desc& m, f2 uv) {
    f3 e = mk3(m.emission_color[0], m.emission_color[1], m.emission_color[2]) * m.emission_strength;
    if (m.emissive_tex >= 0) {
        uv = lx_apply_xf(uv, m.emissive_xf);
        float t[4];
        lx_tex_sample(S.texs[m.emissive_tex], uv.x, uv.y, t);
        e = e * mk3(lxm_pow(t[0], 2.2f), lxm_pow(t[1], 2.2f), lxm_pow(t[2], 2.2f));
    }
    return e;
}

/* -------------------------------------------------------- Cook-Torrance */
#define LX_CT_EPS 1e-30f
#define LX_PDF_CLAMP 1e-5f
#define LX_INV_PI_CT 0.3183098f
#define LX_GLM_PI ((float)3.14159265358979323846264338327950288)
#define LX_PI_F 3.1415926535897932384626422832795028841971f
#define LX_TWO_PI_F 6.2831853071795864769252867665590057683943f
#define LX_SQRT_THIRD 0.5773502691896257645091487805019574556476f

/* D_TrowbridgeReitz, src/shaders/cook_torrance.cu:8-25 */
LXD float lx_D(f3 h, f3 n, float alpha) {
    float ndh = dot(n, h);
    if (ndh <= 0.0f) return 0.0f;
    float a2 = alpha * alpha;
    float ndh2 = ndh * ndh;
    float dc = (ndh2 * a2) + (1.0f - ndh2);
    float den = dc * dc * LX_GLM_PI;
    return a2 / lxm_max(den, LX_CT_EPS);
}
/* F_SchlickApprox, cook_torrance.cu:27-31 */
LXD f3 lx_schlick(float vdh, f3 f0) {
    float m = 1.0f - vdh;
    float m2 = m * m;
    return f0 + ((mk3(1.0f, 1.0f, 1.0f) - f0) * (m2 * m2 * m));
}
/* FresnelDielectric, cook_torrance.cu:34-50 */
LXD float lx_fdiel(float ci, float ei, float et) {
    ci = lxm_clamp(ci, 0.0f, 1.0f);
    float si = lxm_sqrt(lxm_max(0.0f, 1.0f - ci * ci));
    float st = (ei / et) * si;
    if (st >= 1.0f) return 1.0f;
    float ct = lxm_sqrt(lxm_max(0.0f, 1.0f - st * st));
    float rpa = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
    float rpe = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
    return (rpa * rpa + rpe * rpe) / 2.0f;
}
/* Smith_GGX, cook_torrance.cu:56-66 */
LXD float lx_G1(f3 w, f3 n, float alpha) {
    float a2 = alpha * alpha;
    float ndw = lxm_max(dot(n, w), 0.0f);
    float num = 2.0f * ndw;
    float under = a2 + (1.0f - a2) * (ndw * ndw);
    float den = ndw + lxm_sqrt(under);
    return num / lxm_max(den, LX_CT_EPS);
}
/* Fresnel mix, cook_torrance.cu:72-77 */
LXD f3 lx_fresnel(float c, f3 albedo, float metallic, float ei, float et) {
    float fd = lx_fdiel(c, ei, et);
    f3 d = mk3(fd, fd, fd);
    f3 cd = lx_schlick(c, albedo);
    return d + metallic * (cd - d);
}
/* The view-side terms of the BRDF for one shading point (v, n, rough, ior):
 * the same for every light direction the bounce evaluates, so a bounce
 * computes them once. */
struct LxBrdfView {
    f3 v, nf;
    float ei, et, alpha, g1v, den4v;
};
LXD LxBrdfView lx_brdf_view(f3 v, f3 n, float rough, float ior) {
    LxBrdfView b;
    ior = lxm_max(ior, 1.0001f);
    bool enter = dot(v, n) > 0.0f;
    b.v = v;
    b.nf = enter ? n : -n;
    b.ei = enter ? 1.0f : ior;
    b.et = enter ? ior : 1.0f;
    rough = lxm_clamp(rough, 0.0001f, 1.0f);
    b.alpha = rough * rough;
    b.g1v = lx_G1(v, b.nf, b.alpha);
    b.den4v = 4.0f * lxm_max(dot(b.nf, v), 0.0f);
    return b;
}
/* BRDF, cook_torrance.cu:79-110, toward l */
LXD f3 lx_brdf_v(const LxBrdfView& b, f3 l, f3 albedo, float metal, float tr) {
    f3 h = normalize(b.v + l);
    float D = lx_D(h, b.nf, b.alpha);
    f3 F = lx_fresnel(lxm_max(dot(b.v, h), 0.0f), albedo, metal, b.ei, b.et);
    float G = b.g1v * lx_G1(l, b.nf, b.alpha);
    f3 num = (D * F) * G;
    float den = b.den4v * lxm_max(dot(b.nf, l), 0.0f);
    f3 spec = num / lxm_max(den, LX_CT_EPS);
    f3 ns = mk3(1.0f, 1.0f, 1.0f) - F;
    f3 kD = ((1.0f - metal) * ns) * (1.0f - tr);
    f3 diff = (kD * albedo) * LX_INV_PI_CT;
    return diff + spec;
}
LXD f3 lx_brdf(f3 v, f3 n, f3 l, f3 albedo, float rough, float metal, float ior, float tr) {
    return lx_brdf_v(lx_brdf_view(v, n, rough, ior), l, albedo, metal, tr);
}
/* BTDF, cook_torrance.cu:112-142 */
LXD f3 lx_btdf(f3 wo, f3 n, f3 wi, f3 albedo, float rough, float metal, float ior, float tr) {
    float ci = dot(wo, n);
    bool enter = ci > 0.f;
    float ei = enter ? 1.0f : ior, et = enter ? ior : 1.0f;
    f3 ht = normalize(-((ei * wo) + (et * wi)));
    float jd = (et * dot(wi, ht) + ei * dot(wo, ht));
    float jac = (ei * ei) / lxm_max(jd * jd, LX_CT_EPS);
    float ctn = lxm_abs(dot(wi, ht)) * lxm_abs(dot(wo, ht));
    float ctd = lxm_abs(dot(wi, n)) * lxm_abs(dot(wo, n));
    rough = lxm_clamp(rough, 0.0001f, 1.0f);
    float alpha = rough * rough;
    float D = lx_D(ht, n, alpha);
    float F = lx_fdiel(lxm_abs(dot(wo, ht)), ei, et);
    bool gv = dot(wi, ht) * dot(wi, n) > 0.0f && dot(wo, ht) * dot(wo, n) > 0.0f;
    float G = gv ? lx_G1(dot(wi, n) > 0.0f ? wi : -wi, n, alpha) * lx_G1(dot(wo, n) > 0.0f ? wo : -wo, n, alpha)
                 : 0.0f;
    float DFG = D * G * (1.0f - F);
    return albedo * (tr * (1.0f - metal) * (ctn / ctd) * jac * DFG);
}
struct LxLobes {
    float pR, pD, pS;
};
/* lobe probabilities shared by sampleCookTorrance (cook_torrance.cu:170-183)
 * and PDF (cook_torrance.cu:272-285) */
LXD void lx_lobes(f3 wo, f3 n, f3 color, float metal, float ior, float tr, float& pR, float& pD, float& pS) {
    bool enter = dot(wo, n) > 0.0f;
    float ei = enter ? 1.0f : ior, et = enter ? ior : 1.0f;
    float cm = lxm_abs(dot(wo, n));
    f3 Fa = lx_fresnel(cm, color, metal, ei, et);
    float sr = lxm_clamp((Fa.x + Fa.y + Fa.z) / 3.0f, 0.01f, 0.99f);
    float Fm = lx_fdiel(cm, ei, et);
    pR = tr * (1.0f - metal) * (1.0f - Fm);
    pD = (1.0f - pR) * (1.0f - sr) * (1.0f - metal);
    pS = 1.0f - pR - pD;
}
/* the lobes of a shading point, with PDF's and the sampler's ior clamp */
LXD LxLobes lx_lobes_of(f3 wo, f3 n, f3 color, float metal, float ior, float tr) {
    LxLobes q;
    lx_lobes(wo, n, color, metal, lxm_max(ior, 1.0001f), tr, q.pR, q.pD, q.pS);
    return q;
}
/* calculateRandomDirectionInHemisphere, src/interactions.cu:7-45 */
LXD f3 lx_cos_hemisphere(f3 n, LxRng& rng) {
    float up = lxm_sqrt(rng.u01());
    float over = lxm_sqrt(1.0f - up * up);
    float around = rng.u01() * LX_TWO_PI_F;
    f3 dnn;
    if (lxm_abs(n.x) < LX_SQRT_THIRD) dnn = mk3(1, 0, 0);
    else if (lxm_abs(n.y) < LX_SQRT_THIRD) dnn = mk3(0, 1, 0);
    else dnn = mk3(0, 0, 1);
    f3 p1 = normalize(cross(n, dnn));
    f3 p2 = normalize(cross(n, p1));
    return ((up * n) + ((lxm_cos(around) * over) * p1)) + ((lxm_sin(around) * over) * p2);
}
/* lx_cos_hemisphere with its draws' sqrt(u1) and the sine and cosine of its
 * azimuth u2 * 2pi given */
LXD f3 lx_cos_hemisphere_at(f3 n, float up, float c_around, float s_around) {
    float over = lxm_sqrt(1.0f - up * up);
    f3 dnn;
    if (lxm_abs(n.x) < LX_SQRT_THIRD) dnn = mk3(1, 0, 0);
    else if (lxm_abs(n.y) < LX_SQRT_THIRD) dnn = mk3(0, 1, 0);
    else dnn = mk3(0, 0, 1);
    f3 p1 = normalize(cross(n, dnn));
    f3 p2 = normalize(cross(n, p1));
    return ((up * n) + ((c_around * over) * p1)) + ((s_around * over) * p2);
}
/* sampleCookTorrance, cook_torrance.cu:156-224, with the lobe probabilities
 * given.  Both lobes draw two more numbers in the same order and take
 * sqrt(u1) and the sine and cosine of the same azimuth 2pi * u2 (2.0f * pi_f
 * is the float 2pi bit for bit), so these are computed once, before the
 * lobe branch: a wave whose lanes pick both lobes evaluates them once. */
LXD f3 lx_sample_ct_q(f3 wo, f3 n, float rough, LxRng& rng, float ior, const LxLobes& q) {
    float r = rng.u01();
    ior = lxm_max(ior, 1.0001f);
    rough = lxm_clamp(rough, 0.0001f, 1.0f);
    float alpha = rough * rough;
    bool enter = dot(wo, n) > 0.0f;
    f3 nf = enter ? n : -n;
    float ei = enter ? 1.0f : ior, et = enter ? ior : 1.0f;
    const float pD = q.pD, pS = q.pS;
    float x1 = rng.u01();
    float x2 = rng.u01();
    float sx1 = lxm_sqrt(x1);
    float ph = 2.0f * LX_PI_F * x2;
    float cph = lxm_cos(ph), sph = lxm_sin(ph);
    if (r > pS && r <= pS + pD) return lx_cos_hemisphere_at(nf, sx1, cph, sph);
    float th = lxm_atan((alpha * sx1) / lxm_sqrt(1.0f - x1));
    float st = lxm_sin(th);
    f3 h = normalize(mk3(1.0f * st * cph, 1.0f * st * sph, 1.0f * lxm_cos(th)));
    f3 t = normalize(lxm_abs(nf.z) < 0.999f ? cross(mk3(0, 0, 1), nf) : cross(mk3(1, 0, 0), nf));
    f3 b = cross(nf, t);
    f3 hw = ((h.x * t) + (h.y * b)) + (h.z * nf);
    f3 I = -wo;
    if (r <= pS) {
        f3 wi = I - ((hw * dot(hw, I)) * 2.0f);
        if (dot(wi, nf) <= 0.0f) return mk3(0, 0, 0);
        return normalize(wi);
    }
    float eta = ei / et;
    float dv = dot(hw, I);
    float k = 1.0f - eta * eta * (1.0f - dv * dv);
    if (!(k >= 0.0f)) return mk3(0, 0, 0);
    f3 wi = (eta * I) - ((eta * dv + lxm_sqrt(k)) * hw);
    if (dot(wi, wi) < LX_CT_EPS || lxm_isnan(wi.x) || lxm_isnan(wi.y) || lxm_isnan(wi.z) || dot(wi, nf) >= 0.0f)
        return mk3(0, 0, 0);
    return normalize(wi);
}
LXD f3 lx_sample_ct(f3 wo, f3 n, float rough, float metal, LxRng& rng, f3 color, float ior, float tr) {
    return lx_sample_ct_q(wo, n, rough, rng, ior, lx_lobes_of(wo, n, color, metal, ior, tr));
}
/* PDF_GGX, cook_torrance.cu:226-237 */
LXD float lx_pdf_ggx(f3 wo, f3 wi, f3 n, float rough) {
    f3 h = normalize(wo + wi);
    rough = lxm_clamp(rough, 0.0001f, 1.0f);
    float alpha = rough * rough;
    float ph = lx_D(h, n, alpha) * dot(n, h);
    return ph / lxm_max((4.0f * lxm_abs(dot(wo, h))), LX_CT_EPS);
}
/* PDF_GGX_Transmission, cook_torrance.cu:239-258 */
LXD float lx_pdf_ggx_t(f3 wo, f3 wi, f3 n, float rough, float ior) {
    float ci = dot(wo, n);
    bool enter = ci > 0.f;
    float ei = enter ? 1.0f : ior, et = enter ? ior : 1.0f;
    f3 ht = normalize(-((ei * wo) + (et * wi)));
    rough = lxm_clamp(rough, 0.0001f, 1.0f);
    float alpha = rough * rough;
    float den = et * dot(wi, ht) + ei * dot(wo, ht);
    float ph = lx_D(ht, n, alpha) * lxm_abs(dot(n, ht));
    return ph * (et * et * lxm_abs(dot(wi, ht))) / lxm_max(den * den, LX_CT_EPS);
}
/* PDF, cook_torrance.cu:260-295, with the lobe probabilities given */
LXD float lx_pdf_q(f3 wo, f3 wi, f3 n, float rough, float ior, const LxLobes& q) {
    ior = lxm_max(ior, 1.0001f);
    f3 nf = dot(wo, n) < 0.0f ? -n : n;
    float dwn = dot(wi, nf);
    float pd = (dwn > 0.0f ? dwn : 0.0f) * LX_INV_PI_CT;
    bool refract = (dot(wi, n) * dot(wo, n)) < 0.0f;
    float pdf;
    if (refract) pdf = q.pR * lx_pdf_ggx_t(wo, wi, n, rough, ior);
    else pdf = q.pD * pd + q.pS * lx_pdf_ggx(wo, wi, nf, rough);
    return lxm_max(pdf, LX_PDF_CLAMP);
}
LXD float lx_pdf(f3 wo, f3 wi, f3 n, float rough, float metal, f3 color, float ior, float tr) {
    return lx_pdf_q(wo, wi, n, rough, ior, lx_lobes_of(wo, n, color, metal, ior, tr));
}
/* shadePathCookTorrance, cook_torrance.cu:298-331 (b: the BRDF's view-side
 * terms of the same wo, n, rough, ior) */
LXD f3 lx_ct_eval_cos_v(const LxBrdfView& b, f3 wo, f3 n, f3 wi, f3 albedo, float rough, float metal, float ior,
                        float tr) {
    ior = lxm_max(ior, 1.0001f);
    bool refract = (dot(wi, n) * dot(wo, n)) < 0.0f;
    f3 f = refract ? lx_btdf(wo, n, wi, albedo, rough, metal, ior, tr) : lx_brdf_v(b, wi, albedo, metal, tr);
    return f * lxm_abs(dot(wi, n));
}
LXD f3 lx_ct_eval_cos(f3 wo, f3 n, f3 wi, f3 albedo, float rough, float metal, float ior, float tr) {
    return lx_ct_eval_cos_v(lx_brdf_view(wo, n, rough, ior), wo, n, wi, albedo, rough, metal, ior, tr);
}

/* ---------------------------------------------------------- CDF sampling */
/* cudaUtils::select_from_cdf, src/common.cu:3-15 */
LXD int lx_select_cdf(const float* cdf_, int size, float xi) {
    const LX_GLOBAL float* cdf = lx_gptr(cdf_);
    int lo = 0, hi = size - 1;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (cdf[mid] < xi) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* envmap_pdf, src/kernels/shade_path.cu:27-42 */
LXD float lx_envmap_pdf(const LxDevScene& S, f3 d) {
    int W = S.env.w, H = S.env.h;
    float phi = lxm_atan2(d.z, d.x);
    float theta = lxm_acos(lxm_clamp(d.y, -1.0f, 1.0f));
    float u = (phi + LX_PI_F) / (2.0f * LX_PI_F);
    float v = theta / LX_PI_F;
    int col = (int)(u * W);
    col = col < 0 ? 0 : (col > W - 1 ? W - 1 : col);
    int row = (int)(v * H);
    row = row < 0 ? 0 : (row > H - 1 ? H - 1 : row);
    const LX_GLOBAL float* marg = lx_gptr(S.marginal);
    const LX_GLOBAL float* cond = lx_gptr(S.conditional);
    float pm = (marg[row] - (row == 0 ? 0.0f : marg[row - 1])) * H;
    size_t ci = (size_t)row * W + col;
    float pc = (cond[ci] - (col == 0 ? 0.0f : cond[ci - 1])) * W;
    float st = lxm_sin(theta);
    return (st > 1e-6f) ? (pm * pc) / (2.0f * LX_PI_F * LX_PI_F * st) : 0.0f;
}

/* Environment radiance lookup of shade_path.cu:205-215 and 309-318 (the file
 * redefines M_PI as the double 3.14159, shade_path.cu:23). */
LXD f3 lx_env_lookup(const LxDevScene& S, f3 d, float intensity) {
    const double MPI = 3.14159;
    float phi = lxm_atan2(d.z, d.x);
    float theta = lxm_acos(lxm_clamp(d.y, -1.0f, 1.0f));
    float u = (float)(((double)phi + MPI) * (double)(1.0f / (2.0f * MPI)));
    float v = (float)((double)theta * (double)(1.0f / MPI));
    float t[4];
    lx_tex_sample(S.env, u, v, t);
    return mk3(t[0], t[1], t[2]) * intensity;
}

/* sampleDirectLight emitter branch, src/kernels/sample_direct_light.cu:32-66 */
LXD f3 lx_sample_light_point(const LxDevScene& S, LxRng& rng) {
    float rnd = rng.u01();
    int sel = lx_select_cdf(S.emissive_prefix, (int)S.num_emissive, rnd);
    const LX_GLOBAL LxInstance& g = lx_gptr(S.inst)[lx_gptr(S.emissive_inst)[sel]];
    const LX_GLOBAL float* prefix = lx_gptr(S.emissive_prefix);
    float lower = (sel == 0) ? 0.0f : prefix[sel - 1];
    float upper = prefix[sel];
    float den = upper - lower;
    float to = (den > 1e-10f) ? (rnd - lower) / den : 0.0f;
    to = lxm_clamp(to, 0.0f, 1.0f);
    uint32_t co = lx_gptr(S.tri_cdf_off)[sel];
    int ntri = (int)(lx_gptr(S.tri_cdf_off)[sel + 1] - co);
    int tri = lx_select_cdf(S.tri_cdf + co, ntri, to);
    const LX_GLOBAL uint32_t* ix = lx_gptr(S.idx) + g.idx_off + 3u * (uint32_t)tri;
    const LX_GLOBAL float* P = lx_gptr(S.pos) + 3u * (size_t)g.vtx_off;
    f3 v0 = mk3(P[3 * ix[0]], P[3 * ix[0] + 1], P[3 * ix[0] + 2]);
    f3 v1 = mk3(P[3 * ix[1]], P[3 * ix[1] + 1], P[3 * ix[1] + 2]);
    f3 v2 = mk3(P[3 * ix[2]], P[3 * ix[2] + 1], P[3 * ix[2] + 2]);
    float r1 = lxm_sqrt(rng.u01());
    float r2 = rng.u01();
    float u = 1.0f - r1;
    float v = r2 * r1;
    f3 local = ((u * v0) + (v * v1)) + ((1.0f - u - v) * v2);
    return xf_point((const float*)g.m, local);
}

/* sampleDirectLight environment branch, sample_direct_light.cu:68-100 */
LXD f3 lx_sample_env_dir(const LxDevScene& S, LxRng& rng) {
    int W = S.env.w, H = S.env.h;
    float rnd = rng.u01();
    int mg = lx_select_cdf(S.marginal, H, rnd);
    const LX_GLOBAL float* marg = lx_gptr(S.marginal);
    float lower = (mg == 0) ? 0.0f : marg[mg - 1];
    float upper = marg[mg];
    float den = upper - lower;
    float off = (den > 1e-10f) ? (rnd - lower) / den : 0.0f;
    off = lxm_clamp(off, 0.0f, 1.0f);
    int cd = lx_select_cdf(S.conditional + (size_t)mg * W, W, off);
    float u = (cd + 0.5f) / W;
    float v = (mg + 0.5f) / H;
    float phi = (float)((double)u * 2.0 * (double)LX_PI_F - (double)LX_PI_F);
    float theta = v * LX_PI_F;
    float st = lxm_sin(theta);
    return mk3(st * lxm_cos(phi), lxm_cos(theta), st * lxm_sin(phi));
}

/* get_alpha, src/material_queries.cu:119-137 */
LXD float lx_get_alpha(const LxDevScene& S, const lx_material_desc& m, f2 uv) {
    if (m.alpha_mode == LX_ALPHA_OPAQUE) return 1.0f;
    float a = m.alpha;
    if (m.albedo_tex >= 0) {
        uv = lx_apply_xf(uv, m.albedo_xf);
        float t[4];
        lx_tex_sample(S.texs[m.albedo_tex], uv.x, uv.y, t);
        a *= t[3];
    }
    if (m.alpha_mode == LX_ALPHA_MASK) return a >= m.alpha_cutoff ? 1.0f : 0.0f;
    return a;
}

/* The raw normal-map texel get_normal reports for the debug view
 * (material_queries.cu:36-39, 83): (0.5, 0.5, 1) without a normal map. */
LXD f3 lx_normal_map_rgb(const LxDevScene& S, const lx_material_desc& m, f2 uv) {
    if (m.normal_tex < 0) return mk3(0.5f, 0.5f, 1.0f);
    uv = lx_apply_xf(uv, m.normal_xf);
    float t[4];
    lx_tex_sample(S.texs[m.normal_tex], uv.x, uv.y, t);
    return mk3(t[0], t[1], t[2]);
}

/* render_material_debug_mode, src/material_debug_render.cu:3-26 */
LXD f3 lx_debug_color(const LxDevScene& S, const lx_material_desc& m, uint32_t mode, f3 albedo, f3 normal, f2 uv,
                      float rough, float metal) {
    switch (mode) {
    case LX_DEBUG_ALBEDO: return albedo;
    case LX_DEBUG_WORLD_NORMAL: return mk3(normal.x * 0.5f + 0.5f, normal.y * 0.5f + 0.5f, normal.z * 0.5f + 0.5f);
    case LX_DEBUG_NORMAL_MAP: return lx_normal_map_rgb(S, m, uv);
    case LX_DEBUG_ROUGHNESS: return mk3(rough, rough, rough);
    case LX_DEBUG_METALLIC: return mk3(metal, metal, metal);
    case LX_DEBUG_ALPHA: {
        float a = lx_get_alpha(S, m, uv);
        return mk3(a, a, a);
    }
    default: return mk3(0.0f, 0.0f, 0.0f);
    }
}

/* ------------------------------------------- legacy (UBER_SHADER 0) shaders */
/* glm::reflect(I, N) = I - N * dot(N, I) * 2 */
LXD f3 lx_reflect(f3 I, f3 N) { return I - ((N * dot(N, I)) * 2.0f); }

/* Lambert::PDF, src/shaders/lambert.cu:9-12 */
LXD float lx_lambert_pdf(f3 wi, f3 n) { return lxm_max(lxm_max(dot(wi, n), 0.0f) * LX_INV_PI_CT, 1e-6f); }

/* TransmissiveGlass::FresnelDielectricEval, src/shaders/glass.cu:27-55 (the
 * glass shader's own Fresnel, which takes the raw ior) */
LXD float lx_glass_fresnel(float ci, float ior) {
    float ei = 1.0f, et = ior;
    ci = lxm_clamp(ci, -1.0f, 1.0f);
    if (!(ci > 0.0f)) {
        float tmp = ei;
        ei = et;
        et = tmp;
        ci = lxm_abs(ci);
    }
    float si = lxm_sqrt(lxm_max(0.0f, 1.0f - ci * ci));
    float st = (ei / et) * si;
    if (st >= 1.0f) return 1.0f;
    float ct = lxm_sqrt(lxm_max(0.0f, 1.0f - st * st));
    float rpa = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
    float rpe = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
    return (rpa * rpa + rpe * rpe) / 2.0f;
}

/* TransmissiveGlass::sampleGlass with sampleSpecularRefl / sampleSpecularTrans
 * (glass.cu:5-25, 57-75); glm::refract yields NaN under total internal
 * reflection, which the shader turns into the mirror direction. */
LXD f3 lx_sample_glass(f3 d, f3 n, float ior, LxRng& rng) {
    f3 wo = -d;
    float r = rng.u01();
    float F = lxm_clamp(lx_glass_fresnel(dot(wo, n), ior), 0.0f, 1.0f);
    bool enter = dot(wo, n) > 0.0f;
    f3 ln = enter ? n : -n;
    f3 I = -wo;
    if (r < F) return lx_reflect(I, ln);
    float eta = enter ? (1.0f / ior) : (ior / 1.0f);
    float dv = dot(ln, I);
    float k = 1.0f - eta * eta * (1.0f - dv * dv);
    if (!(k >= 0.0f)) return lx_reflect(I, ln);
    f3 wi = (eta * I) - ((eta * dv + lxm_sqrt(k)) * ln);
    if (dot(wi, wi) < 1e-6f || lxm_isnan(wi.x) || lxm_isnan(wi.y) || lxm_isnan(wi.z)) return lx_reflect(I, ln);
    return normalize(wi);
}

/* sample_materials legacy dispatch, src/sample_materials.cu:24-37 (the
 * microfacet class takes the uber-shader sampler) */
LXD f3 lx_sample_legacy(int type, f3 d, f3 normal, float ior, LxRng& rng) {
    if (type == LX_MATERIAL_DIFFUSE || type == LX_MATERIAL_EMISSIVE) return lx_cos_hemisphere(normal, rng);
    if (type == LX_MATERIAL_SPECULAR) return lx_reflect(-(-d), normal); /* sampleMirror, specular.cu:14-19 */
    return lx_sample_glass(d, normal, ior, rng);                         /* LX_MATERIAL_GLASS */
}

/* get_pdf_bsdf / get_brdf legacy branches, src/kernels/shade_path.cu:60-92:
 * Lambert for diffuse, zero for the delta and emissive classes. */
LXD float lx_legacy_pdf(int type, f3 wi, f3 normal) {
    return type == LX_MATERIAL_DIFFUSE ? lx_lambert_pdf(wi, normal) : 0.0f;
}
LXD f3 lx_legacy_brdf(int type, f3 color) {
    return type == LX_MATERIAL_DIFFUSE ? color / LX_PI_F : mk3(0.0f, 0.0f, 0.0f);
}

/* update_throughput_materials legacy branches for the non-microfacet
 * classes, src/update_throughput_materials.cu:36-62; returns last_pdf (the
 * emissive class leaves the path and last_pdf as they are). */
LXD float lx_legacy_update(int type, f3 d, f3 normal, f3 wi, f3 color, float ior, float tr, f3& thr, float last_pdf) {
    if (type == LX_MATERIAL_DIFFUSE) {
        /* Lambert::shadePathLambert / PDF, lambert.cu:5-30 */
        f3 f = (color * LX_INV_PI_CT) * lxm_max(dot(wi, normal), 0.0f);
        float pdf = lx_lambert_pdf(wi, normal);
        thr = thr * (f / pdf);
        return pdf;
    }
    if (type == LX_MATERIAL_SPECULAR) { /* shadePathSpecular, specular.cu:5-12 */
        thr = thr * color;
        return 1.0f;
    }
    if (type == LX_MATERIAL_GLASS) { /* shadePathGlass, glass.cu:77-95 */
        f3 wo = -d;
        float ci = dot(wo, normal);
        bool enter = ci > 0.f;
        float ei = enter ? 1.0f : ior, et = enter ? ior : 1.0f;
        bool refr = dot(wo, normal) * dot(wi, normal) < 0.0f;
        float at = refr ? (ei / et) : 1.0f;
        at = at * at;
        f3 tint = refr ? mk3(1.0f + tr * (color.x - 1.0f), 1.0f + tr * (color.y - 1.0f), 1.0f + tr * (color.z - 1.0f))
                       : mk3(1.0f, 1.0f, 1.0f);
        thr = thr * (at * tint);
        return 1.0f;
    }
    return last_pdf;
}

#endif
