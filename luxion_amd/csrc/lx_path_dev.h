/*
 * lx_path_dev.h -- one bounce of one path, in registers.
 *
 * lx_trace_path: sampleDirectLight (src/kernels/sample_direct_light.cu:9-101)
 *   + the raygen's radiance / direct-light / environment rays and the
 *   closest-hit attribute programs (src/optixshaders/optix_triangle.cu:52-305).
 * lx_shade_path: shadePath (src/kernels/shade_path.cu:94-364) with
 *   update_throughput_materials (src/update_throughput_materials.cu:8-67).
 * The wavefront kernels store the record between the two in SoA buffers; the
 * tail kernel keeps it in registers.  Both give the same bits.
 */
#ifndef LX_PATH_DEV_H
#define LX_PATH_DEV_H

#include "lx_trace_dev.h"

struct LxRec {
    float t;
    f3 N;
    f2 uv;
    int mat;
    f3 Tg;
    float env_t;
    f3 ed;
    float dl_t;
    f3 dlN;
    f2 dluv;
    int dlmat;
    f3 lp;
    f3 so; /* origin of the environment shadow ray (hit point + 1e-4 N) */
};

struct LxRayCount {
    uint32_t rad, dl, env;
};

/* Sink receives the record piece by piece (hit, emitter shadow hit, light
 * point, environment occlusion) so a caller that stores to memory can free
 * the registers early; LxRecSink keeps them in an LxRec. */
struct LxRecSink {
    LxRec& r;
    LXD void hit(float t, f3 N, f2 uv, int mat, bool tan, f3 Tg) {
        r.t = t;
        r.N = N;
        r.uv = uv;
        r.mat = mat;
        if (tan) r.Tg = Tg;
    }
    LXD void miss() { r.t = -1.0f; }
    LXD void dl(float t, f3 N, f2 uv, int mat) {
        r.dl_t = t;
        r.dlN = N;
        r.dluv = uv;
        r.dlmat = mat;
    }
    LXD void dl_miss() { r.dl_t = -1.0f; }
    LXD void light(f3 lp) { r.lp = lp; }
    LXD void env_dir(f3 ed) { r.ed = ed; }
    LXD void env(float env_t) { r.env_t = env_t; }
    LXD void env_origin(f3 so) { r.so = so; }
};

template <bool CNT, int W, class Sink>
LXD void lx_trace_path(const LxJob& J, const LxDevScene& S, f3 o, f3 d, uint32_t pix, int bounce, LxStack& st,
                       uint32_t* overflow, Sink& out, LxRayCount& rc, LxWork* wc, LxWork* wa) {
    f3 lp = mk3(0, 0, 0), ed = mk3(0, 0, 0);
    if (J.dmis || J.emis) {
        /* sampleDirectLight: RNG keyed on the pre-increment depth */
        LxRng rng = lx_make_rng(J.iter, (int)pix, bounce, J.seed);
        if (J.dmis) lp = lx_sample_light_point(S, rng);
        if (J.emis) ed = lx_sample_env_dir(S, rng);
    }
    LxHitD h = lx_trace<false, CNT, W>(S, o, d, 1e16f, pix, (uint32_t)J.iter, LX_INST_ALPHA, st, overflow, wc);
    rc.rad++;
    if (J.dmis) out.light(lp);
    if (J.emis) out.env_dir(ed);
    if (h.inst != 0xffffffffu) {
        bool want_t = lx_gptr(S.mats)[lx_gptr(S.inst)[h.inst].material].normal_tex >= 0;
        f3 N, Tg;
        f2 uv;
        int mat;
        lx_hit_attrs(S, h, want_t, N, Tg, uv, mat);
        out.hit(h.t, N, uv, mat, want_t, Tg);
        f3 hp = o + (h.t * d);
        f3 so = hp + (N * 1e-4f);
        if (J.dmis) {
            f3 ld = lp - so;
            float dist = length(ld);
            ld = ld * (1.0f / dist);
            LxHitD hd = lx_trace<false, CNT, W>(S, so, ld, dist * 1.0001f + 1e-4f, pix, (uint32_t)J.iter, LX_INST_ALPHA,
                                             st, overflow, wc);
            rc.dl++;
            if (hd.inst != 0xffffffffu) {
                f3 dN;
                f2 duv;
                int dmat;
                lx_dl_attrs(S, hd, dN, duv, dmat);
                out.dl(hd.t, dN, duv, dmat);
            } else {
                out.dl_miss();
            }
        }
        if (J.emis) {
            LxHitD he = lx_trace<true, CNT, W>(S, so, ed, 1e16f, pix, (uint32_t)J.iter, LX_INST_ALPHA, st, overflow, wa);
            rc.env++;
            out.env(he.inst != 0xffffffffu ? 1.0f : -1.0f);
        } else {
            out.env(1.0f);
        }
    } else {
        out.miss();
        if (J.dmis) out.dl_miss();
        out.env(-0.5f);
    }
}

/* The tail kernel's form of lx_trace_path: the environment shadow ray of
 * the previous bounce (pend_*; its result goes to pend_occluded) is traced in
 * the same traversal loop as this bounce's radiance ray, and this bounce's
 * environment shadow ray is not traced here but left in the record (so, ed)
 * for the next call.  has_rad false traces only the pending ray (a path that
 * ended with one outstanding).  The direct-light ray follows; with DLQ it is
 * answered by lx_trace_dl, whose hit is kept only when it emits (a
 * non-emitting closest hit adds nothing to the colour), so the per-bounce
 * form's records, not this form's, carry the occluders. */
template <bool CNT, int W, bool DLQ = false>
LXD void lx_trace_path_d(const LxJob& J, const LxDevScene& S, f3 o, f3 d, uint32_t pix, int bounce, LxStack& st,
                         uint32_t* overflow, bool pend, f3 pend_o, f3 pend_d, bool has_rad, bool& pend_occluded,
                         LxRecSink& out, LxRayCount& rc, LxWork* wc, LxWork* wa) {
    LxHitD h = lx_trace_pair<CNT, W>(S, pend, pend_o, pend_d, 1e16f, has_rad, o, d, 1e16f, pix, (uint32_t)J.iter,
                                      LX_INST_ALPHA, st, overflow, pend_occluded, wc, wa);
    if (pend) rc.env++;
    if (!has_rad) return;
    rc.rad++;
    /* the miss values first, overwritten below on hits: stores of the same
     * constant to different fields in different branches would otherwise
     * be merged into one store through a selected address, which keeps the
     * record in scratch memory */
    out.miss();
    if (J.dmis) out.dl_miss();
    out.env(-0.5f);
    if (h.inst != 0xffffffffu) {
        bool want_t = lx_gptr(S.mats)[lx_gptr(S.inst)[h.inst].material].normal_tex >= 0;
        f3 N, Tg;
        f2 uv;
        int mat;
        lx_hit_attrs(S, h, want_t, N, Tg, uv, mat);
        out.hit(h.t, N, uv, mat, want_t, Tg);
        /* sampleDirectLight: RNG keyed on the pre-increment depth, a stream
         * of its own.  Its samples do not depend on the trace and only a hit
         * uses them, so they are drawn after it and only on hits (fewer
         * registers live across the traversal loop, no sampling in waves
         * whose rays all miss; the same values) */
        f3 lp = mk3(0, 0, 0), ed = mk3(0, 0, 0);
        if (J.dmis || J.emis) {
            LxRng rng = lx_make_rng(J.iter, (int)pix, bounce, J.seed);
            if (J.dmis) lp = lx_sample_light_point(S, rng);
            if (J.emis) ed = lx_sample_env_dir(S, rng);
        }
        if (J.dmis) out.light(lp);
        if (J.emis) out.env_dir(ed);
        f3 hp = o + (h.t * d);
        f3 so = hp + (N * 1e-4f);
        if (J.dmis) {
            f3 ld = lp - so;
            float dist = length(ld);
            ld = ld * (1.0f / dist);
            /* DLQ: only an emitting closest hit reaches the colour */
            LxHitD hd = DLQ ? lx_trace_dl<CNT, W>(S, so, ld, dist * 1.0001f + 1e-4f, pix, (uint32_t)J.iter, st,
                                                  overflow, wc)
                            : lx_trace<false, CNT, W>(S, so, ld, dist * 1.0001f + 1e-4f, pix, (uint32_t)J.iter,
                                                      LX_INST_ALPHA, st, overflow, wc);
            rc.dl++;
            if (hd.inst != 0xffffffffu) {
                f3 dN;
                f2 duv;
                int dmat;
                lx_dl_attrs(S, hd, dN, duv, dmat);
                out.dl(hd.t, dN, duv, dmat);
            }
        }
        out.env_origin(so);
        out.env(1.0f);
    }
}

/* Updates the path registers; returns true if the path continues to the
 * next bounce (compaction push), false when it stops.  DEFER: the
 * environment-light term is not added here but returned in pend_add /
 * pend_col, to be added once the environment shadow ray (traced later) is
 * known to be unoccluded -- before anything else is added to the colour, so
 * the sums are the same. */
template <bool DEFER = false>
LXD bool lx_shade_path(const LxJob& J, const LxDevScene& S, const LxRec& R, int depth, bool compact, f3& o, f3& d,
                       f3& thr, f3& col, float& last_pdf, bool& spec, uint32_t pix, bool* pend_add = nullptr,
                       f3* pend_col = nullptr) {
    if (DEFER) *pend_add = false;
    /* only a hit draws from the bounce's stream (sampling, roulette) */
    LxRng rng;
    rng.x = 1u;
    bool kill = false;
    float t = R.t;
    f3 N = mk3(0, 0, 0);
    f3 sample_dir = mk3(0, 0, 0);
    bool has_exr = S.has_env != 0;
    /* The environment terms (the MIS weight of an emitter hit, the
     * environment light sample at a hit, a miss's radiance) are requested
     * per lane in the branches and evaluated after them, once for every lane
     * of the wave: a wave whose lanes both hit and miss runs envmap_pdf and
     * the lookup once instead of once per branch.  Same inputs, same
     * expressions, same bits. */
    int env_case = 0; /* 1: emitter hit under -emis, 2: environment light sample, 3: miss */
    bool want_pe = false, want_L = false;
    f3 pe_dir = mk3(0, 0, 0), L_dir = mk3(0, 0, 0);
    f3 ev = mk3(0, 0, 0); /* case 1: the emission; case 2: the BRDF */
    float e_pb = 0.0f, e_cs = 0.0f;
    /* the hit's shading terms, used again by the throughput update below */
    bool shaded = false, legacy = false, is_specular = false;
    int mt = LX_MATERIAL_MICROFACET;
    f3 color = mk3(0, 0, 0), normal = mk3(0, 0, 0);
    float rough = 0.0f, metal = 0.0f, m_ior = 1.0f, m_tr = 0.0f;
    LxLobes lobes{};
    LxBrdfView bview{};
    if (t > 0.0f) {
        rng = lx_make_rng(J.iter, (int)pix, depth, J.seed);
        N = R.N;
        f2 uv = R.uv;
        const lx_material_desc& mat = S.mats[R.mat];
        /* legacy shading (UBER_SHADER 0) dispatches on the material class; the
         * uber shader treats every material as microfacet */
        legacy = (J.flags & LX_JF_LEGACY) != 0u;
        mt = legacy ? mat.material_type : LX_MATERIAL_MICROFACET;
        color = lx_get_albedo(S, mat, uv);
        /* the legacy loader gives an emissive material its emission colour as
         * base colour (src/scene/scene_load_gltf.cpp:211-213) */
        if (mt == LX_MATERIAL_EMISSIVE && mat.albedo_tex < 0)
            color = mk3(mat.emission_color[0], mat.emission_color[1], mat.emission_color[2]);
        normal = N;
        if (mat.normal_tex >= 0) normal = lx_get_normal(S, mat, N, R.Tg, uv);
        lx_get_mr(S, mat, uv, rough, metal);
        m_ior = mat.ior;
        m_tr = mat.transmission;
        f3 emission = lx_get_emission(S, mat, uv);
        is_specular = legacy ? (mt == LX_MATERIAL_SPECULAR || mt == LX_MATERIAL_GLASS) : rough < 0.05f;
        f3 nf = dot(-d, normal) < 0.0f ? -normal : normal;
        const uint32_t dbg = J.flags & LX_JF_DEBUG_MASK;
        if (dbg != 0u) {
            /* render_material_debug_mode replaces the colour and ends the path */
            col = lx_debug_color(S, mat, dbg, color, normal, uv, rough, metal);
            kill = true;
        } else {
            shaded = true;
            /* the lobe probabilities and the BRDF's view-side terms of this
             * shading point, shared by the sampler and every PDF / BRDF below */
            if (mt == LX_MATERIAL_MICROFACET) {
                lobes = lx_lobes_of(-d, normal, color, metal, m_ior, m_tr);
                bview = lx_brdf_view(-d, normal, rough, m_ior);
                sample_dir = lx_sample_ct_q(-d, normal, rough, rng, m_ior, lobes);
            }
            else
                sample_dir = lx_sample_legacy(mt, d, normal, m_ior, rng);
            bool emissive = legacy ? mt == LX_MATERIAL_EMISSIVE : length(emission) > 0.00001f;
            if (J.emis) {
                if (emissive) {
                    env_case = 1;
                    ev = emission;
                    if (depth > 1 && !spec) {
                        want_pe = true;
                        pe_dir = d;
                        e_pb = last_pdf;
                    }
                    kill = true;
                } else if ((DEFER || R.env_t < 0.0f) && !is_specular && has_exr) {
                    f3 ed = R.ed;
                    env_case = 2;
                    want_pe = want_L = true;
                    pe_dir = ed;
                    L_dir = normalize(ed);
                    e_pb = mt == LX_MATERIAL_MICROFACET ? lx_pdf_q(-d, ed, normal, rough, m_ior, lobes)
                                                        : lx_legacy_pdf(mt, ed, normal);
                    ev = mt == LX_MATERIAL_MICROFACET ? lx_brdf_v(bview, ed, color, metal, m_tr)
                                                      : lx_legacy_brdf(mt, color);
                    e_cs = dot(nf, ed);
                }
            } else if (J.dmis) {
                if (emissive) {
                    float w = 1.0f;
                    if (depth > 1 && !spec) {
                        float pb = last_pdf;
                        float dist_sq = t * t;
                        float cl = dot(N, -d);
                        if (cl > 0.0001f) {
                            float pa = 1.0f / S.total_area;
                            float psa = pa * (dist_sq / cl);
                            w = (pb * pb) / (pb * pb + psa * psa);
                        } else {
                            w = 0.0f;
                        }
                    }
                    col = col + thr * (w * emission);
                    kill = true;
                } else if (R.dl_t > 0.0f && !is_specular) {
                    f3 sp = (t * d) + o;
                    f3 lv = R.lp - sp;
                    float dsq = dot(lv, lv);
                    float dist = lxm_sqrt(dsq);
                    f3 ld = lv / dist;
                    float cs = dot(nf, ld);
                    float cl = dot(R.dlN, -ld);
                    if (cs > 0.0f && cl > 0.0f) {
                        float pa = 1.0f / S.total_area;
                        float psa = pa * (dsq / cl);
                        float pb = mt == LX_MATERIAL_MICROFACET
                                       ? lx_pdf_q(-d, ld, normal, rough, m_ior, lobes)
                                       : lx_legacy_pdf(mt, ld, normal);
                        float w = (psa * psa) / (psa * psa + pb * pb);
                        f3 brdf = mt == LX_MATERIAL_MICROFACET
                                      ? lx_brdf_v(bview, ld, color, metal, m_tr)
                                      : lx_legacy_brdf(mt, color);
                        f3 Le = lx_get_emission(S, S.mats[R.dlmat], R.dluv);
                        f3 contrib = ((Le * brdf) * cs) / psa;
                        col = col + (thr * contrib) * w;
                    }
                }
            } else if (emissive) {
                col = col + thr * emission;
                kill = true;
            }
        }
    } else if (has_exr) {
        env_case = 3;
        f3 dn = normalize(d);
        want_L = true;
        L_dir = dn;
        if (J.emis && depth > 1 && !spec) {
            want_pe = true;
            pe_dir = dn;
            e_pb = last_pdf;
        }
    }
    /* the environment terms, with the throughput before this bounce's update */
    float pe = 0.0f;
    f3 L = mk3(0, 0, 0);
    if (want_pe) pe = lx_envmap_pdf(S, pe_dir);
    if (want_L) L = lx_env_lookup(S, L_dir, J.intensity);
    if (env_case == 1) {
        float w = 1.0f;
        if (want_pe) w = (e_pb * e_pb) / (e_pb * e_pb + pe * pe);
        col = col + thr * (w * ev);
    } else if (env_case == 2) {
        float w = (pe * pe) / (pe * pe + e_pb * e_pb);
        if (e_cs > 0.0f && pe > 1e-8f) {
            f3 contrib = ((L * ev) * e_cs) / pe;
            if (DEFER) {
                *pend_add = true;
                *pend_col = (thr * contrib) * w;
            } else {
                col = col + (thr * contrib) * w;
            }
        }
    } else if (env_case == 3) {
        float w = 1.0f;
        if (want_pe) w = (e_pb * e_pb) / (e_pb * e_pb + pe * pe);
        col = col + (thr * L) * w;
    }
    if (shaded && !kill) {
        /* update_throughput_materials, src/update_throughput_materials.cu:8-67 */
        if (mt == LX_MATERIAL_MICROFACET) {
            /* only the uber build counts a transmission as specular */
            bool transmitted = !legacy && dot(-d, normal) * dot(sample_dir, normal) < 0.0f;
            bool nspec = is_specular || transmitted;
            f3 ct = lx_ct_eval_cos_v(bview, -d, normal, sample_dir, color, rough, metal, m_ior, m_tr);
            float pdf = lx_pdf_q(-d, sample_dir, normal, rough, m_ior, lobes);
            thr = thr * (ct / pdf);
            last_pdf = pdf;
            spec = nspec;
        } else {
            last_pdf = lx_legacy_update(mt, d, normal, sample_dir, color, m_ior, m_tr, thr, last_pdf);
            spec = is_specular;
        }
        if (depth >= 3) {
            float mx = lxm_max(thr.x, lxm_max(thr.y, thr.z));
            float sv = lxm_clamp(mx, 0.0f, 1.0f);
            if (rng.u01() >= sv) kill = true;
            else thr = thr / sv;
        }
    }
    if (t == -1.0f) {
        kill = true;
    } else {
        /* getPointOnRay (include/intersections.h:29-32) and the spawn offset */
        f3 hp = o + ((t - .0001f) * normalize(d));
        float l2 = dot(sample_dir, sample_dir);
        if (!(l2 > 1e-12f) || !lxm_isfinite(l2)) return false;
        d = sample_dir;
        if (dot(d, N) > 0.0f) o = hp + (N * 0.01f);
        else o = hp - (N * 0.01f);
    }
    return compact && !kill;
}

#endif
