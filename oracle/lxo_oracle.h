/*
 * lxo_oracle.h -- C interface of the CPU oracle (test infrastructure only).
 *
 * The oracle is a from-scratch CPU restatement of the reference renderer's
 * hot path (seabiscuit-iv/luxion, src/pathtrace.cu and the files it calls).
 * Only tests, the smoke check and the benchmark's CPU-baseline leg load it;
 * the product library never links or calls it.
 */
#ifndef LXO_ORACLE_H
#define LXO_ORACLE_H

#include <stdint.h>
#include "../include/luxion_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lxo_scene lxo_scene;

enum { LXO_TRAVERSE_BVH = 0, LXO_TRAVERSE_BRUTE = 1 };

/* cdf_mode: LX_CDF_FP64 or LX_CDF_FP32_REFERENCE. build_bvh: build the
 * reference-style binary SAH BVH per mesh (needed for LXO_TRAVERSE_BVH). */
lxo_scene* lxo_scene_create(const lx_scene_desc* desc, int cdf_mode, int build_bvh);
void lxo_scene_destroy(lxo_scene* s);
double lxo_scene_bvh_build_ms(const lxo_scene* s);
uint32_t lxo_scene_bvh_nodes(const lxo_scene* s, uint32_t mesh);
int lxo_scene_bvh_depth(const lxo_scene* s, uint32_t mesh);

int lxo_read_env_cdfs(const lxo_scene* s, float* marginal, float* conditional);
int lxo_read_emitter_cdfs(const lxo_scene* s, uint32_t* emissive_instances, float* instance_prefix,
                          uint32_t* num_emissive, float* total_area);
int lxo_read_instance_tri_cdf(const lxo_scene* s, uint32_t emissive_index, float* tri_prefix, uint32_t* n);

/* pixel: per-ray key of the blend any-hit hash (may be NULL = ray index). */
int lxo_trace(const lxo_scene* s, const lx_ray* rays, uint32_t n, int mode, int alpha, uint32_t iteration,
              const uint32_t* pixel, lx_hit* hits, uint8_t* occluded);

typedef struct lxo_render_opts {
    int32_t width, height;
    int32_t max_depth;
    int32_t dmis, emis;
    float envmap_intensity;
    uint32_t seed;
    lx_camera_desc camera;
    int32_t traverse_mode;
    int32_t threads;
    int32_t material_debug_mode; /* LX_DEBUG_* */
    int32_t shading_model;       /* LX_SHADING_* */
} lxo_render_opts;

typedef struct lxo_stats {
    uint64_t radiance_rays, dl_shadow_rays, env_shadow_rays, path_bounces;
    double wall_ms;
} lxo_stats;

/* Render iterations first_iter .. first_iter+n_iters-1 for the listed pixels
 * (NULL = all W*H in order); rgb_sum gets 3 floats per listed pixel. */
int lxo_render(const lxo_scene* s, const lxo_render_opts* o, uint32_t first_iter, uint32_t n_iters,
               const uint32_t* pixels, uint32_t n_pixels, float* rgb_sum, lxo_stats* stats);

/* One iteration, bounce-resolved: survivors (pixel indices, sorted per
 * bounce) and hit records (8 floats per shaded path: t, nx, ny, nz, u, v,
 * material, pixel; sorted by pixel within a bounce). */
int lxo_debug_bounce(const lxo_scene* s, const lxo_render_opts* o, uint32_t iter, const uint32_t* pixels,
                     uint32_t n_pixels, uint32_t* survivors, uint32_t* survivor_counts, float* records,
                     uint32_t* record_counts, float* rgb_sum);

/* Building blocks exposed for known-answer and property tests. */
uint32_t lxo_utilhash(uint32_t a);
void lxo_rng_draws(int iter, int index, int depth, uint32_t seed, int n, float* out, uint32_t* states);
void lxo_camera_ray(const lx_camera_desc* cam, int iter, int x, int y, int trace_depth, uint32_t seed,
                    float* dir3);
float lxo_ct_pdf(const float* wo, const float* wi, const float* n, float roughness, float metallic,
                 const float* color, float ior, float transmission);
void lxo_ct_brdf(const float* v, const float* n, const float* l, const float* albedo, float roughness,
                 float metallic, float ior, float transmission, float* out3);
void lxo_ct_btdf(const float* wo, const float* n, const float* wi, const float* albedo, float roughness,
                 float metallic, float ior, float transmission, float* out3);
float lxo_ct_fresnel_dielectric(float cos_i, float eta_i, float eta_t);
/* Legacy (UBER_SHADER 0) shaders: the glass shader's Fresnel; one sample of
 * class `type` from RNG(iter, index, depth) (returns the u01 draws used);
 * the throughput factor and last_pdf of update_throughput_materials. */
float lxo_glass_fresnel(float cos_i, float ior);
int lxo_legacy_sample(int type, const float* dir, const float* n, float ior, int iter, int index, int depth,
                      uint32_t seed, float* dir3);
float lxo_legacy_update(int type, const float* dir, const float* n, const float* wi, const float* color, float ior,
                        float transmission, float* thr3);
float lxo_lambert_pdf(const float* wi, const float* n);
float lxo_ct_d(const float* h, const float* n, float alpha);
float lxo_ct_smith_g1(const float* w, const float* n, float alpha);
/* Sample the uber-shader lobe from RNG(iter, index, depth); returns the
 * number of u01 draws consumed. */
int lxo_ct_sample(const float* wo, const float* n, float roughness, float metallic, const float* color, float ior,
                  float transmission, int iter, int index, int depth, uint32_t seed, float* dir3);
int lxo_select_from_cdf(const float* cdf, int size, float xi);
void lxo_texture_sample(const lx_texture_desc* tex, float u, float v, float* out4);
float lxo_envmap_pdf(const lxo_scene* s, const float* d);
void lxo_sample_env(const lxo_scene* s, int iter, int index, int depth, uint32_t seed, int dmis, float* dir3);
void lxo_sample_light(const lxo_scene* s, int iter, int index, int depth, uint32_t seed, float* p3);

/* Batched forms for Monte-Carlo and quadrature checks.  lxo_ct_eval_batch:
 * what 0 = PDF (1 float per direction), 1 = BTDF, 2 = BRDF (3 floats) at
 * n_dirs directions wi.  The samplers draw RNG(iter, index0 + i, depth). */
void lxo_ct_eval_batch(int what, const float* wo, const float* wi, uint32_t n_dirs, const float* n, float roughness,
                       float metallic, const float* color, float ior, float transmission, float* out);
void lxo_ct_sample_batch(const float* wo, const float* n, float roughness, float metallic, const float* color,
                         float ior, float transmission, int iter, int depth, uint32_t seed, int index0, uint32_t count,
                         float* dirs);
void lxo_envmap_pdf_batch(const lxo_scene* s, const float* dirs, uint32_t count, float* out);
void lxo_sample_env_batch(const lxo_scene* s, int iter, int depth, uint32_t seed, int index0, uint32_t count,
                          float* dirs);
void lxo_sample_light_batch(const lxo_scene* s, int iter, int depth, uint32_t seed, int index0, uint32_t count,
                            float* pts);
/* get_normal (material_queries.cu:33-89) of material `mat` for a hit with
 * normal N, tangent T and uv. */
void lxo_get_normal(const lxo_scene* s, uint32_t mat, const float* N, const float* T, const float* uv, float* out3);

/* Deterministic math, exported for the accuracy sweep. */
void lxo_math_eval(int fn, const float* x, const float* y, float* out, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
