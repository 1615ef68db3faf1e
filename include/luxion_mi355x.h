/*
 * luxion_mi355x.h -- C ABI of the MI355X-native luxion wavefront path tracer.
 *
 * The reference renderer has no plugin API; its de-facto boundary is the
 * Scene constructor, pathtraceInit / pathtrace / pathtraceFree /
 * copyImageToHost and the -dmis / -emis options.  Each entry point below
 * names the reference interface it replaces (paths relative to the
 * reference repository root).
 *
 * Conventions
 *  - Every call returns lx_status; lx_last_error() gives a message for the
 *    calling thread.  Nothing calls exit().
 *  - Handles are opaque.  A context owns one HIP device and one stream; a
 *    context and everything created from it is used from one host thread at a
 *    time (the served form runs its own request thread internally).
 *  - Scene creation copies every caller array: the caller may free its
 *    buffers when the call returns.
 *  - All host arrays are plain C arrays; no torch or HIP types cross the ABI.
 */
#ifndef LUXION_MI355X_H
#define LUXION_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LX_ABI_VERSION 4

typedef enum lx_status {
    LX_OK = 0,
    LX_ERR_INVALID = 1,     /* bad argument or malformed descriptor */
    LX_ERR_OOM = 2,         /* device or host allocation failed */
    LX_ERR_HIP = 3,         /* a HIP runtime call failed */
    LX_ERR_IO = 4,          /* file could not be read */
    LX_ERR_FORMAT = 5,      /* file content is malformed or unsupported */
    LX_ERR_BUDGET = 6,      /* the context's HBM budget would be exceeded */
    LX_ERR_CANCELLED = 7,   /* job cancelled before completion */
    LX_ERR_TIMEOUT = 8,     /* job deadline expired */
    LX_ERR_UNSUPPORTED = 9, /* feature not available in this build */
    LX_ERR_STATE = 10,      /* call not valid in the object's current state */
    LX_ERR_NOT_FOUND = 11   /* unknown job id or missing item */
} lx_status;

typedef struct lx_context lx_context;
typedef struct lx_scene lx_scene;
typedef struct lx_render lx_render;
typedef struct lx_server lx_server;

/* Message for the last failed call on this thread ("" if none). */
const char* lx_last_error(void);
int lx_abi_version(void);

/* ---------------------------------------------------------------- context */
/* Replaces the implicit CUDA device-0 setup of the reference
 * (src/app/gl_resources.cpp:130-155, initCuda).  hbm_budget_bytes == 0 means
 * "all free device memory". */
lx_status lx_context_create(int device_ordinal, uint64_t hbm_budget_bytes, lx_context** out);
void lx_context_destroy(lx_context* ctx);
/* Bytes currently charged against the context budget. */
uint64_t lx_context_bytes_in_use(const lx_context* ctx);

/* ------------------------------------------------------------ scene input */
enum { LX_TEX_RGBA8 = 0, LX_TEX_RGBA16 = 1, LX_TEX_RGBA32F = 2 };
enum { LX_ALPHA_OPAQUE = 0, LX_ALPHA_MASK = 1, LX_ALPHA_BLEND = 2 };

/* KHR_texture_transform: uv * scale, rotate by `rotation` if non-zero, + offset
 * (src/material_queries.cu:3-17; sceneStructs.h TextureTransform). */
typedef struct lx_tex_transform {
    float offset[2];
    float scale[2];
    float rotation;
} lx_tex_transform;

/* One glTF-style indexed triangle mesh (one glTF primitive). Positions,
 * normals and uvs share the index buffer, as the reference's glTF loader uses
 * one index list for all three (src/scene/scene_load_gltf.cpp:427). */
typedef struct lx_mesh_desc {
    uint32_t num_verts;
    const float* positions; /* 3 * num_verts */
    const float* normals;   /* 3 * num_verts, or NULL */
    const float* uvs;       /* 2 * num_verts, or NULL */
    uint32_t num_tris;
    const uint32_t* indices; /* 3 * num_tris */
} lx_mesh_desc;

/* One placed mesh: the reference's Geom (sceneStructs.h Geom; the OptiX IAS
 * instance built at src/main.cpp:239-270).  transform is the object-to-world
 * affine map as a row-major 3x4 matrix. */
typedef struct lx_instance_desc {
    uint32_t mesh;
    uint32_t material;
    float transform[12];
} lx_instance_desc;

/* Legacy material classes (sceneStructs.h MaterialType, used when the
 * reference is compiled with UBER_SHADER 0).  Only LX_SHADING_LEGACY renders
 * read material_type; the uber shader treats every material as microfacet.
 * lx_scene_load_glb assigns it with the reference loader's rule
 * (src/scene/scene_load_gltf.cpp:211-228, include/scene/gltf_material_utils.h):
 * emissive, else glass, else microfacet. */
enum {
    LX_MATERIAL_MICROFACET = 0, /* Cook-Torrance uber shader */
    LX_MATERIAL_DIFFUSE = 1,    /* Lambert (src/shaders/lambert.cu) */
    LX_MATERIAL_SPECULAR = 2,   /* perfect mirror (src/shaders/specular.cu) */
    LX_MATERIAL_EMISSIVE = 3,   /* light: emits, Lambert-sampled */
    LX_MATERIAL_GLASS = 4       /* smooth dielectric (src/shaders/glass.cu) */
};

/* The reference's uber-shader Material (sceneStructs.h Material). Texture
 * indices are -1 when absent. */
typedef struct lx_material_desc {
    float base_color[3];
    int32_t albedo_tex;
    lx_tex_transform albedo_xf;
    float roughness;
    float metallic;
    int32_t metallic_roughness_tex;
    lx_tex_transform metallic_roughness_xf;
    int32_t normal_tex;
    lx_tex_transform normal_xf;
    float emission_color[3];
    float emission_strength;
    int32_t emissive_tex;
    lx_tex_transform emissive_xf;
    float ior;
    float transmission;
    float alpha;
    int32_t alpha_mode;
    float alpha_cutoff;
    int32_t material_type; /* LX_MATERIAL_*, read by LX_SHADING_LEGACY only */
} lx_material_desc;

/* Texel data, row-major, 4 channels.  RGBA8 / RGBA16 decode to float as
 * value / 255 or value / 65535 (src/scene/scene_load_gltf.cpp:118-145). */
typedef struct lx_texture_desc {
    uint32_t width;
    uint32_t height;
    int32_t format;
    const void* data;
} lx_texture_desc;

/* Pinhole camera with the reference's Camera fields (sceneStructs.h Camera;
 * ray construction src/kernels/generate_ray_from_camera.cu:26-35). */
typedef struct lx_camera_desc {
    int32_t width;
    int32_t height;
    float position[3];
    float view[3];
    float up[3];
    float right[3];
    float pixel_length[2];
} lx_camera_desc;

typedef struct lx_scene_desc {
    const lx_mesh_desc* meshes;
    uint32_t num_meshes;
    const lx_instance_desc* instances;
    uint32_t num_instances;
    const lx_material_desc* materials;
    uint32_t num_materials;
    const lx_texture_desc* textures;
    uint32_t num_textures;
    const lx_texture_desc* envmap; /* RGBA32F equirect, or NULL */
    lx_camera_desc camera;
    int32_t trace_depth; /* reference default 8 (scene_load_gltf.cpp:475) */
} lx_scene_desc;

/* Replaces Scene::Scene(std::string, const char*) for procedural input
 * (src/scene/scene.cpp:22-49) plus the emitter / HDRI CDF precompute
 * (src/scene/scene_precompute.cpp:22-144). */
lx_status lx_scene_create(lx_context* ctx, const lx_scene_desc* desc, lx_scene** out);

/* Replaces Scene::Scene for .glb input (src/scene/scene.cpp:22-49 ->
 * src/scene/scene_load_gltf.cpp:61-529): glTF binary + optional EXR
 * environment map (NULL or "" for none).  Textures: PNG and baseline JPEG;
 * EXR: scanline NONE / RLE / ZIPS / ZIP / PIZ. */
lx_status lx_scene_load_glb(lx_context* ctx, const char* glb_path, const char* exr_path, lx_scene** out);
/* Test hook, no device needed: the descriptor lx_scene_load_glb would pass to
 * lx_scene_create.  Its arrays stay valid until the next call on the same
 * thread.  env_rgba_head receives the first head_n floats of the EXR. */
lx_status lx_debug_parse_glb(const char* glb_path, const char* exr_path, lx_scene_desc* out,
                             float* env_rgba_head, uint32_t head_n);

enum { LX_CDF_FP64 = 0, LX_CDF_FP32_REFERENCE = 1 };
typedef struct lx_build_opts {
    int32_t bvh_width;      /* 4 or 8; 0 = the context tuning's bvh_width */
    int32_t leaf_size;      /* max triangles per leaf, 1..8 (default 4) */
    int32_t builder;        /* 0 = GPU (top-down binned SAH on the device, the host builder's
                               split and leaf rules), 1 = host binned SAH, 2 = auto (default):
                               GPU for BLASes of 2^14 triangles or more, host otherwise */
    int32_t sah_bins;       /* binned-SAH bins (default 16) */
    int32_t cdf_precision;  /* LX_CDF_FP64 (default) or LX_CDF_FP32_REFERENCE */
} lx_build_opts;

/* Replaces the per-mesh GAS build + compaction and the IAS build
 * (src/myoptix/optix_accel.cpp:9-110, src/myoptix/optix_ias.cpp:9-63). */
lx_status lx_scene_build_accel(lx_scene* scene, const lx_build_opts* opts);
lx_status lx_scene_get_camera(const lx_scene* scene, lx_camera_desc* out, int32_t* trace_depth);

typedef struct lx_scene_info {
    uint32_t num_meshes, num_instances, num_materials, num_textures;
    uint64_t num_unique_tris;   /* sum over meshes */
    uint64_t num_instanced_tris; /* sum over instances */
    uint32_t num_emissive_instances;
    float total_emissive_area;
    int32_t env_width, env_height;
    uint64_t device_bytes;
    uint32_t bvh_nodes;
    int32_t bvh_depth;
    double build_ms;
    double sah_cost;
    char name[128]; /* glTF scene name ("GLTF_DEFAULT_NAME" when unnamed), or "" */
} lx_scene_info;
lx_status lx_scene_get_info(const lx_scene* scene, lx_scene_info* out);

/* Read back the precomputed CDFs (for parity checks).  Any pointer may be
 * NULL; sizes in elements. */
lx_status lx_scene_read_env_cdfs(const lx_scene* scene, float* marginal, float* conditional);
lx_status lx_scene_read_emitter_cdfs(const lx_scene* scene, uint32_t* emissive_instances,
                                     float* instance_prefix, uint32_t* num_emissive);
lx_status lx_scene_read_instance_tri_cdf(const lx_scene* scene, uint32_t emissive_index,
                                         float* tri_prefix, uint32_t* num_tris);
void lx_scene_destroy(lx_scene* scene);

/* ----------------------------------------------------------- render jobs */
enum { LX_TONEMAP_REINHARD = 0, LX_TONEMAP_AGX = 1, LX_TONEMAP_ACES = 2 };
/* Material debug views (src/material_debug_render.cu:3-26; names from
 * src/app/render_imgui.cpp:336).  A view traces one bounce (src/pathtrace.cu:107),
 * writes the queried value as the pixel colour, and lx_render_read_ldr shows
 * it without tonemap or gamma (src/kernels/image_kernels.cu:20). */
enum {
    LX_DEBUG_OFF = 0,
    LX_DEBUG_ALBEDO = 1,
    LX_DEBUG_WORLD_NORMAL = 2, /* shading normal * 0.5 + 0.5 */
    LX_DEBUG_NORMAL_MAP = 3,   /* raw normal-map texel, (0.5, 0.5, 1) without one */
    LX_DEBUG_ROUGHNESS = 4,
    LX_DEBUG_METALLIC = 5,
    LX_DEBUG_ALPHA = 6
};
/* LX_SHADING_UBER: every material is the Cook-Torrance uber shader (the
 * reference's default build, include/config.h:23).  LX_SHADING_LEGACY: the
 * per-material_type shaders of the UBER_SHADER 0 build
 * (src/sample_materials.cu, src/update_throughput_materials.cu,
 * src/kernels/shade_path.cu:44-92). */
enum { LX_SHADING_UBER = 0, LX_SHADING_LEGACY = 1 };

/* PathTracerOptions (include/pathtrace/pathtracer_options.h:3-21) plus the
 * CLI flags -dmis / -emis (src/main.cpp:117-124). */
typedef struct lx_render_opts {
    int32_t max_depth;        /* 0 = scene trace_depth */
    int32_t dmis;             /* direct-light (emitter) NEE + MIS */
    int32_t emis;             /* environment importance sampling + MIS; wins over dmis */
    float envmap_intensity;   /* default 1 */
    uint32_t seed;            /* 0 = the reference's RNG streams */
    int32_t color_mode;       /* LX_TONEMAP_* for lx_render_read_ldr */
    const lx_camera_desc* camera; /* NULL = scene camera */
    int32_t material_debug_mode; /* LX_DEBUG_* (PathTracerOptions::material_debug_mode) */
    int32_t shading_model;       /* LX_SHADING_* */
} lx_render_opts;

/* Replaces pathtraceInit(Scene*) (src/pathtrace/pathtrace_buffers.cu:21-140). */
lx_status lx_render_create(lx_scene* scene, const lx_render_opts* opts, lx_render** out);
/* Replaces n_iters calls of pathtrace(pbo, frame, iter) with iter running
 * first_iter .. first_iter + n_iters - 1 (1-based, src/app/render_loop.cpp:131;
 * bounce loop src/pathtrace.cu:97-367).  Asynchronous on the context stream;
 * lx_render_sync waits. */
lx_status lx_render_iterate(lx_render* r, uint32_t first_iter, uint32_t n_iters);
lx_status lx_render_sync(lx_render* r);
/* Replaces copyImageToHost (pathtrace_buffers.cu:180-190): the per-pixel SUM
 * of radiance (3 floats per pixel, row-major) and the spp it holds. */
lx_status lx_render_read_accum(lx_render* r, float* rgb_sum, uint32_t* spp);
/* Display path: average, tonemap (Reinhard / AgX / ACES), gamma 0.45, RGBA8
 * (sendImageToPBO, src/kernels/image_kernels.cu:6-46). */
lx_status lx_render_read_ldr(lx_render* r, uint8_t* rgba, int32_t tonemap);
lx_status lx_render_reset(lx_render* r);

typedef struct lx_render_stats {
    uint64_t radiance_rays;   /* closest-hit rays traced */
    uint64_t dl_shadow_rays;  /* emitter shadow rays traced (closest hit) */
    uint64_t env_shadow_rays; /* environment occlusion rays traced (any hit) */
    uint64_t path_bounces;    /* paths shaded, summed over bounces */
    uint32_t iterations;      /* spp accumulated */
    double trace_ms;          /* device time in trace kernels (event-timed) */
    double shade_ms;          /* device time in shade kernels */
    double other_ms;          /* camera / setup kernels */
    uint32_t trace_launches;  /* kernel launches counted in trace_ms */
    uint32_t shade_launches;
    /* filled when the render runs with LX_TIMING_COUNT_WORK: BVH nodes,
     * triangles and instances visited by closest-hit (radiance + emitter
     * shadow) and any-hit (environment shadow) rays */
    uint64_t closest_nodes, closest_tris, closest_insts;
    uint64_t any_nodes, any_tris, any_insts;
    /* the tail kernel's share of the above (bounces at and after
     * lx_tuning.wavefront_bounces), its rays, and its event-timed device time */
    uint64_t tail_closest_rays, tail_any_rays;
    uint64_t tail_closest_nodes, tail_closest_tris, tail_closest_insts;
    uint64_t tail_any_nodes, tail_any_tris, tail_any_insts;
    double tail_ms;
    uint32_t tail_launches;
    uint64_t tail_bounces; /* path-bounces shaded by the tail kernel */
} lx_render_stats;
/* Counters since the last reset. */
lx_status lx_render_stats_get(lx_render* r, lx_render_stats* out);
/* LX_TIMING_EVENTS: launch eagerly with a HIP event between consecutive
 * kernels (per-kernel device time in the stats).  LX_TIMING_COUNT_WORK: run
 * the counting variant of the trace kernel (same results, plus visit counts).
 * 0 restores graph replay. */
enum { LX_TIMING_EVENTS = 1, LX_TIMING_COUNT_WORK = 2 };
lx_status lx_render_set_timing(lx_render* r, int32_t flags);
/* saveImage (src/app/save_image.cpp:43-84, src/image.cpp:23-43): average the
 * accumulation, host ACES, gamma 0.45, clamp * 255 truncated, mirrored
 * horizontally, written as an 8-bit RGB PNG at `path` (extension included). */
lx_status lx_render_save_png(lx_render* r, const char* path);
/* Test hook, no device: the same output path over a host accumulation
 * buffer (w * h * 3 radiance sums, index x + y * w). */
lx_status lx_debug_write_png(const char* path, const float* accum, uint32_t w, uint32_t h, uint32_t spp);
void lx_render_destroy(lx_render* r);

/* -------------------------------------------------------------- test hooks */
typedef struct lx_ray {
    float origin[3];
    float tmin;        /* unused by the triangle test (t > 2^-23 always) */
    float dir[3];
    float tmax;
} lx_ray;

typedef struct lx_hit {
    float t;           /* -1 on a miss */
    uint32_t instance; /* 0xffffffff on a miss */
    uint32_t prim;
    float u, v;        /* barycentric weights of vertices 1 and 2 */
} lx_hit;

/* Closest hit (lexicographic minimum of (t, instance, prim) over every
 * triangle the Moller-Trumbore test of src/intersections/mesh_isect.cu:77-137
 * accepts) for n host rays.  any_hit_alpha applies the alpha any-hit rules of
 * src/optixshaders/optix_triangle.cu:317-366 with the given iteration. */
lx_status lx_trace_closest(lx_scene* scene, const lx_ray* rays, uint32_t n, int32_t any_hit_alpha,
                           uint32_t iteration, lx_hit* hits);
/* Any hit: occluded[i] = 1 if some accepted triangle has t < tmax. */
lx_status lx_trace_occluded(lx_scene* scene, const lx_ray* rays, uint32_t n, int32_t any_hit_alpha,
                            uint32_t iteration, uint8_t* occluded);
/* No counterpart in the reference, whose direct-light ray is the closest-hit
 * ray of lx_trace_closest (src/kernels/sample_direct_light.cu:9-66,
 * src/optixshaders/optix_triangle.cu:259-305): the query the tail kernel
 * answers that ray with for -dmis jobs.  hits[i] is lx_trace_closest's hit
 * (alpha any-hit rules on, iteration as given) when the hit instance's
 * material emits (emission colour x strength non-zero in a channel), else a
 * miss; a test hook for the query's parity with the closest-hit ray. */
lx_status lx_trace_emitting(lx_scene* scene, const lx_ray* rays, uint32_t n, uint32_t iteration, lx_hit* hits);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lx_trace_closest on DEVICE arrays (d_rays: n lx_ray,
 * d_hits: n lx_hit, both allocations of this process on the scene's device),
 * with no hipMalloc / hipFree and no allocating library call.  The traversal
 * stack's spill columns and its overflow flag live in a fixed scratch array
 * of 32 MiB + 256 B per device (held until the process ends), so n is at
 * most 131072 rays per call (larger n: LX_ERR_INVALID).  Runs on the
 * context's stream and returns once the hits are written.  Every call on a
 * device shares that device's array: calls that overlap on one device (from
 * different contexts or threads) race; calls through one context are
 * serialised. */
lx_status lx_trace_closest_noalloc(lx_scene* scene, const lx_ray* d_rays, uint32_t n, int32_t any_hit_alpha,
                                   uint32_t iteration, lx_hit* d_hits);

/* Run one bounce-resolved iteration and dump, per bounce d = 1..depth, the
 * sorted pixel indices that survive compaction (survivors_out, with counts in
 * survivor_counts[d-1]) plus the radiance hit record of every shaded path
 * (records_out: 8 floats per shaded path in shade order: t, nx, ny, nz, u, v,
 * material (as float), pixel (as float)).  Buffers hold max_depth * W * H
 * entries.  The render's accumulation buffer is updated as by
 * lx_render_iterate. */
lx_status lx_debug_bounce(lx_render* r, uint32_t iter, uint32_t* survivors_out, uint32_t* survivor_counts,
                          float* records_out, uint32_t* record_counts);

/* Evaluate one deterministic math function of include/luxion_detmath.h on
 * the device (fn: 0 sin, 1 cos, 2 atan, 3 atan2(x, y), 4 acos, 5 log2,
 * 6 exp2, 7 pow(x, y), 8 sqrt, 9 x / y) for n host values, to prove the
 * GPU and host sequences agree bit for bit. */
lx_status lx_debug_math_eval(lx_context* ctx, int32_t fn, const float* x, const float* y, float* out, uint32_t n);

/* Plain device-to-device copy kernel of `bytes` bytes run `iters` times on
 * the context stream (16-byte loads / stores); returns the event-timed copy
 * bandwidth (read + write bytes per second / 1e9).  Used to calibrate HBM
 * traffic counters and as the achievable-bandwidth reference. */
lx_status lx_debug_copy_bandwidth(lx_context* ctx, uint64_t bytes, uint32_t iters, double* gbps);

/* ----------------------------------------------------------- batched form */
/* One render job of the batched / served forms.  All jobs of a batch share
 * one persistent launch per bounce. */
typedef struct lx_job {
    lx_scene* scene;
    lx_camera_desc camera;
    int32_t spp;
    int32_t max_depth;
    int32_t dmis, emis;
    float envmap_intensity;
    uint32_t seed;        /* per-job RNG stream; 0 = reference stream */
    uint32_t first_iter;  /* 1-based iteration of the first sample */
    double deadline_ms;   /* served form: relative deadline, 0 = none */
    int32_t material_debug_mode; /* LX_DEBUG_* */
    int32_t shading_model;       /* LX_SHADING_* */
} lx_job;

typedef struct lx_job_result {
    lx_status status;
    float* rgb_sum;       /* caller-owned, 3 * W * H floats, or NULL */
    uint32_t spp_done;
    uint64_t rays;
    double submit_ms, start_ms, finish_ms; /* context clock */
} lx_job_result;

typedef struct lx_tuning {
    int32_t block_size;    /* threads per workgroup of the tail kernel: 128 or 256 */
    int32_t bvh_width;     /* 4 or 8: scenes built with bvh_width 0 and servers */
    int32_t queue_chunk;   /* tail kernel: a wave claims paths from the queue once this
                              many of its lanes (1..64) are idle */
    int32_t waves_per_cu;  /* persistent grid size */
    int32_t use_graph;     /* capture the bounce loop in a hipGraph */
    uint32_t max_paths_per_round; /* batched form path capacity, 0 = from the budget */
    int32_t wavefront_bounces; /* bounces run as one launch each; the rest run in the
                                  tail kernel, one lane per path to its end (same bits);
                                  negative = every bounce its own launch */
    int32_t tail_min_waves;    /* 0, or 4: the tail kernel is compiled for at least
                                  4 waves per SIMD (fewer registers, some spills) */
    int32_t sort_mode;         /* LX_SORT_* bits: queue-reordering stages (same bits) */
    int32_t pixel_tile;        /* camera paths enter the queue tile x tile pixel
                                  square by square (1, 8, 16 or 32; 1 = row-major;
                                  same bits) */
    int32_t tile_band;         /* squares follow each other in bands of this many
                                  square rows, column by column (0 = row by row;
                                  0..64; same bits) */
    int32_t overlap_rounds;    /* 1..3 (default 1): lx_render_iterate of two or more
                                  samples without timing keeps this many rounds in
                                  flight besides the current one, each on a stream
                                  of its own, so one round's last long paths
                                  overlap the next rounds; each round's colours go
                                  to a buffer of their own, added to the
                                  accumulation in round order (same bits).  The
                                  batched form overlaps pairs of rounds and a
                                  server four rounds whenever this is nonzero.
                                  0: one round after another on the render's
                                  stream (captured in a hipGraph with use_graph) */
    int32_t round_frames;      /* 1..32 (default 1): lx_render_iterate of two or more
                                  samples (without work counting; event timing
                                  brackets each round) renders up to this many
                                  consecutive samples in one round (one job copy
                                  per sample, each into a colour buffer of its own,
                                  added to the accumulation in sample order; same
                                  bits), so a round's last long paths are shared
                                  by that many frames; with overlap_rounds, rounds
                                  of this width overlap.  Path buffers grow by
                                  this factor; a render whose HBM budget cannot
                                  hold them renders as with 1.  lx_batch_run
                                  takes up to this many samples of every job per
                                  round (at most 2048 job copies) when its
                                  budget holds them (same bits) */
} lx_tuning;

/* Coherence stages (the reference's RAY_SORTING and MATERIAL_SORTING,
 * src/pathtrace.cu:157-185 and 266-276, include/config.h:3-4): LX_SORT_RAYS
 * orders the live-path queue by a Morton code of each ray before every traced
 * bounce (the per-bounce traces and the tail kernel's first one);
 * LX_SORT_MATERIALS orders the per-bounce shading by hit material.  Both only
 * permute the queue, so the rendered bits do not change. */
enum { LX_SORT_NONE = 0, LX_SORT_RAYS = 1, LX_SORT_MATERIALS = 2 };

lx_status lx_tuning_get(const lx_context* ctx, lx_tuning* out);
lx_status lx_tuning_set(lx_context* ctx, const lx_tuning* t);
/* Autotuner: sweeps the launch knobs (per-bounce / tail split, occupancy,
 * block size, grid size, queue granularity, sort stages and, for scenes it
 * may rebuild, BVH width) on the given job set, keeps the fastest, and writes
 * it as JSON to json_path (if non-NULL). */
lx_status lx_autotune(lx_context* ctx, const lx_job* jobs, uint32_t n, const char* json_path, lx_tuning* best);
lx_status lx_tuning_load(lx_context* ctx, const char* json_path);

/* Render n jobs to completion, all sharing each bounce launch.  No
 * counterpart in the reference (its only batch is a serial CLI loop,
 * run_best_scenes.ps1:22-50). */
lx_status lx_batch_run(lx_context* ctx, const lx_job* jobs, uint32_t n, lx_job_result* results);

/* ------------------------------------------------------------ served form */
typedef struct lx_server_opts {
    uint64_t hbm_budget_bytes; /* admission budget for per-job buffers, 0 = context budget */
    int32_t max_jobs_in_flight;
    int32_t spp_per_round;     /* samples each admitted job advances per round (1) */
    uint32_t max_paths_per_round; /* 0 = sized from the budget */
    int32_t max_depth;         /* deepest job accepted; bounces per round (0 = 16) */
} lx_server_opts;

lx_status lx_server_start(lx_context* ctx, const lx_server_opts* opts, lx_server** out);
/* Queue a job; returns its id.  result->rgb_sum must stay valid until the job
 * is reported done by lx_server_poll. */
lx_status lx_server_submit(lx_server* s, const lx_job* job, lx_job_result* result, uint64_t* job_id);
/* done = 1 once the job finished (status in the result). timeout_ms < 0 waits. */
lx_status lx_server_poll(lx_server* s, uint64_t job_id, double timeout_ms, int32_t* done);
lx_status lx_server_cancel(lx_server* s, uint64_t job_id);
typedef struct lx_server_stats {
    uint64_t jobs_done, jobs_failed, jobs_rejected, rounds;
    uint64_t rays;
    double busy_ms;
    double p50_ms, p99_ms;
    uint64_t peak_bytes;
    uint32_t peak_jobs_in_flight;
} lx_server_stats;
lx_status lx_server_stats_get(lx_server* s, lx_server_stats* out);
/* Test hook: fail the next n device-buffer admissions as if the budget were
 * exhausted. */
lx_status lx_server_inject_budget_failure(lx_server* s, uint32_t n);
lx_status lx_server_stop(lx_server* s);

/* Context clock in milliseconds (monotonic). */
double lx_now_ms(void);

#ifdef __cplusplus
}
#endif

#endif /* LUXION_MI355X_H */
