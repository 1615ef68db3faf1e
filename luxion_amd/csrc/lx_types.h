/*
 * lx_types.h -- device-visible data layout of the MI355X wavefront path
 * tracer (shared by host code and the gfx950 kernels).
 *
 * HBM layout (one scene):
 *   nodes   : LxNode4[]  all BLAS trees followed by the TLAS, 128 B per node
 *   tris    : float4[3n] leaf-ordered triangles (v0 | prim, e1, e2), 48 B each
 *   inst    : LxInstance[] world<->object maps, BLAS root, attribute offsets
 *   idx/pos/nrm/uv : per-mesh attribute arrays for hit attributes
 *   mats / texs / texels : uber-shader materials and texel pools
 *   env + marginal/conditional CDFs, emitter CDFs
 * Path state is struct-of-arrays (LxPathBuf), one slot per live path,
 * compacted every bounce so every kernel reads and writes coalesced.
 */
#ifndef LX_TYPES_H
#define LX_TYPES_H

#include <stdint.h>
#include "../../include/luxion_mi355x.h"

#include <hip/hip_runtime.h>

/* BVH4 node: four child boxes stored per axis so one lane reads the node as
 * seven 16-byte loads.  child[k]: node index (bit 31 clear), leaf
 * (bit 31 set: count-1 in bits 28..30, first triangle / instance in bits
 * 0..27) or LX_CHILD_EMPTY. */
#define LX_CHILD_EMPTY 0xffffffffu
#define LX_LEAF_BIT 0x80000000u
#define LX_LEAF_FIRST_MASK 0x0fffffffu
#define LX_SENTINEL 0x7ffffffeu /* stack marker: return from an instance */

struct alignas(128) LxNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 child;
    int4 meta; /* meta.x: 1 for TLAS nodes */
};

/* BVH8 node (256 B): the same scheme with eight children; each bound is
 * eight consecutive floats (children 0-3, then 4-7). */
struct alignas(256) LxNode8 {
    float4 lox[2], hix[2], loy[2], hiy[2], loz[2], hiz[2];
    int4 child[2];
    int4 meta[2];
};
static_assert(sizeof(LxNode4) == 128 && sizeof(LxNode8) == 256, "node layout");

/* Builder output node, up to eight children (host and GPU builders); packed
 * into LxNode4 or LxNode8 by the scene once every tree's indices are final.
 * Child boxes are already padded. */
struct LxWideNode {
    float lo[8][3], hi[8][3];
    uint32_t child[8]; /* node index, leaf ref or LX_CHILD_EMPTY */
    uint32_t n;
};

struct alignas(16) LxInstance {
    float inv[12]; /* world -> object, row-major 3x4 */
    float m[12];   /* object -> world */
    uint32_t mesh, material, blas_root, flags;
    uint32_t idx_off;  /* first index (x3) of the mesh in the index pool */
    uint32_t vtx_off;  /* first vertex of the mesh in the attribute pools */
    uint32_t num_tris;
    uint32_t has_n_uv; /* bit 0: normals, bit 1: uvs */
};
#define LX_INST_ALPHA 1u    /* material needs the any-hit alpha test */
#define LX_INST_EMISSIVE 2u /* material emits: emission colour x strength not all zero */

/* A TLAS leaf: instances that the traversal enters together.  Instances with
 * one bitwise-identical transform whose meshes have no other user share one
 * BLAS over all their triangles (a tagged group: each triangle record carries
 * its instance in e1.w); any other instance is a group of its own over its
 * mesh's BLAS. */
struct alignas(16) LxGroup {
    float inv[12];      /* world -> object of every member */
    uint32_t blas_root;
    uint32_t flags;     /* LX_INST_ALPHA if any member needs the any-hit alpha test;
                           LX_GROUP_EMIS_ANY / _ALL if some / every member emits */
    uint32_t inst;      /* the member instance, or LX_GROUP_TAGGED */
    uint32_t pad;
};
#define LX_GROUP_TAGGED 0xffffffffu
#define LX_GROUP_EMIS_ANY 2u
#define LX_GROUP_EMIS_ALL 4u

struct LxTexDesc {
    const void* data;
    int32_t w, h, fmt, pad;
};

struct LxDevScene {
    const void* nodes; /* LxNode4* or LxNode8* */
    const float4* tris;
    const LxInstance* inst;
    const LxGroup* groups; /* TLAS leaves */
    const uint32_t* idx;
    const float* pos;
    const float* nrm;
    const float* uv;
    const lx_material_desc* mats;
    const LxTexDesc* texs;
    LxTexDesc env;
    const float* marginal;
    const float* conditional;
    const uint32_t* emissive_inst;
    const float* emissive_prefix;
    const float* tri_cdf;
    const uint32_t* tri_cdf_off;
    uint32_t tlas_root;
    uint32_t emis_tlas_root; /* TLAS over the groups with an emitting member (LX_CHILD_EMPTY: none) */
    uint32_t num_inst;
    uint32_t num_emissive;
    float total_area;
    int32_t has_env;
    int32_t bvh_width;
    /* ray sort keys: world box of all instances and a reach along the ray */
    float box_lo[3], box_inv[3];
    float sort_reach;
    uint32_t num_groups;
};

/* One render job (single renders are a batch of one). */
struct LxJob {
    const LxDevScene* scene; /* device address */
    lx_camera_desc cam;
    uint32_t path_off;  /* first path slot of this job in a round */
    uint32_t num_pix;   /* W * H */
    uint32_t accum_off; /* first pixel of this job in the accumulation pool */
    int32_t depth;
    int32_t dmis, emis;
    float intensity;
    uint32_t seed;
    int32_t iter;   /* current 1-based iteration */
    int32_t active; /* takes part in this round */
    int32_t spp_done;
    uint32_t flags; /* LX_JF_*: material debug view, legacy shading */
};
#define LX_JF_DEBUG_MASK 7u /* LX_DEBUG_* view, 0 = off */
#define LX_JF_LEGACY 8u     /* LX_SHADING_LEGACY */
/* bits 8..15: the job's rounds alternate with this many copies of it in all
 * (rounds in flight); each of its rounds then advances iter by that step */
#define LX_JF_STEP_SHIFT 8u
#define LX_JF_STEP_MASK 0xff00u

/* Path state (struct of arrays, capacity N). */
struct LxPathBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tx, *ty, *tz;
    float *cx, *cy, *cz;
    float* pdf;
    uint32_t* pix;  /* pixel index within the job; bit 31 = last bounce specular */
    uint32_t* job;
};

/* Trace -> shade records (struct of arrays, capacity N). */
struct LxRecBuf {
    float *t, *nx, *ny, *nz, *u, *v;
    int32_t* mat;
    float *tgx, *tgy, *tgz;
    float* env_t;
    float *edx, *edy, *edz;
    float *dl_t, *dl_nx, *dl_ny, *dl_nz, *dl_u, *dl_v;
    int32_t* dl_mat;
    float *lpx, *lpy, *lpz;
};

/* Per-round counters in device memory. */
#define LX_MAX_DEPTH 64
/* counting replay: rays and node / triangle / instance visits of closest-hit
 * (radiance + direct-light) and any-hit (environment) rays of one kernel */
struct LxWorkCnt {
    unsigned long long closest_rays, any_rays;
    unsigned long long closest_nodes, closest_tris, closest_insts, any_nodes, any_tris, any_insts;
};
struct LxCounters {
    uint32_t count[LX_MAX_DEPTH + 1]; /* live paths entering bounce d */
    uint32_t tail_next;               /* tail kernel's queue head */
    unsigned long long rad_rays, dl_rays, env_rays, bounces;
    unsigned long long tail_bounces; /* path-bounces shaded by the tail kernel */
    LxWorkCnt work[2]; /* [0] wavefront trace kernel, [1] tail kernel */
    uint32_t stack_overflow;
    uint32_t pad[3];
};

#endif
