/*
 * lx_kernels.h -- host-side launch entry points of the gfx950 kernels.
 */
#ifndef LX_KERNELS_H
#define LX_KERNELS_H

#include <hip/hip_runtime.h>
#include "lx_types.h"

#ifndef LX_TRACE_BLOCK
#define LX_TRACE_BLOCK 128
#endif
#ifndef LX_SHADE_BLOCK
#define LX_SHADE_BLOCK 256
#endif
#define LX_SCHED_MAX 2048
#ifndef LX_STACK
#define LX_STACK 24 /* per-lane LDS stack entries */
#endif
#define LX_SPILL 64 /* per-lane global spill entries beyond LX_STACK */
/* lx_trace_closest_noalloc: at most this many rays per call (one lane per
 * ray, a multiple of LX_TRACE_BLOCK); its scratch array holds their spill
 * columns and one 256-byte line for the stack-overflow flag (32 MiB + 256 B) */
#define LX_NOALLOC_MAX_RAYS 131072u
#define LX_NOALLOC_SPILL_WORDS ((size_t)LX_SPILL * LX_NOALLOC_MAX_RAYS)
#define LX_NOALLOC_SCRATCH_WORDS (LX_NOALLOC_SPILL_WORDS + 64)

struct LxSched {
    uint32_t n_active;
    uint32_t pad[3];
    uint32_t idx[LX_SCHED_MAX];
    uint32_t off[LX_SCHED_MAX];
};

hipError_t lx_launch_begin_round(LxJob* jobs, const int32_t* job_spp, uint32_t njobs, LxCounters* C, LxSched* sched,
                                 uint32_t cap, hipStream_t s);
hipError_t lx_launch_end_round(LxJob* jobs, uint32_t njobs, hipStream_t s);
/* tile: camera paths of a job cover tile x tile pixel squares in turn (1 =
 * row-major); band: squares in bands of this many square rows, walked
 * column by column (0 = row by row) */
hipError_t lx_launch_camera(const LxJob* jobs, const LxSched* sched, const LxCounters* C, const LxPathBuf& P,
                            uint32_t grid, hipStream_t s, uint32_t tile = 1, uint32_t band = 0);
/* width: the BVH width (4 or 8) of every scene the jobs reference */
/* order: NULL, or the queue permutation from lx_launch_sort (thread q takes
 * the path in slot order[q]) */
hipError_t lx_launch_trace(const LxJob* jobs, const LxPathBuf& P, const LxRecBuf& R, LxCounters* C, int bounce,
                           uint32_t* spill, uint32_t grid, int width, hipStream_t s, bool count = false,
                           const uint32_t* order = nullptr);
hipError_t lx_launch_shade(const LxJob* jobs, const LxPathBuf& P, const LxPathBuf& Q, const LxRecBuf& R, LxCounters* C,
                           int bounce, float* accum, uint32_t grid, hipStream_t s, const uint32_t* order = nullptr);
/* Queue permutation of the count[bounce] live paths by ray cell
 * (LX_SORT_RAYS) or hit material (LX_SORT_MATERIALS) into order;
 * keys: path capacity entries, hist: LX_SORT_HIST entries. */
#define LX_SORT_HIST 4097
hipError_t lx_launch_sort(const LxJob* jobs, const LxPathBuf& P, const LxRecBuf& R, const LxCounters* C, int bounce,
                          int mode, uint32_t* keys, uint32_t* hist, uint32_t* order, uint32_t grid, hipStream_t s);
/* threads: total lanes (a multiple of 128); block: 128 or 256; occ: 0 or 4
 * minimum waves per SIMD; chunk: a wave refills once this many of its lanes
 * (1..64) are idle */
hipError_t lx_launch_finish(const LxJob* jobs, const LxPathBuf& P, LxCounters* C, int bounce0, uint32_t* spill,
                            float* accum, uint32_t threads, int block, int occ, int width, int chunk,
                            const LxDevScene* uscene /* the scene of every job, or NULL */,
                            bool single_job /* njobs == 1 */, hipStream_t s, bool count = false,
                            const uint32_t* order = nullptr,
                            /* non-NULL (bounce0 0, no order): the kernel makes the
                             * camera paths of the scheduled jobs itself, as the
                             * camera kernel would with this tile and band, instead
                             * of reading them from P */
                            const LxSched* cam_sched = nullptr, uint32_t tile = 1, uint32_t band = 0,
                            /* some job has -dmis: the kernel whose direct-light rays
                             * are occlusion queries (same colours) */
                            bool dl_query = false,
                            /* accum was filled with -0.0 for this round and gets one
                             * path per pixel: colours are stored, not added */
                            bool colour_fresh = false);
hipError_t lx_launch_trace_rays(const LxDevScene* S, const lx_ray* rays, uint32_t n, int any, int alpha,
                                uint32_t iteration, lx_hit* hits, uint8_t* occ, uint32_t* spill, uint32_t* overflow,
                                uint32_t grid, int width, hipStream_t s);
/* Device address of the allocation-free trace's scratch array on the current
 * device (ordinal `device`, which must be current); looked up once per device
 * with hipGetSymbolAddress and cached. */
hipError_t lx_noalloc_scratch_address(int device, uint32_t** out);
/* World-space bounds of transformed vertex sets: item k maps the vertices
 * [first, first + count) of pos (xyz float triples) by its object -> world
 * 3x4 matrix, in double, and writes min xyz, max xyz (double) to out[6k..]. */
struct LxBoundsItem {
    float m[12];
    uint32_t first, count;
};
hipError_t lx_launch_inst_bounds(const float* pos, const LxBoundsItem* items, uint32_t n, double* out, hipStream_t s);
hipError_t lx_launch_tonemap(const float* accum, uint32_t npix, float spp, int mode, uint8_t* out, hipStream_t s);
hipError_t lx_launch_math_eval(int fn, const float* x, const float* y, float* out, uint32_t n, hipStream_t s);
hipError_t lx_launch_copy16(const void* src, void* dst, size_t n16, uint32_t grid, hipStream_t s);
hipError_t lx_launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
hipError_t lx_launch_add_f32(float* dst, const float* src, size_t n, hipStream_t s);
/* dst[i] += src[0 * n + i], then += src[1 * n + i], ... in region order */
hipError_t lx_launch_add_regions(float* dst, const float* src, size_t n, uint32_t regions, hipStream_t s);
/* ranges: nranges pairs (first pixel, pixel count) of an RGB float pool, in device memory
 * (nranges <= 65535) */
hipError_t lx_launch_fill_ranges(uint32_t* p, const uint32_t* ranges, uint32_t nranges, uint32_t v, hipStream_t s);
hipError_t lx_launch_add_ranges(float* dst, const float* src, const uint32_t* ranges, uint32_t nranges, hipStream_t s);

#endif
