/*
 * lx_stubs.cpp -- entry points whose implementation is not in this build yet.
 */
#include "lx_internal.h"

bool lx_build_blas_gpu(lx_context*, const float*, const uint32_t*, uint32_t, int, int, int, BvhBuildResult&) {
    return false;
}

