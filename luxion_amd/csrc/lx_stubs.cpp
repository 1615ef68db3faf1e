/*
 * lx_stubs.cpp -- entry points whose implementation is not in this build yet.
 */
#include "lx_internal.h"

bool lx_build_blas_gpu(lx_context*, const float*, const uint32_t*, uint32_t, int, int, int, BvhBuildResult&) {
    return false;
}

extern "C" {
lx_status lx_scene_load_glb(lx_context*, const char*, const char*, lx_scene**) {
    lx_set_error("glTF loading is not available in this build");
    return LX_ERR_UNSUPPORTED;
}
}
