// A C++ caller of include/luxion_mi355x.h, written the way the reference's
// render loop would call the library (src/app/render_loop.cpp:83-148,
// src/main.cpp:71-286).  The block between the INTEGRATION markers is the
// C++ example of INTEGRATION.md, verbatim (tests/test_capi_cpu.py checks
// that the two stay identical).
//
//   lx_capi_consumer SCENE.glb ITERATIONS OUT.bin [-dmis]
// loads the scene, renders ITERATIONS 1-spp iterations, and writes
// width, height, spp (3 x int32) and the radiance sums (3 floats per pixel).
#include "luxion_mi355x.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

// INTEGRATION-BEGIN
static lx_context* g_ctx; static lx_scene* g_scene; static lx_render* g_render;

static void check(lx_status s) {
    if (s != LX_OK) { fprintf(stderr, "luxion: %s\n", lx_last_error()); std::exit(1); }
}

void initRenderer(const char* scene_file, const char* envmap, bool dmis, bool emis) {
    check(lx_context_create(/*device*/0, /*hbm_budget*/0, &g_ctx));
    check(lx_scene_load_glb(g_ctx, scene_file, envmap, &g_scene));
    lx_build_opts b = {4, 4, 2 /* auto: GPU builder for big meshes */, 16, LX_CDF_FP64};
    check(lx_scene_build_accel(g_scene, &b));
    lx_render_opts o = {};
    o.dmis = dmis; o.emis = emis; o.envmap_intensity = 1.0f; o.color_mode = LX_TONEMAP_ACES;
    check(lx_render_create(g_scene, &o, &g_render));
}

void runFrame(int iteration /* 1-based, as app.iteration */) {
    check(lx_render_iterate(g_render, (uint32_t)iteration, 1));
}

void saveImage(std::vector<float>& rgb_sum, uint32_t& spp) {
    check(lx_render_read_accum(g_render, rgb_sum.data(), &spp));   // sum; divide by spp
}
// INTEGRATION-END

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s SCENE.glb ITERATIONS OUT.bin [-dmis]\n", argv[0]);
        return 2;
    }
    bool dmis = argc > 4 && std::strcmp(argv[4], "-dmis") == 0;
    int iters = std::atoi(argv[2]);
    initRenderer(argv[1], nullptr, dmis, false);
    for (int it = 1; it <= iters; ++it) runFrame(it);
    lx_camera_desc cam;
    int32_t depth = 0;
    check(lx_scene_get_camera(g_scene, &cam, &depth));
    std::vector<float> rgb((size_t)cam.width * cam.height * 3);
    uint32_t spp = 0;
    saveImage(rgb, spp);
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 3;
    int32_t head[3] = {cam.width, cam.height, (int32_t)spp};
    std::fwrite(head, sizeof(head), 1, f);
    std::fwrite(rgb.data(), sizeof(float), rgb.size(), f);
    std::fclose(f);
    lx_render_destroy(g_render);
    lx_scene_destroy(g_scene);
    lx_context_destroy(g_ctx);
    std::printf("rendered %dx%d, %u spp\n", cam.width, cam.height, spp);
    return 0;
}
