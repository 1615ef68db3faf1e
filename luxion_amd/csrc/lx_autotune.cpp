/*
 * lx_autotune.cpp -- launch-shape autotuner and its JSON record.
 *
 * The reference fixes its launch shapes at compile time (128-thread blocks,
 * src/pathtrace.cu:197-300).  Here the knobs live in lx_tuning and every
 * choice renders the same bits (per-path RNG streams, one accumulation per
 * pixel, tree-independent closest hits), so the tuner may pick whatever is
 * fastest on the caller's own job mix: it times lx_batch_run over the jobs
 * for each candidate and keeps the best, one knob at a time (coordinate
 * descent, two passes).  BVH width is a knob only when every job's scene can
 * be rebuilt (no other renders or jobs hold it); the scenes end up built at
 * the chosen width.
 */
#include "lx_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <set>
#include <sstream>

namespace {

double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Knob {
    const char* name;
    int32_t lx_tuning::*field;
    std::vector<int32_t> values;
};

std::vector<Knob> knobs() {
    return {
        {"wavefront_bounces", &lx_tuning::wavefront_bounces, {0, 1, 2, -1}},
        {"tail_min_waves", &lx_tuning::tail_min_waves, {0, 4}},
        {"block_size", &lx_tuning::block_size, {128, 256}},
        {"waves_per_cu", &lx_tuning::waves_per_cu, {8, 12, 16, 24}},
        {"queue_chunk", &lx_tuning::queue_chunk, {1, 8, 32}},
        {"sort_mode", &lx_tuning::sort_mode, {0, 1, 2, 3}},
        {"pixel_tile", &lx_tuning::pixel_tile, {1, 8, 16}},
        {"tile_band", &lx_tuning::tile_band, {0, 4, 8}},
    };
}

const char* kFields[] = {"block_size", "bvh_width", "queue_chunk", "waves_per_cu", "use_graph",
                         "max_paths_per_round", "wavefront_bounces", "tail_min_waves", "sort_mode",
                         "pixel_tile", "tile_band", "overlap_rounds", "round_frames"};

int32_t* field_ptr(lx_tuning& t, const std::string& k) {
    if (k == "block_size") return &t.block_size;
    if (k == "bvh_width") return &t.bvh_width;
    if (k == "queue_chunk") return &t.queue_chunk;
    if (k == "waves_per_cu") return &t.waves_per_cu;
    if (k == "use_graph") return &t.use_graph;
    if (k == "max_paths_per_round") return (int32_t*)&t.max_paths_per_round;
    if (k == "wavefront_bounces") return &t.wavefront_bounces;
    if (k == "tail_min_waves") return &t.tail_min_waves;
    if (k == "sort_mode") return &t.sort_mode;
    if (k == "pixel_tile") return &t.pixel_tile;
    if (k == "tile_band") return &t.tile_band;
    if (k == "overlap_rounds") return &t.overlap_rounds;
    if (k == "round_frames") return &t.round_frames;
    return nullptr;
}

std::string to_json(const lx_tuning& t, double ms) {
    std::ostringstream o;
    lx_tuning c = t;
    o << "{";
    for (size_t i = 0; i < sizeof(kFields) / sizeof(kFields[0]); i++) {
        int32_t v = *field_ptr(c, kFields[i]);
        if (std::string(kFields[i]) == "max_paths_per_round") o << "\"" << kFields[i] << "\": " << (uint32_t)v;
        else o << "\"" << kFields[i] << "\": " << v;
        o << ", ";
    }
    o << "\"measured_ms\": " << ms << "}\n";
    return o.str();
}

/* Flat JSON object of integer (or number) members; unknown keys ignored. */
void parse_tuning_json(const std::string& text, lx_tuning& t) {
    size_t p = text.find('{');
    LX_REQUIRE(p != std::string::npos, LX_ERR_INVALID, "tuning file: expected a JSON object");
    p++;
    while (true) {
        while (p < text.size() && isspace((unsigned char)text[p])) p++;
        LX_REQUIRE(p < text.size(), LX_ERR_INVALID, "tuning file: unterminated object");
        if (text[p] == '}') break;
        if (text[p] == ',') {
            p++;
            continue;
        }
        LX_REQUIRE(text[p] == '"', LX_ERR_INVALID, "tuning file: expected a key");
        size_t e = text.find('"', p + 1);
        LX_REQUIRE(e != std::string::npos, LX_ERR_INVALID, "tuning file: unterminated key");
        std::string key = text.substr(p + 1, e - p - 1);
        p = text.find(':', e);
        LX_REQUIRE(p != std::string::npos, LX_ERR_INVALID, "tuning file: expected ':'");
        p++;
        while (p < text.size() && isspace((unsigned char)text[p])) p++;
        size_t q = p;
        while (q < text.size() && (isdigit((unsigned char)text[q]) || text[q] == '-' || text[q] == '+' ||
                                   text[q] == '.' || text[q] == 'e' || text[q] == 'E'))
            q++;
        LX_REQUIRE(q > p, LX_ERR_INVALID, "tuning file: values must be numbers");
        double v = std::stod(text.substr(p, q - p));
        p = q;
        int32_t* f = field_ptr(t, key);
        if (f) *f = key == "max_paths_per_round" ? (int32_t)(uint32_t)v : (int32_t)v;
    }
}

void check_tuning(const lx_tuning& t) {
    LX_REQUIRE(t.waves_per_cu >= 1 && t.waves_per_cu <= 32, LX_ERR_INVALID, "waves_per_cu out of range");
    LX_REQUIRE(t.bvh_width == 4 || t.bvh_width == 8, LX_ERR_INVALID, "bvh_width must be 4 or 8");
    LX_REQUIRE(t.block_size == 128 || t.block_size == 256, LX_ERR_INVALID, "block_size must be 128 or 256");
    LX_REQUIRE(t.tail_min_waves == 0 || t.tail_min_waves == 4, LX_ERR_INVALID, "tail_min_waves must be 0 or 4");
    LX_REQUIRE(t.queue_chunk >= 1 && t.queue_chunk <= 64, LX_ERR_INVALID, "queue_chunk must be 1..64");
    LX_REQUIRE(t.sort_mode >= 0 && t.sort_mode <= (LX_SORT_RAYS | LX_SORT_MATERIALS), LX_ERR_INVALID,
               "sort_mode must be a combination of LX_SORT_RAYS and LX_SORT_MATERIALS");
    LX_REQUIRE(t.pixel_tile == 1 || t.pixel_tile == 8 || t.pixel_tile == 16 || t.pixel_tile == 32, LX_ERR_INVALID,
               "pixel_tile must be 1, 8, 16 or 32");
    LX_REQUIRE(t.tile_band >= 0 && t.tile_band <= 64, LX_ERR_INVALID, "tile_band must be 0..64");
    LX_REQUIRE(t.overlap_rounds >= 0 && t.overlap_rounds <= 3, LX_ERR_INVALID, "overlap_rounds must be 0..3");
    LX_REQUIRE(t.round_frames >= 1 && t.round_frames <= 32, LX_ERR_INVALID, "round_frames must be 1..32");
}

} // namespace

void lx_tuning_check(const lx_tuning& t) { check_tuning(t); }

extern "C" {

lx_status lx_autotune(lx_context* ctx, const lx_job* jobs, uint32_t n, const char* json_path, lx_tuning* best_out) {
    LX_TRY
    LX_REQUIRE(ctx && jobs && n > 0, LX_ERR_INVALID, "null argument");
    for (uint32_t i = 0; i < n; i++) lx_validate_job(jobs[i]);
    std::set<lx_scene*> scenes;
    for (uint32_t i = 0; i < n; i++) scenes.insert(jobs[i].scene);
    bool can_rebuild = true;
    for (lx_scene* s : scenes) can_rebuild &= s->refs.load() == 1;
    std::vector<lx_job_result> res(n);
    auto set_tuning = [&](const lx_tuning& t) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->tuning = t;
    };
    auto measure = [&](const lx_tuning& t) {
        set_tuning(t);
        double best = 1e300;
        for (int rep = 0; rep < 2; rep++) {
            for (auto& r : res) std::memset(&r, 0, sizeof(r));
            double t0 = wall_ms();
            lx_status st = lx_batch_run(ctx, jobs, n, res.data());
            if (st != LX_OK) throw LxError{st, std::string("autotune: batch failed: ") + lx_last_error()};
            best = std::min(best, wall_ms() - t0);
        }
        return best;
    };
    auto rebuild = [&](int width) {
        for (lx_scene* s : scenes) {
            if (s->dev.bvh_width == width) continue;
            lx_build_opts o = s->build_opts;
            o.bvh_width = width;
            std::lock_guard<std::mutex> lk(ctx->mu);
            lx_scene_accel(s, &o);
        }
    };
    lx_tuning start;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        start = ctx->tuning;
    }
    lx_tuning cur = start;
    cur.bvh_width = jobs[0].scene->dev.bvh_width;
    /* a candidate replaces the current choice only if it is at least 1 %
     * faster: closer shapes differ by less than the run-to-run noise */
    constexpr double kSwitch = 0.99;
    try {
        if (can_rebuild) rebuild(cur.bvh_width);
        double cur_ms = measure(cur);
        for (int pass = 0; pass < 2; pass++) {
            if (can_rebuild) {
                for (int w : {4, 8}) {
                    if (w == cur.bvh_width) continue;
                    rebuild(w);
                    lx_tuning c = cur;
                    c.bvh_width = w;
                    double ms = measure(c);
                    if (ms < cur_ms * kSwitch) {
                        cur_ms = ms;
                        cur = c;
                    }
                }
                rebuild(cur.bvh_width);
            }
            for (const Knob& k : knobs()) {
                for (int32_t v : k.values) {
                    if (cur.*(k.field) == v) continue;
                    lx_tuning c = cur;
                    c.*(k.field) = v;
                    double ms = measure(c);
                    if (ms < cur_ms * kSwitch) {
                        cur_ms = ms;
                        cur = c;
                    }
                }
            }
        }
        set_tuning(cur);
        if (best_out) *best_out = cur;
        if (json_path) {
            std::ofstream f(json_path);
            LX_REQUIRE(f.good(), LX_ERR_IO, std::string("cannot write ") + json_path);
            f << to_json(cur, cur_ms);
        }
    } catch (...) {
        set_tuning(start);
        throw;
    }
    LX_CATCH
}

lx_status lx_tuning_load(lx_context* ctx, const char* json_path) {
    LX_TRY
    LX_REQUIRE(ctx && json_path, LX_ERR_INVALID, "null argument");
    std::ifstream f(json_path);
    LX_REQUIRE(f.good(), LX_ERR_IO, std::string("cannot read ") + json_path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::lock_guard<std::mutex> lk(ctx->mu);
    lx_tuning t = ctx->tuning;
    parse_tuning_json(ss.str(), t);
    check_tuning(t);
    ctx->tuning = t;
    LX_CATCH
}
}
