/*
 * lx_scene.cpp -- scene ingest, emitter / HDRI CDF precompute, device upload
 * and acceleration-structure assembly.
 *
 * Replaces Scene::Scene + Scene::precompute_emissive_mesh_area +
 * Scene::precompute_hdri_emission (src/scene/scene.cpp:22-49,
 * src/scene/scene_precompute.cpp:22-144), Mesh::make_mesh_device
 * (src/mesh/mesh_device.cu:11-45) and the GAS / IAS builds
 * (src/myoptix/optix_accel.cpp:9-110, src/myoptix/optix_ias.cpp:9-63).
 *
 * The CDF prefix sums run in FP64 by default (normalised values rounded to
 * fp32 once); LX_CDF_FP32_REFERENCE reproduces the reference's fp32 sums.
 */
#include "lx_internal.h"

#include <algorithm>
#include <array>
#include <map>
#include <chrono>
#include <cmath>
#include <limits>
#include <cfloat>

#include "../../include/luxion_detmath.h"

namespace {

struct H3 {
    float x, y, z;
};
static inline H3 h3(float x, float y, float z) { return H3{x, y, z}; }
static inline H3 hsub(H3 a, H3 b) { return h3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline H3 hcross(H3 x, H3 y) {
    return h3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
static inline H3 hxf(const float* m, H3 p) {
    return h3((m[0] * p.x + m[1] * p.y) + (m[2] * p.z + m[3]), (m[4] * p.x + m[5] * p.y) + (m[6] * p.z + m[7]),
              (m[8] * p.x + m[9] * p.y) + (m[10] * p.z + m[11]));
}

/* Whether a material can emit: lx_get_emission is emission colour x
 * strength (times a texture), so a material whose product is zero in every
 * channel adds exactly zero wherever it is seen. */
static bool lx_material_emits(const lx_material_desc& m) {
    for (int c = 0; c < 3; c++) {
        float e = m.emission_color[c] * m.emission_strength;
        if (e != 0.0f || e != e) return true;
    }
    return false;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

/* Affine inverse in double, rounded once to float (the reference inverts in
 * fp32 with glm::inverse, src/scene/scene_load_gltf.cpp:436). */
void lx_affine_inverse(const float* m, float* out) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    LX_REQUIRE(det != 0.0 && std::isfinite(det), LX_ERR_INVALID, "instance transform is singular");
    double inv_det = 1.0 / det;
    double r[9] = {A * inv_det,
                   -(b * i - c * h) * inv_det,
                   (b * f - c * e) * inv_det,
                   B * inv_det,
                   (a * i - c * g) * inv_det,
                   -(a * f - c * d) * inv_det,
                   C * inv_det,
                   -(a * h - b * g) * inv_det,
                   (a * e - b * d) * inv_det};
    double tx = m[3], ty = m[7], tz = m[11];
    for (int row = 0; row < 3; row++) {
        out[row * 4 + 0] = (float)r[row * 3 + 0];
        out[row * 4 + 1] = (float)r[row * 3 + 1];
        out[row * 4 + 2] = (float)r[row * 3 + 2];
        out[row * 4 + 3] = (float)(-(r[row * 3 + 0] * tx + r[row * 3 + 1] * ty + r[row * 3 + 2] * tz));
    }
}

static size_t texel_bytes(int fmt) { return fmt == LX_TEX_RGBA8 ? 4 : (fmt == LX_TEX_RGBA16 ? 8 : 16); }

static void copy_tex(const lx_texture_desc& d, HostTex& t) {
    LX_REQUIRE(d.width > 0 && d.height > 0 && d.data, LX_ERR_INVALID, "texture without data");
    LX_REQUIRE(d.format >= LX_TEX_RGBA8 && d.format <= LX_TEX_RGBA32F, LX_ERR_INVALID, "unknown texture format");
    t.w = d.width;
    t.h = d.height;
    t.fmt = d.format;
    size_t n = (size_t)d.width * d.height * texel_bytes(d.format);
    t.data.assign((const uint8_t*)d.data, (const uint8_t*)d.data + n);
}

static float texel_f(const HostTex& t, size_t k, int c) {
    if (t.fmt == LX_TEX_RGBA8) return ((float)t.data[4 * k + c]) / 255.0f;
    if (t.fmt == LX_TEX_RGBA16) {
        uint16_t v;
        std::memcpy(&v, &t.data[8 * k + 2 * c], 2);
        return ((float)v) / 65535.0f;
    }
    float v;
    std::memcpy(&v, &t.data[16 * k + 4 * c], 4);
    return v;
}

/* precompute_emissive_mesh_area, scene_precompute.cpp:22-74 */
static void precompute_emitters(lx_scene* s) {
    s->emissive.clear();
    s->emissive_prefix.clear();
    s->tri_cdf.clear();
    s->tri_cdf_off.assign(1, 0u);
    float area_f = 0.0f;
    double area_d = 0.0;
    std::vector<double> pref_d;
    for (uint32_t gi = 0; gi < s->insts.size(); gi++) {
        const lx_instance_desc& g = s->insts[gi];
        const lx_material_desc& mat = s->mats[g.material];
        H3 ec = h3(mat.emission_color[0], mat.emission_color[1], mat.emission_color[2]);
        if (!(lxm_sqrt(hdot(ec, ec)) * mat.emission_strength > 1e-5f)) continue;
        const HostMesh& m = s->meshes[g.mesh];
        size_t nt = m.idx.size() / 3;
        size_t base = s->tri_cdf.size();
        s->tri_cdf.resize(base + nt);
        float cum_f = 0.0f;
        double cum_d = 0.0;
        std::vector<double> cd(nt);
        for (size_t t = 0; t < nt; t++) {
            const float* P = m.pos.data();
            uint32_t i0 = m.idx[3 * t], i1 = m.idx[3 * t + 1], i2 = m.idx[3 * t + 2];
            H3 A = hxf(g.transform, h3(P[3 * i0], P[3 * i0 + 1], P[3 * i0 + 2]));
            H3 B = hxf(g.transform, h3(P[3 * i1], P[3 * i1 + 1], P[3 * i1 + 2]));
            H3 C = hxf(g.transform, h3(P[3 * i2], P[3 * i2 + 1], P[3 * i2 + 2]));
            H3 cr = hcross(hsub(B, A), hsub(C, A));
            float a = 0.5f * lxm_sqrt(hdot(cr, cr));
            cum_f += a;
            cum_d += (double)a;
            s->tri_cdf[base + t] = cum_f;
            cd[t] = cum_d;
        }
        if (s->cdf_mode == LX_CDF_FP32_REFERENCE) {
            for (size_t t = 0; t < nt; t++) s->tri_cdf[base + t] /= cum_f;
            area_f += cum_f;
            s->emissive_prefix.push_back(area_f);
        } else {
            for (size_t t = 0; t < nt; t++) s->tri_cdf[base + t] = (float)(cd[t] / cum_d);
            area_d += cum_d;
            pref_d.push_back(area_d);
        }
        s->emissive.push_back(gi);
        s->tri_cdf_off.push_back((uint32_t)s->tri_cdf.size());
    }
    if (s->cdf_mode == LX_CDF_FP32_REFERENCE) {
        for (auto& p : s->emissive_prefix) p /= area_f;
        s->total_area = area_f;
    } else {
        for (double p : pref_d) s->emissive_prefix.push_back((float)(p / area_d));
        s->total_area = (float)area_d;
    }
}

/* precompute_hdri_emission, scene_precompute.cpp:76-144 */
static void precompute_env(lx_scene* s) {
    int W = (int)s->env.w, H = (int)s->env.h;
    const float PIF = 3.1415926535897932384626422832795028841971f;
    s->conditional.assign((size_t)W * H, 0.0f);
    s->marginal.assign(H, 0.0f);
    float total_f = 0.0f;
    double total_d = 0.0;
    std::vector<double> marg_d(H), row_d(W);
    for (int y = 0; y < H; y++) {
        float row_f = 0.0f;
        double rs = 0.0;
        float theta = PIF * (y + 0.5f) / H;
        float st = lxm_sin(theta);
        for (int x = 0; x < W; x++) {
            size_t k = (size_t)y * W + x;
            float r = texel_f(s->env, k, 0), g = texel_f(s->env, k, 1), b = texel_f(s->env, k, 2),
                  a = texel_f(s->env, k, 3);
            float e = lxm_sqrt((r * r + g * g) + (b * b + a * a));
            float w = e * st;
            row_f += w;
            rs += (double)w;
            s->conditional[k] = row_f;
            row_d[x] = rs;
        }
        float* row = &s->conditional[(size_t)y * W];
        if (s->cdf_mode == LX_CDF_FP32_REFERENCE) {
            if (row_f > 0.0f)
                for (int x = 0; x < W; x++) row[x] /= row_f;
            else
                for (int x = 0; x < W; x++) row[x] = (float)(x + 1) / (float)W;
            row[W - 1] = 1.0f;
            total_f += row_f;
            s->marginal[y] = total_f;
        } else {
            if (rs > 0.0)
                for (int x = 0; x < W; x++) row[x] = (float)(row_d[x] / rs);
            else
                for (int x = 0; x < W; x++) row[x] = (float)(x + 1) / (float)W;
            row[W - 1] = 1.0f;
            total_d += rs;
            marg_d[y] = total_d;
        }
    }
    if (s->cdf_mode == LX_CDF_FP32_REFERENCE) {
        LX_REQUIRE(total_f != 0.0f, LX_ERR_INVALID, "environment map has zero emission");
        for (int y = 0; y < H; y++) s->marginal[y] /= total_f;
    } else {
        LX_REQUIRE(total_d != 0.0, LX_ERR_INVALID, "environment map has zero emission");
        for (int y = 0; y < H; y++) s->marginal[y] = (float)(marg_d[y] / total_d);
    }
    s->marginal[H - 1] = 1.0f;
}

static void upload_textures(lx_scene* s) {
    lx_context* ctx = s->ctx;
    std::vector<LxTexDesc> td(s->texs.size());
    s->d_tex.resize(s->texs.size());
    for (size_t i = 0; i < s->texs.size(); i++) {
        s->d_tex[i].upload(ctx, s->texs[i].data.data(), s->texs[i].data.size());
        td[i].data = s->d_tex[i].ptr;
        td[i].w = (int)s->texs[i].w;
        td[i].h = (int)s->texs[i].h;
        td[i].fmt = s->texs[i].fmt;
    }
    if (!td.empty()) s->d_texdesc.upload(ctx, td.data(), td.size() * sizeof(LxTexDesc));
    if (s->has_env) {
        s->d_env.upload(ctx, s->env.data.data(), s->env.data.size());
        s->dev.env.data = s->d_env.ptr;
        s->dev.env.w = (int)s->env.w;
        s->dev.env.h = (int)s->env.h;
        s->dev.env.fmt = s->env.fmt;
        s->d_marg.upload(ctx, s->marginal.data(), s->marginal.size() * 4);
        s->d_cond.upload(ctx, s->conditional.data(), s->conditional.size() * 4);
    }
}

void lx_scene_init(lx_scene* s, const lx_scene_desc* d) {
    LX_REQUIRE(d != nullptr, LX_ERR_INVALID, "null scene descriptor");
    LX_REQUIRE(d->num_meshes > 0 && d->meshes, LX_ERR_INVALID, "scene needs at least one mesh");
    LX_REQUIRE(d->num_instances > 0 && d->instances, LX_ERR_INVALID, "scene needs at least one instance");
    LX_REQUIRE(d->num_materials > 0 && d->materials, LX_ERR_INVALID, "scene needs at least one material");
    s->meshes.resize(d->num_meshes);
    for (uint32_t i = 0; i < d->num_meshes; i++) {
        const lx_mesh_desc& md = d->meshes[i];
        LX_REQUIRE(md.positions && md.indices && md.num_tris > 0 && md.num_verts > 0, LX_ERR_INVALID,
                   "mesh " + std::to_string(i) + " has no geometry");
        HostMesh& m = s->meshes[i];
        m.pos.assign(md.positions, md.positions + 3 * (size_t)md.num_verts);
        m.idx.assign(md.indices, md.indices + 3 * (size_t)md.num_tris);
        for (uint32_t v : m.idx) LX_REQUIRE(v < md.num_verts, LX_ERR_INVALID, "mesh index out of range");
        m.has_n = md.normals != nullptr;
        if (m.has_n) m.nrm.assign(md.normals, md.normals + 3 * (size_t)md.num_verts);
        m.has_uv = md.uvs != nullptr;
        if (m.has_uv) m.uv.assign(md.uvs, md.uvs + 2 * (size_t)md.num_verts);
    }
    s->mats.assign(d->materials, d->materials + d->num_materials);
    s->texs.resize(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; i++) copy_tex(d->textures[i], s->texs[i]);
    for (const auto& m : s->mats) {
        int ts[4] = {m.albedo_tex, m.metallic_roughness_tex, m.normal_tex, m.emissive_tex};
        for (int t : ts) LX_REQUIRE(t < (int)d->num_textures, LX_ERR_INVALID, "material texture index out of range");
    }
    s->insts.assign(d->instances, d->instances + d->num_instances);
    s->inv.resize(12 * s->insts.size());
    for (size_t i = 0; i < s->insts.size(); i++) {
        LX_REQUIRE(s->insts[i].mesh < d->num_meshes, LX_ERR_INVALID, "instance mesh index out of range");
        LX_REQUIRE(s->insts[i].material < d->num_materials, LX_ERR_INVALID, "instance material out of range");
        lx_affine_inverse(s->insts[i].transform, &s->inv[12 * i]);
    }
    s->cam = d->camera;
    s->trace_depth = d->trace_depth > 0 ? d->trace_depth : 8;
    LX_REQUIRE(s->trace_depth <= LX_MAX_DEPTH, LX_ERR_INVALID, "trace depth above the supported maximum");
    s->has_env = d->envmap != nullptr;
    if (s->has_env) {
        LX_REQUIRE(d->envmap->format == LX_TEX_RGBA32F, LX_ERR_INVALID, "environment map must be RGBA32F");
        copy_tex(*d->envmap, s->env);
    }
    s->info.num_meshes = d->num_meshes;
    s->info.num_instances = d->num_instances;
    s->info.num_materials = d->num_materials;
    s->info.num_textures = d->num_textures;
    for (auto& m : s->meshes) s->info.num_unique_tris += m.idx.size() / 3;
    for (auto& g : s->insts) s->info.num_instanced_tris += s->meshes[g.mesh].idx.size() / 3;
}

/* Nodes in[0, n) packed: child boxes in per-axis SoA float4 groups (four
 * children per group), children refs, meta.x = 1 for TLAS nodes, meta.y =
 * the child count.  Empty child slots get inverted boxes (lo = +inf,
 * hi = -inf, missed by every slab test) and LX_CHILD_EMPTY. */
template <class N, int W>
static std::vector<N> pack_nodes(const LxWideNode* in, size_t count, bool tlas) {
    std::vector<N> out(count);
    for (size_t i = 0; i < count; i++) {
        const LxWideNode& w = in[i];
        N& n = out[i];
        std::memset(&n, 0, sizeof(N));
        float* lox = (float*)&n.lox;
        float* hix = (float*)&n.hix;
        float* loy = (float*)&n.loy;
        float* hiy = (float*)&n.hiy;
        float* loz = (float*)&n.loz;
        float* hiz = (float*)&n.hiz;
        uint32_t* ch = (uint32_t*)&n.child;
        for (int k = 0; k < W; k++) {
            bool used = k < 8 && w.child[k] != LX_CHILD_EMPTY;
            const float inf = std::numeric_limits<float>::infinity();
            lox[k] = used ? w.lo[k][0] : inf;
            loy[k] = used ? w.lo[k][1] : inf;
            loz[k] = used ? w.lo[k][2] : inf;
            hix[k] = used ? w.hi[k][0] : -inf;
            hiy[k] = used ? w.hi[k][1] : -inf;
            hiz[k] = used ? w.hi[k][2] : -inf;
            ch[k] = used ? w.child[k] : LX_CHILD_EMPTY;
        }
        int* meta = (int*)&n.meta;
        meta[0] = tlas ? 1 : 0;
        meta[1] = (int)w.n;
    }
    return out;
}

/* A run of BLAS nodes and triangles in the scene's arrays: built on the host
 * (held in the host arrays from host_node / host_tri) or left on the device
 * by the GPU builder (dev >= 0: its result). */
struct BlasSegment {
    uint32_t node_first, node_count, tri_first, tri_count;
    size_t host_node, host_tri;
    int dev;
};

/* Writes the BLAS segments, then the TLAS nodes (packed as type N), into
 * dst_nodes / dst_tris. */
template <class N, int W>
static void upload_bvh(lx_context* ctx, const std::vector<BlasSegment>& segs, const std::vector<LxWideNode>& nodes,
                       const std::vector<float4>& tris, std::vector<BvhBuildResult>& dev, const BvhBuildResult& tl,
                       uint32_t tlas_first, uint32_t tri_total, DevBuf& dst_nodes, DevBuf& dst_tris,
                       std::vector<std::vector<N>>& hold) {
    hipStream_t st = ctx->stream;
    dst_nodes.alloc(ctx, ((size_t)tlas_first + tl.nodes.size()) * sizeof(N));
    dst_tris.alloc(ctx, (size_t)tri_total * 48);
    char* dn = (char*)dst_nodes.ptr;
    char* dt = (char*)dst_tris.ptr;
    for (const BlasSegment& g : segs) {
        char* n_at = dn + (size_t)g.node_first * sizeof(N);
        char* t_at = dt + (size_t)g.tri_first * 48;
        if (g.dev >= 0) {
            const BvhBuildResult& r = dev[(size_t)g.dev];
            if (g.node_count)
                LX_HIP(hipMemcpyAsync(n_at, r.d_nodes.ptr, (size_t)g.node_count * sizeof(N), hipMemcpyDeviceToDevice,
                                      st));
            LX_HIP(hipMemcpyAsync(t_at, r.d_tris.ptr, (size_t)g.tri_count * 48, hipMemcpyDeviceToDevice, st));
        } else {
            if (g.node_count) {
                hold.push_back(pack_nodes<N, W>(nodes.data() + g.host_node, g.node_count, false));
                LX_HIP(hipMemcpyAsync(n_at, hold.back().data(), (size_t)g.node_count * sizeof(N),
                                      hipMemcpyHostToDevice, st));
            }
            LX_HIP(hipMemcpyAsync(t_at, tris.data() + 3 * g.host_tri, (size_t)g.tri_count * 48, hipMemcpyHostToDevice,
                                  st));
        }
    }
    if (!tl.nodes.empty()) {
        hold.push_back(pack_nodes<N, W>(tl.nodes.data(), tl.nodes.size(), true));
        LX_HIP(hipMemcpyAsync(dn + (size_t)tlas_first * sizeof(N), hold.back().data(), tl.nodes.size() * sizeof(N),
                              hipMemcpyHostToDevice, st));
    }
}

/* Precompute and upload everything; builds the BVHs. */
void lx_scene_accel(lx_scene* s, const lx_build_opts* o) {
    lx_build_opts opt{0, 4, 2, 16, LX_CDF_FP64};
    if (o) opt = *o;
    if (opt.bvh_width == 0) opt.bvh_width = s->ctx->tuning.bvh_width;
    LX_REQUIRE(opt.bvh_width == 4 || opt.bvh_width == 8, LX_ERR_INVALID, "bvh_width must be 4 or 8 (0: the tuning's)");
    if (opt.leaf_size <= 0) opt.leaf_size = 4;
    if (opt.sah_bins <= 0) opt.sah_bins = 16;
    lx_context* ctx = s->ctx;
    ctx->set_device();
    double t0 = now_ms();
    s->build_opts = opt;
    s->cdf_mode = opt.cdf_precision;
    if (s->has_env) precompute_env(s);
    precompute_emitters(s);

    /* attribute pools */
    std::vector<uint32_t> idx_off(s->meshes.size()), vtx_off(s->meshes.size());
    std::vector<uint32_t> all_idx;
    std::vector<float> all_pos, all_nrm, all_uv;
    bool any_n = false, any_uv = false;
    for (auto& m : s->meshes) {
        any_n |= m.has_n;
        any_uv |= m.has_uv;
    }
    for (size_t i = 0; i < s->meshes.size(); i++) {
        const HostMesh& m = s->meshes[i];
        idx_off[i] = (uint32_t)all_idx.size();
        vtx_off[i] = (uint32_t)(all_pos.size() / 3);
        all_idx.insert(all_idx.end(), m.idx.begin(), m.idx.end());
        all_pos.insert(all_pos.end(), m.pos.begin(), m.pos.end());
        size_t nv = m.pos.size() / 3;
        if (any_n) {
            if (m.has_n) all_nrm.insert(all_nrm.end(), m.nrm.begin(), m.nrm.end());
            else all_nrm.resize(all_nrm.size() + 3 * nv, 0.0f);
        }
        if (any_uv) {
            if (m.has_uv) all_uv.insert(all_uv.end(), m.uv.begin(), m.uv.end());
            else all_uv.resize(all_uv.size() + 2 * nv, 0.0f);
        }
    }

    /* traversal groups: instances with one bitwise-identical transform (so
     * one world -> object transform of the ray, with the same bits for all of
     * them) whose meshes have no other user share a BLAS over all their
     * triangles, each triangle record tagged with its instance; a triangle
     * test there gives the bits it gives under its own instance.  Any other
     * instance is a group of its own over its mesh's BLAS. */
    std::vector<uint32_t> mesh_users(s->meshes.size(), 0);
    for (const lx_instance_desc& g : s->insts) mesh_users[g.mesh]++;
    std::vector<std::vector<uint32_t>> groups;
    {
        std::map<std::array<uint32_t, 12>, size_t> by_xf;
        for (size_t i = 0; i < s->insts.size(); i++) {
            const lx_instance_desc& g = s->insts[i];
            if (mesh_users[g.mesh] == 1) {
                std::array<uint32_t, 12> key;
                std::memcpy(key.data(), g.transform, 48);
                auto it = by_xf.find(key);
                if (it != by_xf.end()) {
                    groups[it->second].push_back((uint32_t)i);
                    continue;
                }
                by_xf[key] = groups.size();
            }
            groups.push_back({(uint32_t)i});
        }
    }

    std::vector<LxWideNode> nodes; /* host-built BLAS nodes, absolute child refs */
    std::vector<float4> tris;      /* host-built BLAS triangles */
    std::vector<BlasSegment> segs;
    std::vector<BvhBuildResult> dev_blas;
    uint32_t node_total = 0, tri_total = 0;
    int max_depth = 0;
    double sah = 0.0;
    /* Builds a BLAS as the next segment of the scene's node / triangle
     * arrays; returns its root and, in `boxes`, object-space boxes bounding
     * it: the boxes two levels below the root (at most 64), or the box of all
     * its vertices.  A rotated instance's world box is the union of these
     * boxes' transformed corners, much tighter than the transformed box of
     * the whole mesh. */
    auto add_blas = [&](const float* pos, size_t nv, const uint32_t* idx, uint32_t nt, const uint32_t* tag_off,
                        const uint32_t* tag_inst, size_t ntag, std::vector<std::array<float, 6>>& boxes) {
        const uint32_t noff = node_total, toff = tri_total;
        LX_REQUIRE((size_t)toff + nt < (1u << 28) - 1, LX_ERR_INVALID, "scene too large for the leaf encoding");
        BvhBuildResult r;
        bool ok = false;
        bool gpu = opt.builder == 0 || (opt.builder == 2 && nt >= LX_GPU_BUILD_MIN_TRIS);
        if (gpu)
            ok = lx_build_blas_gpu(ctx, pos, nv, idx, nt, opt.leaf_size, opt.sah_bins, opt.bvh_width, noff, toff,
                                   tag_off, tag_inst, ntag, r);
        BlasSegment g{noff, 0, toff, nt, nodes.size(), tris.size() / 3, -1};
        if (ok) {
            g.node_count = r.dev_count;
            g.dev = (int)dev_blas.size();
            boxes = r.top_boxes;
        } else {
            r = BvhBuildResult();
            lx_build_blas_host(pos, idx, nt, opt.leaf_size, opt.sah_bins, opt.bvh_width, r);
            if (ntag) {
                /* group-wide triangle index -> (instance, triangle of its mesh) */
                for (size_t k = 0; k < r.tris.size() / 3; k++) {
                    uint32_t t;
                    std::memcpy(&t, &r.tris[3 * k].w, 4);
                    size_t m = (size_t)(std::upper_bound(tag_off, tag_off + ntag, t) - tag_off) - 1;
                    uint32_t local = t - tag_off[m];
                    std::memcpy(&r.tris[3 * k].w, &local, 4);
                    std::memcpy(&r.tris[3 * k + 1].w, &tag_inst[m], 4);
                }
            }
            g.node_count = (uint32_t)r.nodes.size();
            LX_REQUIRE((size_t)noff + r.nodes.size() < LX_LEAF_BIT, LX_ERR_INVALID,
                       "scene too large for the node encoding");
            for (LxWideNode n : r.nodes) {
                for (int k = 0; k < 8; k++) {
                    uint32_t v = n.child[k];
                    if (v == LX_CHILD_EMPTY) continue;
                    if (v & LX_LEAF_BIT) v = (v & ~LX_LEAF_FIRST_MASK) | ((v & LX_LEAF_FIRST_MASK) + toff);
                    else v += noff;
                    n.child[k] = v;
                }
                nodes.push_back(n);
            }
            if (!(r.root & LX_LEAF_BIT) && r.root != LX_CHILD_EMPTY) {
                const LxWideNode& rn = r.nodes[r.root];
                for (int k = 0; k < 8; k++) {
                    uint32_t c = rn.child[k];
                    if (c == LX_CHILD_EMPTY) continue;
                    if (c & LX_LEAF_BIT) {
                        boxes.push_back(
                            {rn.lo[k][0], rn.lo[k][1], rn.lo[k][2], rn.hi[k][0], rn.hi[k][1], rn.hi[k][2]});
                        continue;
                    }
                    const LxWideNode& cn = r.nodes[c];
                    for (int q = 0; q < 8; q++)
                        if (cn.child[q] != LX_CHILD_EMPTY)
                            boxes.push_back(
                                {cn.lo[q][0], cn.lo[q][1], cn.lo[q][2], cn.hi[q][0], cn.hi[q][1], cn.hi[q][2]});
                }
            }
            tris.insert(tris.end(), r.tris.begin(), r.tris.end());
        }
        uint32_t rr = r.root;
        if (rr & LX_LEAF_BIT) rr = (rr & ~LX_LEAF_FIRST_MASK) | ((rr & LX_LEAF_FIRST_MASK) + toff);
        else rr += noff;
        max_depth = std::max(max_depth, r.depth);
        sah += r.sah;
        segs.push_back(g);
        if (ok) dev_blas.push_back(std::move(r));
        node_total += g.node_count;
        tri_total += nt;
        if (boxes.empty()) {
            std::array<float, 6> b = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (size_t v = 0; v < nv; v++)
                for (int a = 0; a < 3; a++) {
                    b[a] = std::min(b[a], pos[3 * v + a]);
                    b[3 + a] = std::max(b[3 + a], pos[3 * v + a]);
                }
            boxes.push_back(b);
        }
        return rr;
    };

    /* BLAS per mesh of an untagged group (shared by all its instances) */
    const uint32_t no_root = 0xffffffffu;
    std::vector<uint32_t> mesh_root(s->meshes.size(), no_root);
    std::vector<std::vector<std::array<float, 6>>> mesh_boxes(s->meshes.size());
    for (const std::vector<uint32_t>& G : groups) {
        if (G.size() != 1) continue;
        uint32_t mi = s->insts[G[0]].mesh;
        if (mesh_root[mi] != no_root) continue;
        const HostMesh& m = s->meshes[mi];
        mesh_root[mi] = add_blas(m.pos.data(), m.pos.size() / 3, m.idx.data(), (uint32_t)(m.idx.size() / 3), nullptr,
                                 nullptr, 0, mesh_boxes[mi]);
    }

    /* group records, their BLASes and conservative world boxes */
    std::vector<LxGroup> dgroups(groups.size());
    std::vector<uint32_t> inst_root(s->insts.size());
    std::vector<float> wboxes(6 * groups.size());
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const std::vector<uint32_t>& G = groups[gi];
        const lx_instance_desc& g0 = s->insts[G[0]];
        std::vector<std::array<float, 6>> tagged_boxes;
        const std::vector<std::array<float, 6>>* boxes;
        uint32_t root;
        if (G.size() == 1) {
            root = mesh_root[g0.mesh];
            boxes = &mesh_boxes[g0.mesh];
        } else {
            std::vector<float> gpos;
            std::vector<uint32_t> gidx, off, who;
            for (uint32_t ii : G) {
                const HostMesh& m = s->meshes[s->insts[ii].mesh];
                uint32_t base = (uint32_t)(gpos.size() / 3);
                off.push_back((uint32_t)(gidx.size() / 3));
                who.push_back(ii);
                gpos.insert(gpos.end(), m.pos.begin(), m.pos.end());
                for (uint32_t v : m.idx) gidx.push_back(v + base);
            }
            root = add_blas(gpos.data(), gpos.size() / 3, gidx.data(), (uint32_t)(gidx.size() / 3), off.data(),
                            who.data(), off.size(), tagged_boxes);
            boxes = &tagged_boxes;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (const std::array<float, 6>& b : *boxes)
            for (int c = 0; c < 8; c++) {
                double p[3] = {(c & 1) ? b[3] : b[0], (c & 2) ? b[4] : b[1], (c & 4) ? b[5] : b[2]};
                for (int r = 0; r < 3; r++) {
                    double w = (double)g0.transform[4 * r] * p[0] + (double)g0.transform[4 * r + 1] * p[1] +
                               (double)g0.transform[4 * r + 2] * p[2] + (double)g0.transform[4 * r + 3];
                    lo[r] = std::min(lo[r], w);
                    hi[r] = std::max(hi[r], w);
                }
            }
        float flo[3], fhi[3];
        for (int r = 0; r < 3; r++) {
            flo[r] = std::nextafter((float)lo[r], -FLT_MAX);
            fhi[r] = std::nextafter((float)hi[r], FLT_MAX);
        }
        lx_pad_box(flo, fhi);
        for (int r = 0; r < 3; r++) {
            wboxes[6 * gi + r] = flo[r];
            wboxes[6 * gi + 3 + r] = fhi[r];
        }
        LxGroup& D = dgroups[gi];
        std::memcpy(D.inv, &s->inv[12 * (size_t)G[0]], 48);
        D.blas_root = root;
        D.flags = 0u;
        bool all_emit = true;
        for (uint32_t ii : G) {
            if (s->mats[s->insts[ii].material].alpha_mode != LX_ALPHA_OPAQUE) D.flags |= LX_INST_ALPHA;
            if (lx_material_emits(s->mats[s->insts[ii].material])) D.flags |= LX_GROUP_EMIS_ANY;
            else all_emit = false;
            inst_root[ii] = root;
        }
        if (all_emit) D.flags |= LX_GROUP_EMIS_ALL;
        D.inst = G.size() == 1 ? G[0] : LX_GROUP_TAGGED;
        D.pad = 0u;
    }

    /* a group of one rotated or scaled instance: the union of its mesh's top
     * BVH boxes, mapped to the world, is loose (on the instanced interior
     * its surface area is about 1.8 times that of the mapped vertices' box),
     * so the mesh's vertices themselves are mapped (in double) and bounded on
     * the device, then rounded outward and padded as above */
    s->d_pos.upload(ctx, all_pos.data(), all_pos.size() * 4);
    {
        std::vector<LxBoundsItem> items;
        std::vector<size_t> item_group;
        for (size_t gi = 0; gi < groups.size(); gi++) {
            if (groups[gi].size() != 1) continue;
            const lx_instance_desc& g0 = s->insts[groups[gi][0]];
            const size_t nv = s->meshes[g0.mesh].pos.size() / 3;
            if (nv == 0) continue;
            LxBoundsItem it;
            std::memcpy(it.m, g0.transform, 48);
            it.first = vtx_off[g0.mesh];
            it.count = (uint32_t)nv;
            items.push_back(it);
            item_group.push_back(gi);
        }
        if (!items.empty()) {
            DevBuf d_items, d_out;
            d_items.upload(ctx, items.data(), items.size() * sizeof(LxBoundsItem));
            d_out.alloc(ctx, items.size() * 6 * sizeof(double));
            LX_HIP(lx_launch_inst_bounds(s->d_pos.as<float>(), d_items.as<LxBoundsItem>(), (uint32_t)items.size(),
                                         d_out.as<double>(), ctx->stream));
            std::vector<double> hb(items.size() * 6);
            LX_HIP(hipMemcpyAsync(hb.data(), d_out.ptr, hb.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
            LX_HIP(hipStreamSynchronize(ctx->stream));
            for (size_t k = 0; k < items.size(); k++) {
                float flo[3], fhi[3];
                for (int r = 0; r < 3; r++) {
                    flo[r] = std::nextafter((float)hb[6 * k + r], -FLT_MAX);
                    fhi[r] = std::nextafter((float)hb[6 * k + 3 + r], FLT_MAX);
                }
                lx_pad_box(flo, fhi);
                const size_t gi = item_group[k];
                for (int r = 0; r < 3; r++) {
                    wboxes[6 * gi + r] = flo[r];
                    wboxes[6 * gi + 3 + r] = fhi[r];
                }
            }
        }
    }

    /* per-instance records for hit attributes and emitter sampling */
    std::vector<LxInstance> dinst(s->insts.size());
    for (size_t i = 0; i < s->insts.size(); i++) {
        const lx_instance_desc& g = s->insts[i];
        LxInstance& I = dinst[i];
        std::memcpy(I.inv, &s->inv[12 * i], 48);
        std::memcpy(I.m, g.transform, 48);
        I.mesh = g.mesh;
        I.material = g.material;
        I.blas_root = inst_root[i];
        I.flags = (s->mats[g.material].alpha_mode != LX_ALPHA_OPAQUE ? LX_INST_ALPHA : 0u) |
                  (lx_material_emits(s->mats[g.material]) ? LX_INST_EMISSIVE : 0u);
        I.idx_off = idx_off[g.mesh];
        I.vtx_off = vtx_off[g.mesh];
        I.num_tris = (uint32_t)(s->meshes[g.mesh].idx.size() / 3);
        I.has_n_uv = (s->meshes[g.mesh].has_n ? 1u : 0u) | (s->meshes[g.mesh].has_uv ? 2u : 0u);
    }
    /* world box for the ray-sort keys */
    float wlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, whi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (size_t i = 0; i < groups.size(); i++)
        for (int r = 0; r < 3; r++) {
            wlo[r] = std::min(wlo[r], wboxes[6 * i + r]);
            whi[r] = std::max(whi[r], wboxes[6 * i + 3 + r]);
        }
    /* TLAS over the groups */
    BvhBuildResult tl;
    lx_build_tlas_host(wboxes, opt.bvh_width, tl);
    /* and a second TLAS over the groups with an emitting member, for
     * direct-light shadow rays (whose closest hit only matters when it
     * emits), appended to the first one's nodes with its leaves naming the
     * groups' indices */
    uint32_t emis_root_local = LX_CHILD_EMPTY;
    {
        std::vector<float> eboxes;
        std::vector<uint32_t> egroup;
        for (size_t gi = 0; gi < groups.size(); gi++)
            if (dgroups[gi].flags & LX_GROUP_EMIS_ANY) {
                eboxes.insert(eboxes.end(), wboxes.begin() + 6 * gi, wboxes.begin() + 6 * gi + 6);
                egroup.push_back((uint32_t)gi);
            }
        if (!egroup.empty()) {
            BvhBuildResult et;
            lx_build_tlas_host(eboxes, opt.bvh_width, et);
            const uint32_t off = (uint32_t)tl.nodes.size();
            auto remap = [&](uint32_t v) {
                if (v == LX_CHILD_EMPTY) return v;
                if (v & LX_LEAF_BIT)
                    return (v & ~LX_LEAF_FIRST_MASK) | egroup[v & LX_LEAF_FIRST_MASK];
                return v + off;
            };
            for (LxWideNode n : et.nodes) {
                for (int k = 0; k < 8; k++) n.child[k] = remap(n.child[k]);
                tl.nodes.push_back(n);
            }
            emis_root_local = remap(et.root);
        }
    }
    const uint32_t tlas_first = node_total;
    LX_REQUIRE((size_t)tlas_first + tl.nodes.size() < LX_LEAF_BIT, LX_ERR_INVALID,
               "scene too large for the node encoding");
    /* the traversal addresses a node's planes by 32-bit byte offsets from the node array */
    LX_REQUIRE(((size_t)tlas_first + tl.nodes.size()) * (opt.bvh_width == 8 ? 256u : 128u) <= (size_t)UINT32_MAX,
               LX_ERR_INVALID, "scene too large: BVH node array above 4 GiB (build at BVH width 4 or split the scene)");
    for (LxWideNode& n : tl.nodes)
        for (int k = 0; k < 8; k++) {
            uint32_t v = n.child[k];
            if (v == LX_CHILD_EMPTY || (v & LX_LEAF_BIT)) continue;
            n.child[k] = v + tlas_first;
        }
    uint32_t tlas_root = (tl.root & LX_LEAF_BIT) ? tl.root : tl.root + tlas_first;
    uint32_t emis_tlas_root = (emis_root_local == LX_CHILD_EMPTY || (emis_root_local & LX_LEAF_BIT))
                                  ? emis_root_local
                                  : emis_root_local + tlas_first;

    /* uploads: nodes packed at the traversal's width; BLASes the GPU builder
     * left on the device are copied there */
    std::vector<std::vector<LxNode8>> hold8;
    std::vector<std::vector<LxNode4>> hold4;
    if (opt.bvh_width == 8)
        upload_bvh<LxNode8, 8>(ctx, segs, nodes, tris, dev_blas, tl, tlas_first, tri_total, s->d_nodes, s->d_tris,
                               hold8);
    else
        upload_bvh<LxNode4, 4>(ctx, segs, nodes, tris, dev_blas, tl, tlas_first, tri_total, s->d_nodes, s->d_tris,
                               hold4);
    s->d_inst.upload(ctx, dinst.data(), dinst.size() * sizeof(LxInstance));
    s->d_groups.upload(ctx, dgroups.data(), dgroups.size() * sizeof(LxGroup));
    s->d_idx.upload(ctx, all_idx.data(), all_idx.size() * 4);
    if (any_n) s->d_nrm.upload(ctx, all_nrm.data(), all_nrm.size() * 4);
    if (any_uv) s->d_uv.upload(ctx, all_uv.data(), all_uv.size() * 4);
    s->d_mats.upload(ctx, s->mats.data(), s->mats.size() * sizeof(lx_material_desc));
    upload_textures(s);
    if (!s->emissive.empty()) {
        s->d_emis.upload(ctx, s->emissive.data(), s->emissive.size() * 4);
        s->d_emis_prefix.upload(ctx, s->emissive_prefix.data(), s->emissive_prefix.size() * 4);
        s->d_tri_cdf.upload(ctx, s->tri_cdf.data(), s->tri_cdf.size() * 4);
        s->d_tri_cdf_off.upload(ctx, s->tri_cdf_off.data(), s->tri_cdf_off.size() * 4);
    }
    LxDevScene& D = s->dev;
    D.nodes = s->d_nodes.ptr;
    D.tris = s->d_tris.as<float4>();
    D.inst = s->d_inst.as<LxInstance>();
    D.groups = s->d_groups.as<LxGroup>();
    D.idx = s->d_idx.as<uint32_t>();
    D.pos = s->d_pos.as<float>();
    D.nrm = s->d_nrm.as<float>();
    D.uv = s->d_uv.as<float>();
    D.mats = s->d_mats.as<lx_material_desc>();
    D.texs = s->d_texdesc.as<LxTexDesc>();
    D.marginal = s->d_marg.as<float>();
    D.conditional = s->d_cond.as<float>();
    D.emissive_inst = s->d_emis.as<uint32_t>();
    D.emissive_prefix = s->d_emis_prefix.as<float>();
    D.tri_cdf = s->d_tri_cdf.as<float>();
    D.tri_cdf_off = s->d_tri_cdf_off.as<uint32_t>();
    D.tlas_root = tlas_root;
    D.emis_tlas_root = emis_tlas_root;
    D.num_inst = (uint32_t)s->insts.size();
    D.num_groups = (uint32_t)groups.size();
    D.num_emissive = (uint32_t)s->emissive.size();
    D.total_area = s->total_area;
    D.has_env = s->has_env ? 1 : 0;
    D.bvh_width = opt.bvh_width;
    double diag2 = 0.0;
    for (int r = 0; r < 3; r++) {
        double e = groups.empty() ? 0.0 : (double)whi[r] - (double)wlo[r];
        D.box_lo[r] = groups.empty() ? 0.0f : wlo[r];
        D.box_inv[r] = e > 0.0 ? (float)(1.0 / e) : 0.0f;
        diag2 += e * e;
    }
    D.sort_reach = (float)(0.25 * std::sqrt(diag2));
    s->d_scene.upload(ctx, &D, sizeof(LxDevScene));
    LX_HIP(hipStreamSynchronize(ctx->stream));
    s->accel_built = true;
    s->info.num_emissive_instances = (uint32_t)s->emissive.size();
    s->info.total_emissive_area = s->total_area;
    s->info.env_width = s->has_env ? (int)s->env.w : 0;
    s->info.env_height = s->has_env ? (int)s->env.h : 0;
    s->info.bvh_nodes = tlas_first + (uint32_t)tl.nodes.size();
    s->info.bvh_depth = max_depth + tl.depth + 1;
    s->info.build_ms = now_ms() - t0;
    s->info.sah_cost = sah;
    uint64_t bytes = 0;
    DevBuf* bufs[] = {&s->d_nodes, &s->d_tris, &s->d_inst, &s->d_groups, &s->d_idx,  &s->d_pos,         &s->d_nrm,
                      &s->d_uv,    &s->d_mats, &s->d_texdesc, &s->d_env, &s->d_marg, &s->d_cond,
                      &s->d_emis,  &s->d_emis_prefix, &s->d_tri_cdf, &s->d_tri_cdf_off, &s->d_scene};
    for (DevBuf* b : bufs) bytes += b->bytes;
    for (auto& t : s->d_tex) bytes += t.bytes;
    s->info.device_bytes = bytes;
}
