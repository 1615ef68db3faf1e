/*
 * lx_gltf.cpp -- lx_scene_load_glb: binary glTF (+ optional EXR) into a scene.
 *
 * Follows the reference's loader, src/scene/scene_load_gltf.cpp:61-529, and
 * the camera set-up the viewer applies before the first frame
 * (src/main.cpp:170-190, src/app/render_loop.cpp:86-117):
 *   - material 0 is the default red material; glTF material k becomes k + 1
 *     (scene_load_gltf.cpp:85-89, 318);
 *   - material fields: baseColorFactor (alpha = its 4th entry), metallic and
 *     roughness factors, KHR_materials_emissive_strength (default 1),
 *     KHR_materials_ior (default 1.55), KHR_materials_transmission,
 *     KHR_texture_transform on every texture slot, alpha mode and cutoff
 *     (150-231);
 *   - textures in glTF texture order, decoded to four channels of 8 or 16
 *     bits (92-147);
 *   - node transforms: matrix, or translate * rotate * scale, composed
 *     parent * child (236-278); triangle primitives only; normals normalised
 *     per vertex; TEXCOORD_0; u8/u16/u32 or implicit indices (279-427);
 *   - one mesh per (glTF mesh, primitive) shared by every node that places it
 *     (the device-mesh reuse of src/main.cpp:199-216);
 *   - exactly one camera: 1000 pixels on the short side, pixel lengths from
 *     yfov; trace depth 8 (440-495); then view re-derived from its spherical
 *     angles and up / right re-orthogonalised against world +Y;
 *   - the EXR, when given, becomes the environment map (497-522).
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>

#include "lx_image.h"
#include "lx_internal.h"
#include "lx_json.h"

namespace {

[[noreturn]] void bad(const std::string& what) { throw LxError{LX_ERR_FORMAT, "glTF: " + what}; }

uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

/* column-major 4x4, as glm stores it: m[c * 4 + r] */
struct Mat4 {
    float m[16];
    static Mat4 identity() {
        Mat4 a;
        std::memset(a.m, 0, sizeof(a.m));
        a.m[0] = a.m[5] = a.m[10] = a.m[15] = 1.0f;
        return a;
    }
    float& at(int c, int r) { return m[c * 4 + r]; }
    float at(int c, int r) const { return m[c * 4 + r]; }
};

/* glm's mat4 * mat4: column c of the product = A * column c of B, each
 * entry summed over k in order */
Mat4 mul(const Mat4& A, const Mat4& B) {
    Mat4 o;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            float s = A.at(0, r) * B.at(c, 0);
            s = s + A.at(1, r) * B.at(c, 1);
            s = s + A.at(2, r) * B.at(c, 2);
            s = s + A.at(3, r) * B.at(c, 3);
            o.at(c, r) = s;
        }
    return o;
}

/* glm::mat4_cast of a unit quaternion (w, x, y, z) */
Mat4 quat_mat(float w, float x, float y, float z) {
    Mat4 r = Mat4::identity();
    float qxx = x * x, qyy = y * y, qzz = z * z;
    float qxz = x * z, qxy = x * y, qyz = y * z;
    float qwx = w * x, qwy = w * y, qwz = w * z;
    r.at(0, 0) = 1.0f - 2.0f * (qyy + qzz);
    r.at(0, 1) = 2.0f * (qxy + qwz);
    r.at(0, 2) = 2.0f * (qxz - qwy);
    r.at(1, 0) = 2.0f * (qxy - qwz);
    r.at(1, 1) = 1.0f - 2.0f * (qxx + qzz);
    r.at(1, 2) = 2.0f * (qyz + qwx);
    r.at(2, 0) = 2.0f * (qxz + qwy);
    r.at(2, 1) = 2.0f * (qyz - qwx);
    r.at(2, 2) = 1.0f - 2.0f * (qxx + qyy);
    return r;
}

struct V3 {
    float x, y, z;
};
V3 v3(float x, float y, float z) { return V3{x, y, z}; }
V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V3 cross(V3 a, V3 b) { return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
float length(V3 a) { return std::sqrt(dot(a, a)); }
V3 normalize(V3 a) { return a * (1.0f / std::sqrt(dot(a, a))); }

struct Glb {
    LxJson doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::string dir;

    const LxJson& arr(const char* k, size_t i) const {
        const std::vector<LxJson>& a = doc.array(k);
        if (i >= a.size()) bad(std::string(k) + " index " + std::to_string(i) + " out of range");
        return a[i];
    }

    /* bytes of buffer view v, with its offset applied */
    const uint8_t* view(int v, size_t& len, size_t& stride) const {
        const LxJson& bv = arr("bufferViews", (size_t)v);
        int b = bv.integer("buffer", -1);
        if (b < 0 || (size_t)b >= buffers.size()) bad("bufferView without a loaded buffer");
        size_t off = (size_t)bv.number("byteOffset", 0);
        len = (size_t)bv.number("byteLength", 0);
        stride = (size_t)bv.number("byteStride", 0);
        if (off + len > buffers[b].size()) bad("bufferView outside its buffer");
        return buffers[b].data() + off;
    }
};

std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) throw LxError{LX_ERR_IO, "cannot read " + path};
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::vector<uint8_t> base64(const std::string& s) {
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (char c : s) {
        int v;
        if (c >= 'A' && c <= 'Z') v = c - 'A';
        else if (c >= 'a' && c <= 'z') v = c - 'a' + 26;
        else if (c >= '0' && c <= '9') v = c - '0' + 52;
        else if (c == '+') v = 62;
        else if (c == '/') v = 63;
        else continue;
        acc = acc << 6 | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

std::vector<uint8_t> load_uri(const std::string& uri, const std::string& dir) {
    if (uri.compare(0, 5, "data:") == 0) {
        size_t c = uri.find(";base64,");
        if (c == std::string::npos) bad("only base64 data URIs are supported");
        return base64(uri.substr(c + 8));
    }
    return read_file(dir + uri);
}

Glb parse_glb(const std::string& path) {
    std::vector<uint8_t> f = read_file(path);
    if (f.size() < 20 || le32(f.data()) != 0x46546C67u) bad(path + " is not a binary glTF (.glb) file");
    if (le32(f.data() + 4) != 2) bad("only glTF 2.0 is supported");
    size_t total = std::min<size_t>(le32(f.data() + 8), f.size());
    Glb g;
    size_t slash = path.find_last_of("/\\");
    g.dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    size_t i = 12;
    bool have_json = false;
    std::vector<uint8_t> bin;
    bool have_bin = false;
    while (i + 8 <= total) {
        uint32_t len = le32(&f[i]), type = le32(&f[i + 4]);
        if (i + 8 + (size_t)len > total) bad("truncated chunk");
        if (type == 0x4E4F534Au) {
            g.doc = LxJsonParser((const char*)&f[i + 8], len).parse();
            have_json = true;
        } else if (type == 0x004E4942u && !have_bin) {
            bin.assign(f.begin() + i + 8, f.begin() + i + 8 + len);
            have_bin = true;
        }
        i += 8 + (size_t)((len + 3) & ~3u);
    }
    if (!have_json) bad("no JSON chunk");
    for (const LxJson& b : g.doc.array("buffers")) {
        std::string uri = b.string("uri");
        if (uri.empty()) {
            if (!have_bin) bad("buffer without uri and no BIN chunk");
            g.buffers.push_back(bin);
        } else {
            g.buffers.push_back(load_uri(uri, g.dir));
        }
    }
    return g;
}

/* float VEC2 / VEC3 accessor */
std::vector<float> read_floats(const Glb& g, int acc, int comps) {
    const LxJson& a = g.arr("accessors", (size_t)acc);
    if (a.integer("componentType", 0) != 5126) bad("vertex attribute is not float");
    std::string type = a.string("type");
    if ((comps == 3 && type != "VEC3") || (comps == 2 && type != "VEC2")) bad("unexpected accessor type " + type);
    size_t count = (size_t)a.number("count", 0);
    std::vector<float> out(count * comps, 0.0f);
    if (!a.has("bufferView")) return out; /* all zero, as glTF defines */
    size_t len, stride;
    const uint8_t* base = g.view(a.integer("bufferView", 0), len, stride);
    size_t off = (size_t)a.number("byteOffset", 0);
    if (!stride) stride = 4 * (size_t)comps;
    if (count && off + (count - 1) * stride + 4 * (size_t)comps > len) bad("accessor outside its buffer view");
    for (size_t i = 0; i < count; i++) std::memcpy(&out[i * comps], base + off + i * stride, 4 * (size_t)comps);
    return out;
}

std::vector<uint32_t> read_indices(const Glb& g, int acc) {
    const LxJson& a = g.arr("accessors", (size_t)acc);
    int ct = a.integer("componentType", 0);
    size_t count = (size_t)a.number("count", 0);
    std::vector<uint32_t> out(count, 0);
    if (!a.has("bufferView")) return out;
    size_t len, stride;
    const uint8_t* base = g.view(a.integer("bufferView", 0), len, stride);
    size_t off = (size_t)a.number("byteOffset", 0);
    size_t es = ct == 5121 ? 1 : (ct == 5123 ? 2 : (ct == 5125 ? 4 : 0));
    if (!es) bad("indices must be unsigned byte, short or int");
    if (!stride) stride = es;
    if (count && off + (count - 1) * stride + es > len) bad("index accessor outside its buffer view");
    for (size_t i = 0; i < count; i++) {
        const uint8_t* p = base + off + i * stride;
        out[i] = es == 1 ? p[0] : (es == 2 ? (uint32_t)(p[0] | p[1] << 8) : le32(p));
    }
    return out;
}

lx_tex_transform tex_xf(const LxJson* info) {
    lx_tex_transform t{{0.0f, 0.0f}, {1.0f, 1.0f}, 0.0f};
    if (!info) return t;
    const LxJson* ext = info->get("extensions");
    const LxJson* tt = ext ? ext->get("KHR_texture_transform") : nullptr;
    if (!tt) return t;
    const std::vector<LxJson>& o = tt->array("offset");
    if (o.size() >= 2) {
        t.offset[0] = (float)o[0].num;
        t.offset[1] = (float)o[1].num;
    }
    const std::vector<LxJson>& s = tt->array("scale");
    if (s.size() >= 2) {
        t.scale[0] = (float)s[0].num;
        t.scale[1] = (float)s[1].num;
    }
    if (tt->has("rotation")) t.rotation = (float)tt->number("rotation", 0.0);
    return t;
}

int tex_index(const LxJson* info) { return info ? info->integer("index", -1) : -1; }

double ext_number(const LxJson& mat, const char* ext, const char* key, double dflt) {
    const LxJson* e = mat.get("extensions");
    const LxJson* x = e ? e->get(ext) : nullptr;
    return x ? x->number(key, dflt) : dflt;
}

/* isGlass, include/scene/gltf_material_utils.h:9-47: a transmission or volume
 * extension, an ior away from 1, a see-through MASK / BLEND base colour, or a
 * see-through specular-glossiness diffuse factor. */
bool legacy_is_glass(const LxJson& mat, double base_alpha) {
    const LxJson* e = mat.get("extensions");
    if (e && (e->has("KHR_materials_transmission") || e->has("KHR_materials_volume"))) return true;
    if (const LxJson* x = e ? e->get("KHR_materials_ior") : nullptr) {
        const LxJson* v = x->get("ior");
        if (v && v->kind == LxJson::NUM && std::fabs(v->num - 1.0) > 0.01) return true;
    }
    std::string am = mat.string("alphaMode", "OPAQUE");
    if ((am == "BLEND" || am == "MASK") && base_alpha < (double)0.99f) return true;
    if (const LxJson* x = e ? e->get("KHR_materials_pbrSpecularGlossiness") : nullptr) {
        const std::vector<LxJson>& df = x->array("diffuseFactor");
        if (df.size() == 4 && df[3].kind == LxJson::NUM && df[3].num < 0.99) return true;
    }
    return false;
}

} // namespace

/* The scene description built from a .glb file, owning every array the
 * descriptor points at. */
struct LxGltfScene {
    std::vector<std::vector<float>> pos, nrm, uv;
    std::vector<std::vector<uint32_t>> idx;
    std::vector<lx_mesh_desc> meshes;
    std::vector<lx_instance_desc> insts;
    std::vector<lx_material_desc> mats;
    std::vector<LxImage> images;
    std::vector<lx_texture_desc> texs;
    LxImageF env;
    lx_texture_desc env_desc{};
    lx_scene_desc desc{};
    std::string name;
};

void lx_gltf_build(const std::string& glb_path, const char* exr_path, LxGltfScene& S) {
    Glb g = parse_glb(glb_path);
    const LxJson& d = g.doc;

    /* textures, in glTF texture order */
    for (const LxJson& t : d.array("textures")) {
        int src = t.integer("source", -1);
        if (src < 0) bad("texture without an image source");
        const LxJson& im = g.arr("images", (size_t)src);
        std::vector<uint8_t> bytes;
        const uint8_t* p;
        size_t n;
        if (im.has("bufferView")) {
            size_t stride;
            p = g.view(im.integer("bufferView", 0), n, stride);
        } else {
            bytes = load_uri(im.string("uri"), g.dir);
            p = bytes.data();
            n = bytes.size();
        }
        S.images.push_back(lx_decode_image(p, n));
    }
    for (const LxImage& im : S.images)
        S.texs.push_back(lx_texture_desc{im.w, im.h, im.bits == 16 ? LX_TEX_RGBA16 : LX_TEX_RGBA8, im.data.data()});

    /* materials: the default red one, then the file's */
    lx_material_desc def;
    std::memset(&def, 0, sizeof(def));
    def.base_color[0] = 1.0f;
    def.albedo_tex = def.metallic_roughness_tex = def.normal_tex = def.emissive_tex = -1;
    def.albedo_xf = def.metallic_roughness_xf = def.normal_xf = def.emissive_xf = tex_xf(nullptr);
    def.alpha = 1.0f;
    def.alpha_mode = LX_ALPHA_OPAQUE;
    def.alpha_cutoff = 0.5f;
    S.mats.push_back(def);
    for (const LxJson& m : d.array("materials")) {
        lx_material_desc md = def;
        md.base_color[0] = md.base_color[1] = md.base_color[2] = 0.0f;
        const LxJson* pbr = m.get("pbrMetallicRoughness");
        static const LxJson empty_obj = [] {
            LxJson j;
            j.kind = LxJson::OBJ;
            return j;
        }();
        if (!pbr) pbr = &empty_obj;
        const std::vector<LxJson>& bc = pbr->array("baseColorFactor");
        double bcf[4] = {1.0, 1.0, 1.0, 1.0};
        for (size_t k = 0; k < bc.size() && k < 4; k++) bcf[k] = bc[k].num;
        for (int k = 0; k < 3; k++) md.base_color[k] = (float)bcf[k];
        md.metallic = (float)pbr->number("metallicFactor", 1.0);
        md.roughness = (float)pbr->number("roughnessFactor", 1.0);
        md.metallic_roughness_tex = tex_index(pbr->get("metallicRoughnessTexture"));
        md.metallic_roughness_xf = tex_xf(pbr->get("metallicRoughnessTexture"));
        md.emission_strength = (float)ext_number(m, "KHR_materials_emissive_strength", "emissiveStrength", 1.0);
        const std::vector<LxJson>& ef = m.array("emissiveFactor");
        for (int k = 0; k < 3; k++) md.emission_color[k] = k < (int)ef.size() ? (float)ef[k].num : 0.0f;
        md.emissive_tex = tex_index(m.get("emissiveTexture"));
        md.emissive_xf = tex_xf(m.get("emissiveTexture"));
        md.ior = (float)ext_number(m, "KHR_materials_ior", "ior", 1.55);
        md.transmission = (float)ext_number(m, "KHR_materials_transmission", "transmissionFactor", 0.0);
        md.alpha = (float)bcf[3];
        std::string am = m.string("alphaMode", "OPAQUE");
        md.alpha_mode = am == "MASK" ? LX_ALPHA_MASK : (am == "BLEND" ? LX_ALPHA_BLEND : LX_ALPHA_OPAQUE);
        md.alpha_cutoff = (float)m.number("alphaCutoff", 0.5);
        md.albedo_tex = tex_index(pbr->get("baseColorTexture"));
        md.albedo_xf = tex_xf(pbr->get("baseColorTexture"));
        md.normal_tex = tex_index(m.get("normalTexture"));
        md.normal_xf = tex_xf(m.get("normalTexture"));
        /* legacy material class (the UBER_SHADER 0 branch,
         * src/scene/scene_load_gltf.cpp:211-228): emissive, else glass,
         * else microfacet */
        float ec = std::sqrt(md.emission_color[0] * md.emission_color[0] + md.emission_color[1] * md.emission_color[1] +
                             md.emission_color[2] * md.emission_color[2]);
        if (md.emission_strength * ec > 0.00001f) md.material_type = LX_MATERIAL_EMISSIVE;
        else if (legacy_is_glass(m, bcf[3]) && md.metallic < 0.01f && md.roughness < 0.01f)
            md.material_type = LX_MATERIAL_GLASS;
        else md.material_type = LX_MATERIAL_MICROFACET;
        for (int t : {md.albedo_tex, md.metallic_roughness_tex, md.normal_tex, md.emissive_tex})
            if (t >= (int)S.texs.size()) bad("material names a missing texture");
        S.mats.push_back(md);
    }

    /* nodes */
    std::map<std::pair<int, int>, uint32_t> mesh_of; /* (glTF mesh, primitive) -> mesh */
    int cam_node = -1;
    Mat4 cam_xf = Mat4::identity();
    std::function<void(int, const Mat4&, int)> visit = [&](int ni, const Mat4& parent, int depth) {
        if (depth > 1000) bad("node hierarchy too deep (cycle?)");
        const LxJson& node = g.arr("nodes", (size_t)ni);
        Mat4 local = Mat4::identity();
        const std::vector<LxJson>& mm = node.array("matrix");
        if (mm.size() == 16) {
            for (int k = 0; k < 16; k++) local.m[k] = (float)mm[k].num;
        } else {
            const std::vector<LxJson>& t = node.array("translation");
            if (t.size() == 3) {
                /* glm::translate(I, t): column 3 = t */
                local.at(3, 0) = (float)t[0].num;
                local.at(3, 1) = (float)t[1].num;
                local.at(3, 2) = (float)t[2].num;
            }
            const std::vector<LxJson>& r = node.array("rotation");
            if (r.size() == 4) local = mul(local, quat_mat((float)r[3].num, (float)r[0].num, (float)r[1].num, (float)r[2].num));
            const std::vector<LxJson>& s = node.array("scale");
            if (s.size() == 3) {
                /* glm::scale(M, s): columns 0..2 scaled */
                for (int c = 0; c < 3; c++)
                    for (int rr = 0; rr < 4; rr++) local.at(c, rr) = local.at(c, rr) * (float)s[c].num;
            }
        }
        Mat4 global = mul(parent, local);
        int mi = node.integer("mesh", -1);
        if (mi >= 0) {
            const LxJson& mesh = g.arr("meshes", (size_t)mi);
            const std::vector<LxJson>& prims = mesh.array("primitives");
            for (size_t pi = 0; pi < prims.size(); pi++) {
                const LxJson& prim = prims[pi];
                if (prim.integer("mode", 4) != 4) continue; /* triangles only */
                const LxJson* attrs = prim.get("attributes");
                if (!attrs || !attrs->has("POSITION")) continue;
                auto key = std::make_pair(mi, (int)pi);
                auto it = mesh_of.find(key);
                uint32_t meshid;
                if (it != mesh_of.end()) {
                    meshid = it->second;
                } else {
                    std::vector<float> P = read_floats(g, attrs->integer("POSITION", 0), 3);
                    std::vector<float> N, T;
                    if (attrs->has("NORMAL")) {
                        N = read_floats(g, attrs->integer("NORMAL", 0), 3);
                        for (size_t k = 0; k + 2 < N.size(); k += 3) {
                            V3 n = normalize(v3(N[k], N[k + 1], N[k + 2]));
                            N[k] = n.x;
                            N[k + 1] = n.y;
                            N[k + 2] = n.z;
                        }
                        if (N.size() != P.size()) bad("NORMAL count differs from POSITION count");
                    }
                    if (attrs->has("TEXCOORD_0")) {
                        T = read_floats(g, attrs->integer("TEXCOORD_0", 0), 2);
                        if (T.size() / 2 != P.size() / 3) bad("TEXCOORD_0 count differs from POSITION count");
                    }
                    std::vector<uint32_t> I;
                    if (prim.has("indices")) {
                        I = read_indices(g, prim.integer("indices", 0));
                    } else {
                        I.resize(P.size() / 3);
                        for (size_t k = 0; k < I.size(); k++) I[k] = (uint32_t)k;
                    }
                    I.resize(I.size() / 3 * 3);
                    for (uint32_t v : I)
                        if (v >= P.size() / 3) bad("index out of range");
                    if (I.empty()) continue;
                    meshid = (uint32_t)S.pos.size();
                    S.pos.push_back(std::move(P));
                    S.nrm.push_back(std::move(N));
                    S.uv.push_back(std::move(T));
                    S.idx.push_back(std::move(I));
                    mesh_of[key] = meshid;
                }
                lx_instance_desc inst;
                inst.mesh = meshid;
                int pm = prim.integer("material", -1);
                if (pm >= (int)S.mats.size() - 1) bad("primitive names a missing material");
                inst.material = pm >= 0 ? (uint32_t)(pm + 1) : 0u;
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 4; c++) inst.transform[4 * r + c] = global.at(c, r);
                S.insts.push_back(inst);
            }
        }
        if (node.integer("camera", -1) >= 0) {
            if (cam_node != -1) bad("multiple cameras defined");
            cam_node = ni;
            cam_xf = global;
        }
        for (const LxJson& c : node.array("children")) visit((int)c.num, global, depth + 1);
    };
    int si = d.integer("scene", 0);
    const LxJson& scene = g.arr("scenes", (size_t)(si < 0 ? 0 : si));
    S.name = scene.string("name", "");
    if (S.name.empty()) S.name = "GLTF_DEFAULT_NAME";
    for (const LxJson& n : scene.array("nodes")) visit((int)n.num, Mat4::identity(), 0);
    if (cam_node < 0) bad("no camera found");
    if (S.insts.empty()) bad("no triangle geometry");

    /* camera (scene_load_gltf.cpp:457-495) */
    const LxJson& cam = g.arr("cameras", (size_t)g.arr("nodes", (size_t)cam_node).integer("camera", 0));
    const LxJson* persp = cam.get("perspective");
    if (!persp) bad("camera is not perspective");
    float aspect = (float)persp->number("aspectRatio", 0.0);
    if (!(aspect > 0.0f)) bad("camera needs an aspectRatio");
    lx_camera_desc C;
    if (aspect >= 1.0f) {
        C.height = 1000;
        C.width = (int)(aspect * (float)C.height);
    } else {
        C.width = 1000;
        C.height = (int)((float)C.width / aspect);
    }
    double yfov = persp->number("yfov", 0.0);
    float fovy = (float)(yfov * (180.0 / M_PI));
    const float PIf = 3.14159265358979323846f;
    V3 eye = v3(cam_xf.at(3, 0), cam_xf.at(3, 1), cam_xf.at(3, 2));
    V3 view = -normalize(v3(cam_xf.at(2, 0), cam_xf.at(2, 1), cam_xf.at(2, 2)));
    V3 look = eye + view;
    float yscaled = std::tan(0.5f * fovy * (PIf / 180.0f));
    float xscaled = (yscaled * (float)C.width) / (float)C.height;
    C.pixel_length[0] = 2.0f * xscaled / (float)C.width;
    C.pixel_length[1] = 2.0f * yscaled / (float)C.height;
    /* src/main.cpp:178-189 then src/app/render_loop.cpp:91-116 */
    V3 vn = normalize(view);
    float phi = std::atan2(vn.x, -vn.z);
    float theta = std::acos(std::min(std::max(vn.y, -1.0f), 1.0f));
    float zoom = length(eye - look);
    V3 sph = v3(-zoom * std::sin(phi) * std::sin(theta), -zoom * std::cos(theta), zoom * std::cos(phi) * std::sin(theta));
    V3 v = -normalize(sph);
    V3 refup = v3(0.0f, 1.0f, 0.0f);
    V3 u = refup - v * dot(refup, v);
    if (std::pow(length(u), 2.0f) < 1e-6f) u = v3(0.0f, 1.0f, 0.0f);
    u = normalize(u);
    V3 r = normalize(cross(v, u));
    u = cross(r, v);
    V3 pos = look - v * zoom;
    C.position[0] = pos.x;
    C.position[1] = pos.y;
    C.position[2] = pos.z;
    C.view[0] = v.x;
    C.view[1] = v.y;
    C.view[2] = v.z;
    C.up[0] = u.x;
    C.up[1] = u.y;
    C.up[2] = u.z;
    C.right[0] = r.x;
    C.right[1] = r.y;
    C.right[2] = r.z;

    if (exr_path && *exr_path) {
        S.env = lx_load_exr(exr_path);
        S.env_desc = lx_texture_desc{S.env.w, S.env.h, LX_TEX_RGBA32F, S.env.rgba.data()};
    }

    for (size_t k = 0; k < S.pos.size(); k++) {
        lx_mesh_desc m;
        m.num_verts = (uint32_t)(S.pos[k].size() / 3);
        m.positions = S.pos[k].data();
        m.normals = S.nrm[k].empty() ? nullptr : S.nrm[k].data();
        m.uvs = S.uv[k].empty() ? nullptr : S.uv[k].data();
        m.num_tris = (uint32_t)(S.idx[k].size() / 3);
        m.indices = S.idx[k].data();
        S.meshes.push_back(m);
    }
    lx_scene_desc& D = S.desc;
    D.meshes = S.meshes.data();
    D.num_meshes = (uint32_t)S.meshes.size();
    D.instances = S.insts.data();
    D.num_instances = (uint32_t)S.insts.size();
    D.materials = S.mats.data();
    D.num_materials = (uint32_t)S.mats.size();
    D.textures = S.texs.data();
    D.num_textures = (uint32_t)S.texs.size();
    D.envmap = (exr_path && *exr_path) ? &S.env_desc : nullptr;
    D.camera = C;
    D.trace_depth = 8;
}

extern "C" lx_status lx_scene_load_glb(lx_context* ctx, const char* glb_path, const char* exr_path, lx_scene** out) {
    LX_TRY
    LX_REQUIRE(ctx && glb_path && out, LX_ERR_INVALID, "null argument");
    *out = nullptr;
    LxGltfScene S;
    lx_gltf_build(glb_path, exr_path, S);
    lx_status st = lx_scene_create(ctx, &S.desc, out);
    if (st != LX_OK) return st;
    std::snprintf((*out)->info.name, sizeof((*out)->info.name), "%s", S.name.c_str());
    LX_CATCH
}

/* Host-only view of what the loader produced (no device needed): counts and
 * the first instance / material / camera, for tests and tools. */
extern "C" lx_status lx_debug_parse_glb(const char* glb_path, const char* exr_path, lx_scene_desc* summary,
                                        float* env_rgba_head, uint32_t head_n) {
    static thread_local LxGltfScene keep; /* arrays stay valid until the next call on this thread */
    LX_TRY
    LX_REQUIRE(glb_path && summary, LX_ERR_INVALID, "null argument");
    keep = LxGltfScene();
    lx_gltf_build(glb_path, exr_path, keep);
    *summary = keep.desc;
    if (env_rgba_head && keep.desc.envmap)
        std::memcpy(env_rgba_head, keep.env.rgba.data(), std::min<size_t>(head_n, keep.env.rgba.size()) * 4);
    LX_CATCH
}
