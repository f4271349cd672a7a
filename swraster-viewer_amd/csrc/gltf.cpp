// gltf.cpp -- glTF 2.0 scene loading into the host arrays of sr_scene_desc.
//
// Follows the reference's loader (load_gltf, src/main.rs:81-97;
// Scene::from_gltf and friends, src/scene.rs:145-815; TextureCache,
// src/texture.rs:866-1015):
//   - .gltf with external or base64 data-URI buffers, or .glb;
//   - every document mesh, its primitives in order: POSITION and indices
//     required, TEXCOORD_0 (float or normalized u8/u16) else zeros, NORMAL
//     else smooth normals (compute_smooth_normals, src/scene.rs:515-546),
//     TANGENT when present (else computed where a normal map needs them);
//     a primitive without a material takes material 0 (src/scene.rs:493);
//   - every document node in document order with its world transform: the
//     local matrix or T * R * S (src/scene.rs:394-418) composed from the
//     roots (nodes that are nobody's child) down (src/scene.rs:306-381);
//   - materials (src/scene.rs:692-781): base colour, metallic-roughness,
//     normal, emissive, occlusion (strength), alpha mode and cutoff,
//     KHR_materials_transmission (factor, texture); texture wrap modes from
//     the sampler (repeat by default), one texture per image URI (the first
//     use decides its type), images from URIs relative to the file only
//     (bufferView images give no texture, src/scene.rs:686).
// Textures are decoded from PNG (png.cpp).

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "png.h"
#include "prep.h"
#include "runtime.h"
#include "swraster.h"

using namespace sr;

namespace {

// ---------------------------------------------------------------------------
// JSON
// ---------------------------------------------------------------------------
struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;

    const JVal *get(const char *key) const {
        if (type != Obj) return nullptr;
        for (auto &kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double num(const char *key, double def) const {
        const JVal *v = get(key);
        return v && v->type == Num ? v->n : def;
    }
    // A number as an int; out of int's range (or NaN) it is -1, which every
    // caller rejects as an index.
    static int to_int(double x) { return x >= -2147483648.0 && x <= 2147483647.0 ? (int)x : -1; }
    int integer(const char *key, int def) const {
        const JVal *v = get(key);
        return v && v->type == Num ? to_int(v->n) : def;
    }
    std::string str(const char *key) const {
        const JVal *v = get(key);
        return v && v->type == Str ? v->s : std::string();
    }
    size_t size() const { return type == Arr ? a.size() : 0; }
    const JVal *at(size_t i) const { return type == Arr && i < a.size() ? &a[i] : nullptr; }
};

struct JsonParser {
    const char *p, *end;
    std::string err;
    int depth = 0;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool fail(const char *m) {
        if (err.empty()) err = m;
        return false;
    }
    static void utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(uint32_t &v) {
        if (end - p < 4) return fail("truncated \\u escape");
        v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string &out) {
        p++;  // opening quote
        while (p < end && *p != '"') {
            char c = *p++;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= end) return fail("truncated escape");
            char e = *p++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t lo;
                        if (!hex4(lo)) return false;
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        if (p >= end) return fail("unterminated string");
        p++;
        return true;
    }
    bool value(JVal &v) {
        if (++depth > 256) return fail("JSON nested too deeply");
        ws();
        if (p >= end) return fail("unexpected end of JSON");
        bool ok = true;
        if (*p == '{') {
            v.type = JVal::Obj;
            p++;
            ws();
            if (p < end && *p == '}') {
                p++;
            } else {
                for (;;) {
                    ws();
                    if (p >= end || *p != '"') { ok = fail("expected object key"); break; }
                    std::string key;
                    if (!string(key)) { ok = false; break; }
                    ws();
                    if (p >= end || *p != ':') { ok = fail("expected ':'"); break; }
                    p++;
                    v.o.emplace_back(std::move(key), JVal());
                    if (!value(v.o.back().second)) { ok = false; break; }
                    ws();
                    if (p < end && *p == ',') { p++; continue; }
                    if (p < end && *p == '}') { p++; break; }
                    ok = fail("expected ',' or '}'");
                    break;
                }
            }
        } else if (*p == '[') {
            v.type = JVal::Arr;
            p++;
            ws();
            if (p < end && *p == ']') {
                p++;
            } else {
                for (;;) {
                    v.a.emplace_back();
                    if (!value(v.a.back())) { ok = false; break; }
                    ws();
                    if (p < end && *p == ',') { p++; continue; }
                    if (p < end && *p == ']') { p++; break; }
                    ok = fail("expected ',' or ']'");
                    break;
                }
            }
        } else if (*p == '"') {
            v.type = JVal::Str;
            ok = string(v.s);
        } else if (end - p >= 4 && !strncmp(p, "true", 4)) {
            v.type = JVal::Bool;
            v.b = true;
            p += 4;
        } else if (end - p >= 5 && !strncmp(p, "false", 5)) {
            v.type = JVal::Bool;
            p += 5;
        } else if (end - p >= 4 && !strncmp(p, "null", 4)) {
            p += 4;
        } else {
            char buf[64];
            size_t k = 0;
            while (p < end && k < 63 && (strchr("+-.eE", *p) || (*p >= '0' && *p <= '9'))) buf[k++] = *p++;
            buf[k] = 0;
            char *e = nullptr;
            v.type = JVal::Num;
            v.n = strtod(buf, &e);
            if (k == 0 || e != buf + k) ok = fail("bad JSON value");
        }
        depth--;
        return ok;
    }
};

// ---------------------------------------------------------------------------
// files and URIs
// ---------------------------------------------------------------------------
bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n < 0) {
        fclose(f);
        return false;
    }
    out.resize((size_t)n);
    size_t got = n > 0 ? fread(out.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    return got == (size_t)n;
}

std::string percent_decode(const std::string &s) {
    std::string o;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '%' && i + 2 < s.size() && isxdigit((unsigned char)s[i + 1]) && isxdigit((unsigned char)s[i + 2])) {
            o += (char)strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            o += s[i];
        }
    }
    return o;
}

bool base64_decode(const std::string &in, size_t from, std::vector<uint8_t> &out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < in.size(); i++) {
        char c = in[i];
        if (c == '=') break;
        int v = val(c);
        if (v < 0) {
            if (c == '\n' || c == '\r' || c == ' ') continue;
            return false;
        }
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// loaded document
// ---------------------------------------------------------------------------
struct sr_gltf {
    std::vector<std::vector<float>> positions, normals, texcoords, tangents;
    std::vector<std::vector<uint32_t>> indices;
    std::vector<sr_primitive_desc> prims;
    std::vector<sr_mesh_desc> meshes;
    std::vector<sr_node_desc> nodes;
    std::vector<sr_material_desc> materials;
    std::vector<sr_texture_desc> textures;
    std::vector<std::vector<uint32_t>> texels;
    sr_light light;
};

namespace {

struct Loader {
    std::string dir;
    JVal doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::vector<uint8_t> glb_bin;
    bool has_glb_bin = false;
    sr_gltf *g = nullptr;
    std::map<std::string, int> tex_by_uri;

    int bad(const char *fmt, const std::string &what = std::string()) {
        return fail(SR_ERR_INVALID_ARGUMENT, fmt, what.c_str());
    }

    int load_buffers() {
        const JVal *bs = doc.get("buffers");
        for (size_t i = 0; bs && i < bs->size(); i++) {
            const JVal &b = *bs->at(i);
            std::string uri = b.str("uri");
            std::vector<uint8_t> data;
            if (uri.empty()) {
                if (!has_glb_bin) return bad("glTF buffer %s has no uri and the file has no GLB BIN chunk", std::to_string(i));
                data = glb_bin;
            } else if (uri.compare(0, 5, "data:") == 0) {
                size_t comma = uri.find(',');
                if (comma == std::string::npos || uri.find(";base64") == std::string::npos || uri.find(";base64") > comma)
                    return bad("glTF buffer data URI is not base64: %s", uri.substr(0, 40));
                if (!base64_decode(uri, comma + 1, data)) return bad("bad base64 in glTF buffer %s", std::to_string(i));
            } else {
                std::string path = dir + percent_decode(uri);
                if (!read_file(path, data)) return fail(SR_ERR_NOT_FOUND, "cannot read glTF buffer '%s'", path.c_str());
            }
            size_t need = 0;
            if (!to_size(b.num("byteLength", 0.0), need) || data.size() < need)
                return bad("glTF buffer %s is shorter than its byteLength", std::to_string(i));
            buffers.push_back(std::move(data));
        }
        return SR_OK;
    }

    // Element e, component c of an accessor as double (normalized integers
    // mapped to [0, 1] as the reference's into_f32 does).
    struct Acc {
        const uint8_t *base = nullptr;
        size_t count = 0, stride = 0;
        int comp = 0, ncomp = 0;
        bool normalized = false;
        size_t comp_size() const { return comp == 5126 || comp == 5125 ? 4 : (comp == 5123 || comp == 5122 ? 2 : 1); }
    };

    // A JSON number used as a byte offset, length, stride or count: a
    // non-negative value below 2^52 (so sums of two stay exact and cannot
    // wrap), else false.
    static bool to_size(double x, size_t &out) {
        if (!(x >= 0.0 && x < 4503599627370496.0)) return false;
        out = (size_t)x;
        return true;
    }

    int accessor(int index, Acc &A) {
        const JVal *as = doc.get("accessors");
        const JVal *a = as ? as->at((size_t)index) : nullptr;
        if (!a) return bad("glTF accessor %s does not exist", std::to_string(index));
        if (a->get("sparse")) return fail(SR_ERR_UNSUPPORTED, "sparse glTF accessors are not supported");
        std::string type = a->str("type");
        static const std::pair<const char *, int> TYPES[] = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}};
        A.ncomp = 0;
        for (auto &t : TYPES)
            if (type == t.first) A.ncomp = t.second;
        if (!A.ncomp) return bad("unsupported glTF accessor type '%s'", type);
        A.comp = a->integer("componentType", 0);
        if (!(A.comp == 5120 || A.comp == 5121 || A.comp == 5122 || A.comp == 5123 || A.comp == 5125 || A.comp == 5126))
            return bad("unsupported glTF component type %s", std::to_string(A.comp));
        A.normalized = a->get("normalized") && a->get("normalized")->b;
        const JVal *bv = a->get("bufferView");
        if (!bv) return fail(SR_ERR_UNSUPPORTED, "glTF accessor %d has no bufferView", index);
        const JVal *views = doc.get("bufferViews");
        size_t vi = 0;
        const JVal *v = views && to_size(bv->n, vi) ? views->at(vi) : nullptr;
        if (!v) return bad("glTF accessor %s refers to a missing bufferView", std::to_string(index));
        int buf = v->integer("buffer", -1);
        if (buf < 0 || (size_t)buf >= buffers.size()) return bad("glTF bufferView refers to a missing buffer");
        size_t voff = 0, vlen = 0, aoff = 0;
        if (!to_size(a->num("count", 0.0), A.count) || !to_size(v->num("byteOffset", 0.0), voff) ||
            !to_size(v->num("byteLength", 0.0), vlen) || !to_size(a->num("byteOffset", 0.0), aoff) ||
            !to_size(v->num("byteStride", 0.0), A.stride))
            return bad("glTF accessor %s has a negative or oversized count, offset, length or stride",
                       std::to_string(index));
        size_t elem = A.comp_size() * (size_t)A.ncomp;
        if (A.stride == 0) A.stride = elem;
        const size_t bsize = buffers[buf].size();
        if (voff > bsize || vlen > bsize - voff) return bad("glTF bufferView overruns its buffer");
        // the last element ends inside the view (no product or sum can wrap)
        if (A.count > 0 && (aoff > vlen || elem > vlen - aoff || A.count - 1 > (vlen - aoff - elem) / A.stride))
            return bad("glTF accessor overruns its bufferView");
        A.base = buffers[buf].data() + voff + aoff;
        return SR_OK;
    }

    static double component(const Acc &A, size_t e, int c) {
        const uint8_t *p = A.base + e * A.stride + (size_t)c * A.comp_size();
        switch (A.comp) {
            case 5126: { float f; memcpy(&f, p, 4); return f; }
            case 5125: { uint32_t u; memcpy(&u, p, 4); return u; }
            case 5123: { uint16_t u; memcpy(&u, p, 2); return A.normalized ? u / 65535.0 : u; }
            case 5122: { int16_t s; memcpy(&s, p, 2); return A.normalized ? std::fmax(s / 32767.0, -1.0) : s; }
            case 5121: return A.normalized ? p[0] / 255.0 : p[0];
            default: { int8_t s; memcpy(&s, p, 1); return A.normalized ? std::fmax(s / 127.0, -1.0) : s; }
        }
    }

    int read_floats(int index, int ncomp, std::vector<float> &out, bool allow_normalized) {
        Acc A;
        int rc = accessor(index, A);
        if (rc) return rc;
        if (A.ncomp != ncomp) return bad("glTF attribute accessor %s has the wrong element type", std::to_string(index));
        if (A.comp != 5126 && !(allow_normalized && A.normalized && (A.comp == 5121 || A.comp == 5123)))
            return fail(SR_ERR_UNSUPPORTED, "glTF accessor %d: attribute component type %d is not supported", index, A.comp);
        out.resize(A.count * (size_t)ncomp);
        if (A.comp == 5126) {
            for (size_t e = 0; e < A.count; e++) memcpy(&out[e * ncomp], A.base + e * A.stride, 4 * (size_t)ncomp);
        } else {
            // into_f32: u8 / 255 and u16 / 65535, in f32
            for (size_t e = 0; e < A.count; e++)
                for (int c = 0; c < ncomp; c++) {
                    const uint8_t *p = A.base + e * A.stride + (size_t)c * A.comp_size();
                    float v;
                    if (A.comp == 5121) {
                        v = (float)p[0] / 255.0f;
                    } else {
                        uint16_t u;
                        memcpy(&u, p, 2);
                        v = (float)u / 65535.0f;
                    }
                    out[e * ncomp + c] = v;
                }
        }
        return SR_OK;
    }

    int read_indices(int index, std::vector<uint32_t> &out) {
        Acc A;
        int rc = accessor(index, A);
        if (rc) return rc;
        if (A.ncomp != 1 || !(A.comp == 5121 || A.comp == 5123 || A.comp == 5125))
            return bad("glTF index accessor %s is not an unsigned scalar", std::to_string(index));
        out.resize(A.count);
        for (size_t e = 0; e < A.count; e++) out[e] = (uint32_t)component(A, e, 0);
        return SR_OK;
    }

    // compute_smooth_normals (src/scene.rs:515-546): area-weighted face
    // normals summed per vertex, then normalized; (0, 0, 1) for a vertex
    // whose sum has zero length.
    static void smooth_normals(const std::vector<float> &pos, const std::vector<uint32_t> &idx, std::vector<float> &n) {
        size_t nv = pos.size() / 3;
        n.assign(nv * 3, 0.0f);
        for (size_t t = 0; t + 2 < idx.size(); t += 3) {
            uint32_t i0 = idx[t], i1 = idx[t + 1], i2 = idx[t + 2];
            if (i0 >= nv || i1 >= nv || i2 >= nv) continue;
            float e1[3], e2[3];
            for (int k = 0; k < 3; k++) {
                e1[k] = pos[3 * i1 + k] - pos[3 * i0 + k];
                e2[k] = pos[3 * i2 + k] - pos[3 * i0 + k];
            }
            float f[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            for (uint32_t v : {i0, i1, i2})
                for (int k = 0; k < 3; k++) n[3 * v + k] += f[k];
        }
        for (size_t v = 0; v < nv; v++) {
            float *q = &n[3 * v];
            float l2 = (q[0] * q[0] + q[1] * q[1]) + q[2] * q[2];
            if (l2 > 0.0f) {
                float r = 1.0f / std::sqrt(l2);
                q[0] *= r; q[1] *= r; q[2] *= r;
            } else {
                q[0] = 0.0f; q[1] = 0.0f; q[2] = 1.0f;
            }
        }
    }

    int load_meshes() {
        const JVal *ms = doc.get("meshes");
        for (size_t mi = 0; ms && mi < ms->size(); mi++) {
            const JVal *ps = ms->at(mi)->get("primitives");
            sr_mesh_desc md;
            md.first_primitive = (int32_t)g->prims.size();
            md.num_primitives = (int32_t)(ps ? ps->size() : 0);
            for (size_t pi = 0; ps && pi < ps->size(); pi++) {
                const JVal &p = *ps->at(pi);
                const JVal *attr = p.get("attributes");
                const JVal *pa = attr ? attr->get("POSITION") : nullptr;
                if (!pa) return bad("glTF mesh %s has a primitive without positions", std::to_string(mi));
                const JVal *ia = p.get("indices");
                if (!ia) return bad("glTF mesh %s has a primitive without indices", std::to_string(mi));
                std::vector<float> pos, nrm, uv, tan;
                std::vector<uint32_t> idx;
                int rc = read_floats(JVal::to_int(pa->n), 3, pos, false);
                if (!rc) rc = read_indices(JVal::to_int(ia->n), idx);
                if (rc) return rc;
                size_t nv = pos.size() / 3;
                const JVal *ta = attr->get("TEXCOORD_0");
                if (ta) {
                    if ((rc = read_floats(JVal::to_int(ta->n), 2, uv, true))) return rc;
                    if (uv.size() / 2 != nv) return bad("glTF TEXCOORD_0 count differs from POSITION in mesh %s", std::to_string(mi));
                } else {
                    uv.assign(nv * 2, 0.0f);
                }
                const JVal *na = attr->get("NORMAL");
                if (na) {
                    if ((rc = read_floats(JVal::to_int(na->n), 3, nrm, false))) return rc;
                    if (nrm.size() / 3 != nv) return bad("glTF NORMAL count differs from POSITION in mesh %s", std::to_string(mi));
                } else {
                    smooth_normals(pos, idx, nrm);
                }
                const JVal *tga = attr->get("TANGENT");
                if (tga) {
                    if ((rc = read_floats(JVal::to_int(tga->n), 4, tan, false))) return rc;
                    if (tan.size() / 4 != nv) return bad("glTF TANGENT count differs from POSITION in mesh %s", std::to_string(mi));
                }
                for (uint32_t i : idx)
                    if (i >= nv) return bad("glTF index out of range in mesh %s", std::to_string(mi));
                g->positions.push_back(std::move(pos));
                g->normals.push_back(std::move(nrm));
                g->texcoords.push_back(std::move(uv));
                g->tangents.push_back(std::move(tan));
                g->indices.push_back(std::move(idx));
                sr_primitive_desc d;
                memset(&d, 0, sizeof(d));
                d.num_vertices = (uint32_t)nv;
                d.num_indices = (uint32_t)g->indices.back().size();
                d.material = p.integer("material", 0);
                g->prims.push_back(d);
            }
            g->meshes.push_back(md);
        }
        return SR_OK;
    }

    // Node::get_local_transform (src/scene.rs:394-418); column-major.
    static void local_transform(const JVal &n, float m[16]) {
        const JVal *mat = n.get("matrix");
        if (mat && mat->size() == 16) {
            for (int k = 0; k < 16; k++) m[k] = (float)mat->at(k)->n;
            return;
        }
        float t[3] = {0, 0, 0}, r[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
        if (const JVal *v = n.get("translation"))
            for (int k = 0; k < 3 && k < (int)v->size(); k++) t[k] = (float)v->at(k)->n;
        if (const JVal *v = n.get("rotation"))
            for (int k = 0; k < 4 && k < (int)v->size(); k++) r[k] = (float)v->at(k)->n;
        if (const JVal *v = n.get("scale"))
            for (int k = 0; k < 3 && k < (int)v->size(); k++) s[k] = (float)v->at(k)->n;
        float T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, t[0], t[1], t[2], 1};
        float S[16] = {s[0], 0, 0, 0, 0, s[1], 0, 0, 0, 0, s[2], 0, 0, 0, 0, 1};
        float R[16], TR[16];
        mat4_from_quat(r, R);
        mat4_mul(T, R, TR);
        mat4_mul(TR, S, m);
    }

    int load_nodes() {
        const JVal *ns = doc.get("nodes");
        size_t nn = ns ? ns->size() : 0;
        std::vector<std::array<float, 16>> local(nn);
        std::vector<char> is_child(nn, 0);
        g->nodes.resize(nn);
        for (size_t i = 0; i < nn; i++) {
            const JVal &n = *ns->at(i);
            local_transform(n, local[i].data());
            sr_node_desc &d = g->nodes[i];
            memcpy(d.transform, local[i].data(), 64);
            d.mesh = n.integer("mesh", -1);
            if (d.mesh >= (int)g->meshes.size()) return bad("glTF node %s refers to a missing mesh", std::to_string(i));
            if (const JVal *c = n.get("children"))
                for (size_t k = 0; k < c->size(); k++) {
                    int ci = JVal::to_int(c->at(k)->n);
                    if (ci < 0 || (size_t)ci >= nn) return bad("glTF node %s has a missing child", std::to_string(i));
                    is_child[(size_t)ci] = 1;
                }
        }
        // compute_final_transform_recursive (src/scene.rs:356-381), with an
        // explicit stack and a depth bound against cyclic files
        for (size_t root = 0; root < nn; root++) {
            if (is_child[root]) continue;
            struct Item { size_t node; std::array<float, 16> parent; int depth; };
            std::vector<Item> stack;
            std::array<float, 16> ident = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
            stack.push_back({root, ident, 0});
            while (!stack.empty()) {
                Item it = stack.back();
                stack.pop_back();
                if (it.depth > 1024) return bad("glTF node hierarchy is cyclic or too deep");
                std::array<float, 16> world;
                mat4_mul(it.parent.data(), local[it.node].data(), world.data());
                memcpy(g->nodes[it.node].transform, world.data(), 64);
                if (const JVal *c = ns->at(it.node)->get("children"))
                    for (size_t k = c->size(); k-- > 0;)
                        stack.push_back({(size_t)JVal::to_int(c->at(k)->n), world, it.depth + 1});
            }
        }
        return SR_OK;
    }

    // get_texture_and_sampler (src/scene.rs:638-690) + TextureCache
    int texref(const JVal *info, int type, sr_texref &ref) {
        ref.texture = -1;
        ref.wrap_s = ref.wrap_t = SR_WRAP_REPEAT;
        if (!info) return SR_OK;
        const JVal *texs = doc.get("textures");
        const JVal *t = texs ? texs->at((size_t)info->integer("index", -1)) : nullptr;
        if (!t) return bad("glTF material refers to a missing texture");
        const JVal *imgs = doc.get("images");
        const JVal *img = imgs ? imgs->at((size_t)t->integer("source", -1)) : nullptr;
        if (!img) return SR_OK;
        std::string uri = img->str("uri");
        if (uri.empty()) return SR_OK;   // bufferView image: no texture (src/scene.rs:686)
        if (t->get("sampler")) {
            const JVal *ss = doc.get("samplers");
            const JVal *s = ss ? ss->at((size_t)t->integer("sampler", -1)) : nullptr;
            if (s) {
                auto wrap = [](int w) {
                    return w == 33071 ? SR_WRAP_CLAMP_TO_EDGE : (w == 33648 ? SR_WRAP_MIRRORED_REPEAT : SR_WRAP_REPEAT);
                };
                ref.wrap_s = wrap(s->integer("wrapS", 10497));
                ref.wrap_t = wrap(s->integer("wrapT", 10497));
            }
        }
        auto it = tex_by_uri.find(uri);
        if (it != tex_by_uri.end()) {
            ref.texture = it->second;
            return SR_OK;
        }
        if (uri.compare(0, 5, "data:") == 0)
            return fail(SR_ERR_UNSUPPORTED, "glTF image data URIs are not loaded (textures are files next to the glTF)");
        std::string path = dir + uri;
        std::vector<uint8_t> bytes;
        if (!read_file(path, bytes)) return fail(SR_ERR_NOT_FOUND, "could not load texture '%s' from path '%s'", uri.c_str(), path.c_str());
        uint32_t w = 0, h = 0;
        std::vector<uint32_t> px;
        std::string err;
        bool jpeg = bytes.size() >= 3 && bytes[0] == 0xFF && bytes[1] == 0xD8 && bytes[2] == 0xFF;
        bool ok = jpeg ? jpeg_decode_rgba8(bytes.data(), bytes.size(), w, h, px, err)
                       : png_decode_rgba8(bytes.data(), bytes.size(), w, h, px, err);
        if (!ok)
            return fail(SR_ERR_INVALID_ARGUMENT, "texture '%s': %s: %s", uri.c_str(), jpeg ? "JPEG" : "PNG", err.c_str());
        sr_texture_desc td;
        td.width = w;
        td.height = h;
        td.type = type;
        td.texels = nullptr;
        g->texels.push_back(std::move(px));
        g->textures.push_back(td);
        ref.texture = (int32_t)g->textures.size() - 1;
        tex_by_uri[uri] = ref.texture;
        return SR_OK;
    }

    int load_materials() {
        const JVal *ms = doc.get("materials");
        for (size_t i = 0; ms && i < ms->size(); i++) {
            const JVal &m = *ms->at(i);
            sr_material_desc d;
            memset(&d, 0, sizeof(d));
            const JVal *pbr = m.get("pbrMetallicRoughness");
            static const JVal empty;
            if (!pbr) pbr = &empty;
            for (int k = 0; k < 4; k++) d.base_color_factor[k] = 1.0f;
            if (const JVal *bc = pbr->get("baseColorFactor"))
                for (int k = 0; k < 4 && k < (int)bc->size(); k++) d.base_color_factor[k] = (float)bc->at(k)->n;
            d.metallic_factor = (float)pbr->num("metallicFactor", 1.0);
            d.roughness_factor = (float)pbr->num("roughnessFactor", 1.0);
            if (const JVal *e = m.get("emissiveFactor"))
                for (int k = 0; k < 3 && k < (int)e->size(); k++) d.emissive_factor[k] = (float)e->at(k)->n;
            const JVal *occ = m.get("occlusionTexture");
            d.occlusion_strength = occ ? (float)occ->num("strength", 1.0) : 1.0f;
            d.alpha_cutoff = (float)m.num("alphaCutoff", 0.5);
            std::string mode = m.str("alphaMode");
            d.alpha_mode = mode == "MASK" ? SR_ALPHA_MASK : (mode == "BLEND" ? SR_ALPHA_BLEND : SR_ALPHA_OPAQUE);
            const JVal *ext = m.get("extensions");
            const JVal *tr = ext ? ext->get("KHR_materials_transmission") : nullptr;
            d.is_translucent = tr ? 1 : 0;
            d.transmission = tr ? (float)tr->num("transmissionFactor", 0.0) : 0.0f;
            int rc = texref(pbr->get("baseColorTexture"), SR_TEX_SRGB, d.base_color);
            if (!rc) rc = texref(pbr->get("metallicRoughnessTexture"), SR_TEX_METALLIC_ROUGHNESS, d.metallic_roughness);
            if (!rc) rc = texref(m.get("normalTexture"), SR_TEX_NORMAL, d.normal);
            if (!rc) rc = texref(m.get("emissiveTexture"), SR_TEX_SRGB, d.emissive);
            if (!rc) rc = texref(occ, SR_TEX_LINEAR, d.occlusion);
            if (!rc) rc = texref(tr ? tr->get("transmissionTexture") : nullptr, SR_TEX_LINEAR, d.transmission_texture);
            if (rc) return rc;
            g->materials.push_back(d);
        }
        if (g->materials.empty()) {
            // no materials: primitives take index 0, so one glTF-default material
            sr_material_desc d;
            memset(&d, 0, sizeof(d));
            for (int k = 0; k < 4; k++) d.base_color_factor[k] = 1.0f;
            d.metallic_factor = d.roughness_factor = 1.0f;
            d.occlusion_strength = 1.0f;
            d.alpha_cutoff = 0.5f;
            sr_texref none = {-1, SR_WRAP_REPEAT, SR_WRAP_REPEAT};
            d.base_color = d.metallic_roughness = d.normal = d.emissive = d.occlusion = d.transmission_texture = none;
            g->materials.push_back(d);
        }
        for (const sr_primitive_desc &p : g->prims)
            if (p.material < 0 || p.material >= (int)g->materials.size())
                return bad("glTF primitive refers to missing material %s", std::to_string(p.material));
        return SR_OK;
    }

    int parse(const std::string &path) {
        size_t slash = path.find_last_of('/');
        dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
        std::vector<uint8_t> bytes;
        if (!read_file(path, bytes)) return fail(SR_ERR_NOT_FOUND, "cannot read '%s'", path.c_str());
        const char *json = (const char *)bytes.data();
        size_t json_len = bytes.size();
        if (bytes.size() >= 12 && !memcmp(bytes.data(), "glTF", 4)) {
            // GLB: header, JSON chunk, optional BIN chunk
            auto u32 = [&](size_t o) { uint32_t v; memcpy(&v, bytes.data() + o, 4); return v; };
            size_t pos = 12, total = std::min<size_t>(u32(8), bytes.size());
            json = nullptr;
            while (pos + 8 <= total) {
                uint32_t len = u32(pos), type = u32(pos + 4);
                if (pos + 8 + (size_t)len > total) return bad("truncated GLB chunk in '%s'", path);
                if (type == 0x4E4F534Au && !json) {
                    json = (const char *)bytes.data() + pos + 8;
                    json_len = len;
                } else if (type == 0x004E4942u && !has_glb_bin) {
                    glb_bin.assign(bytes.data() + pos + 8, bytes.data() + pos + 8 + len);
                    has_glb_bin = true;
                }
                pos += 8 + (size_t)((len + 3) & ~3u);
            }
            if (!json) return bad("GLB file '%s' has no JSON chunk", path);
        }
        JsonParser jp{json, json + json_len};
        if (!jp.value(doc)) return bad("glTF JSON: %s", jp.err);
        if (doc.type != JVal::Obj) return bad("glTF JSON root is not an object in '%s'", path);
        int rc = load_buffers();
        if (!rc) rc = load_meshes();
        if (!rc) rc = load_nodes();
        if (!rc) rc = load_materials();
        return rc;
    }
};

}  // namespace

extern "C" {

SR_API int sr_gltf_load(const char *path, sr_gltf **out) {
    if (!path || !out) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::unique_ptr<sr_gltf> g(new sr_gltf());
    Loader L;
    L.g = g.get();
    int rc = L.parse(path);
    if (rc) return rc;
    // default sun (src/scene.rs:236-239): normalize(-0.2, 1, 0.5), (1, 1, 0.95) * 5
    float d[3] = {-0.2f, 1.0f, 0.5f};
    float r = 1.0f / std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    for (int k = 0; k < 3; k++) g->light.direction[k] = d[k] * r;
    g->light.color[0] = 5.0f;
    g->light.color[1] = 5.0f;
    g->light.color[2] = 0.95f * 5.0f;
    *out = g.release();
    return SR_OK;
}

SR_API int sr_gltf_scene_desc(sr_gltf *g, sr_scene_desc *desc) {
    if (!g || !desc) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    for (size_t i = 0; i < g->prims.size(); i++) {
        sr_primitive_desc &p = g->prims[i];
        p.positions = g->positions[i].data();
        p.normals = g->normals[i].data();
        p.texcoords = g->texcoords[i].data();
        p.tangents = g->tangents[i].empty() ? nullptr : g->tangents[i].data();
        p.indices = g->indices[i].data();
    }
    for (size_t i = 0; i < g->textures.size(); i++) g->textures[i].texels = g->texels[i].data();
    memset(desc, 0, sizeof(*desc));
    desc->primitives = g->prims.data();
    desc->num_primitives = (int32_t)g->prims.size();
    desc->meshes = g->meshes.data();
    desc->num_meshes = (int32_t)g->meshes.size();
    desc->nodes = g->nodes.data();
    desc->num_nodes = (int32_t)g->nodes.size();
    desc->materials = g->materials.data();
    desc->num_materials = (int32_t)g->materials.size();
    desc->textures = g->textures.data();
    desc->num_textures = (int32_t)g->textures.size();
    desc->light = g->light;
    return SR_OK;
}

SR_API int sr_gltf_destroy(sr_gltf *g) {
    delete g;
    return SR_OK;
}

SR_API int sr_png_decode(const uint8_t *data, size_t size, uint32_t *width, uint32_t *height, uint32_t *texels,
                         size_t capacity) {
    if (!data || !width || !height) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t w = 0, h = 0;
    std::vector<uint32_t> px;
    std::string err;
    if (!png_decode_rgba8(data, size, w, h, px, err)) return fail(SR_ERR_INVALID_ARGUMENT, "PNG: %s", err.c_str());
    *width = w;
    *height = h;
    if (texels) {
        if (capacity < px.size()) return fail(SR_ERR_INVALID_ARGUMENT, "texel buffer holds %zu, image has %zu", capacity, px.size());
        memcpy(texels, px.data(), px.size() * 4);
    }
    return SR_OK;
}

SR_API int sr_jpeg_decode(const uint8_t *data, size_t size, uint32_t *width, uint32_t *height, uint32_t *texels,
                          size_t capacity) {
    if (!data || !width || !height) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t w = 0, h = 0;
    std::vector<uint32_t> px;
    std::string err;
    if (!jpeg_decode_rgba8(data, size, w, h, px, err)) return fail(SR_ERR_INVALID_ARGUMENT, "JPEG: %s", err.c_str());
    *width = w;
    *height = h;
    if (texels) {
        if (capacity < px.size()) return fail(SR_ERR_INVALID_ARGUMENT, "texel buffer holds %zu, image has %zu", capacity, px.size());
        memcpy(texels, px.data(), px.size() * 4);
    }
    return SR_OK;
}

}  // extern "C"
