// Host-side loaders for the reference's on-disk formats: .v models, .p
// palettes and .pf prefabs spawned into a flat entity list with world
// matrices, light records and the rebuild order the reference's ECS produces.
//
// Reference: VoxAsset.h:42-50 (.v), PalleteAsset.h:63-69 (.p),
// PrefabAsset.cpp:30-142 (Spawn), TransformSystem.cpp:124-135
// (RealculateMatrix), Assets.h:207-210 + Mod.h:86-94 (GUIDs of mod paths),
// WorldRenderer.cpp:243-257 (light draws), ShadowVoxSystem.cpp:116-120
// (rebuild iteration).
//
// No device code here: vx_model_load goes through the public model calls.
#include <charconv>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <map>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <sys/stat.h>
#include <unordered_map>
#include <vector>

#include "vx/vx_api.h"
#include "vx_error.hpp"

namespace {

vx_status failf(vx_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
vx_status failf(vx_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return vx_set_error(code, buf);
}

// Runs an entry point's body; C++ exceptions do not cross the C ABI.
template <class F>
vx_status guarded(const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return failf(VX_ERR_OUT_OF_MEMORY, "%s: host allocation failed", what);
    } catch (const std::exception& e) {
        return failf(VX_ERR_INVALID_ARGUMENT, "%s: %s", what, e.what());
    } catch (...) {
        return failf(VX_ERR_INVALID_ARGUMENT, "%s failed", what);
    }
}

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    bool ok = !ferror(f);
    fclose(f);
    return ok;
}

uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

// ---------------------------------------------------------------------------
// A small JSON reader: the subset json11 produces and parses (objects,
// arrays, strings with escapes, numbers, true/false/null). Object members
// keep their keys sorted, as json11's std::map does.
// ---------------------------------------------------------------------------
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json& operator[](const std::string& k) const {
        static const Json null_value;
        if (kind != OBJ) return null_value;
        auto it = obj.find(k);
        return it == obj.end() ? null_value : it->second;
    }
    bool is_null() const { return kind == NUL; }
    // json11 Json::int_value / number_value / string_value: 0 or "" for other kinds
    int int_value() const { return kind == NUM ? (int)num : 0; }
    double number_value() const { return kind == NUM ? num : 0.0; }
    const std::string& string_value() const {
        static const std::string empty;
        return kind == STR ? str : empty;
    }
};

struct JsonParser {
    const char* p;
    const char* end;
    std::string err;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* what) {
        if (err.empty()) err = what;
        return false;
    }
    bool lit(const char* s) {
        size_t n = strlen(s);
        if ((size_t)(end - p) < n || memcmp(p, s, n) != 0) return fail("bad literal");
        p += n;
        return true;
    }
    static void put_utf8(std::string& o, uint32_t c) {
        if (c < 0x80) {
            o += (char)c;
        } else if (c < 0x800) {
            o += (char)(0xC0 | c >> 6);
            o += (char)(0x80 | (c & 63));
        } else if (c < 0x10000) {
            o += (char)(0xE0 | c >> 12);
            o += (char)(0x80 | (c >> 6 & 63));
            o += (char)(0x80 | (c & 63));
        } else {
            o += (char)(0xF0 | c >> 18);
            o += (char)(0x80 | (c >> 12 & 63));
            o += (char)(0x80 | (c >> 6 & 63));
            o += (char)(0x80 | (c & 63));
        }
    }
    bool hex4(uint32_t& v) {
        if (end - p < 4) return fail("short \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& o) {
        ++p;  // opening quote
        o.clear();
        while (p < end && *p != '"') {
            char c = *p++;
            if (c != '\\') {
                o += c;
                continue;
            }
            if (p >= end) return fail("unterminated escape");
            char e = *p++;
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t c1;
                    if (!hex4(c1)) return false;
                    if (c1 >= 0xD800 && c1 < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t c2;
                        if (!hex4(c2)) return false;
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    put_utf8(o, c1);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool number(double& v) {
        const char* s = p;
        if (p < end && *p == '-') ++p;
        while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-'))
            ++p;
        std::string t(s, p);
        char* q = nullptr;
        v = strtod(t.c_str(), &q);
        if (t.empty() || q != t.c_str() + t.size()) return fail("bad number");
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 200) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        char c = *p;
        if (c == '{') {
            v.kind = Json::OBJ;
            ++p;
            ws();
            if (p < end && *p == '}') {
                ++p;
                return true;
            }
            for (;;) {
                ws();
                if (p >= end || *p != '"') return fail("expected key");
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                Json m;
                if (!value(m, depth + 1)) return false;
                v.obj[k] = std::move(m);
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::ARR;
            ++p;
            ws();
            if (p < end && *p == ']') {
                ++p;
                return true;
            }
            for (;;) {
                Json m;
                if (!value(m, depth + 1)) return false;
                v.arr.push_back(std::move(m));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::STR;
            return string(v.str);
        }
        if (c == 't') {
            v.kind = Json::BOOL;
            v.b = true;
            return lit("true");
        }
        if (c == 'f') {
            v.kind = Json::BOOL;
            return lit("false");
        }
        if (c == 'n') return lit("null");
        v.kind = Json::NUM;
        return number(v.num);
    }
};

bool parse_json(const std::vector<uint8_t>& text, Json& out, std::string& err) {
    JsonParser jp{(const char*)text.data(), (const char*)text.data() + text.size(), {}};
    if (!jp.value(out, 0)) {
        err = jp.err;
        return false;
    }
    jp.ws();
    if (jp.p != jp.end) {
        err = "trailing characters";
        return false;
    }
    return true;
}

// readVec3 (PrefabAsset.cpp:8-18): three floats separated by single spaces,
// each parsed with std::from_chars (correctly rounded decimal to float).
bool read_vec3(const std::string& s, float v[3]) {
    size_t s1 = s.find(' ');
    if (s1 == std::string::npos) return false;
    size_t s2 = s.find(' ', s1 + 1);
    if (s2 == std::string::npos) return false;
    const char* b = s.data();
    const char* spans[3][2] = {{b, b + s1}, {b + s1 + 1, b + s2}, {b + s2 + 1, b + s.size()}};
    for (int i = 0; i < 3; ++i) {
        auto r = std::from_chars(spans[i][0], spans[i][1], v[i]);
        if (r.ec != std::errc()) return false;
    }
    return true;
}

// readGUID (PrefabAsset.cpp:24-28): hexadecimal digits up to the first other
// character; an empty or invalid string gives the null GUID 0.
uint64_t read_guid(const std::string& s) {
    uint64_t g = 0;
    auto r = std::from_chars(s.data(), s.data() + s.size(), g, 16);
    return r.ec == std::errc() ? g : 0;
}

// ---------------------------------------------------------------------------
// glm-order fp32 matrix arithmetic (column-major, m[c*4 + r]); each vec4
// expression is evaluated left to right as glm's scalar path does.
// ---------------------------------------------------------------------------
void mat_identity(float* m) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

// translate(m, v): column 3 = m0*v.x + m1*v.y + m2*v.z + m3
void mat_translate(float* m, const float v[3]) {
    float r[4];
    for (int k = 0; k < 4; ++k) r[k] = m[0 + k] * v[0] + m[4 + k] * v[1] + m[8 + k] * v[2] + m[12 + k];
    for (int k = 0; k < 4; ++k) m[12 + k] = r[k];
}

// rotate(m, angle, axis) for a unit basis axis: the axis-angle matrix
// R[c][r] = c + (1-c)a_c a_r, +-s a_k off the diagonal, then
// column j of the result = m0*R[j][0] + m1*R[j][1] + m2*R[j][2].
void mat_rotate(float* m, float angle, const float axis_in[3]) {
    const float c = cosf(angle);
    const float s = sinf(angle);
    const float len2 = axis_in[0] * axis_in[0] + axis_in[1] * axis_in[1] + axis_in[2] * axis_in[2];
    const float inv = 1.0f / sqrtf(len2);
    const float a[3] = {axis_in[0] * inv, axis_in[1] * inv, axis_in[2] * inv};
    const float one_c = 1.0f - c;
    const float t[3] = {one_c * a[0], one_c * a[1], one_c * a[2]};
    float R[3][3];
    R[0][0] = c + t[0] * a[0];
    R[0][1] = t[0] * a[1] + s * a[2];
    R[0][2] = t[0] * a[2] - s * a[1];
    R[1][0] = t[1] * a[0] - s * a[2];
    R[1][1] = c + t[1] * a[1];
    R[1][2] = t[1] * a[2] + s * a[0];
    R[2][0] = t[2] * a[0] + s * a[1];
    R[2][1] = t[2] * a[1] - s * a[0];
    R[2][2] = c + t[2] * a[2];
    float out[12];
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 4; ++k) out[j * 4 + k] = m[0 + k] * R[j][0] + m[4 + k] * R[j][1] + m[8 + k] * R[j][2];
    for (int i = 0; i < 12; ++i) m[i] = out[i];
}

// scale(m, v): column i *= v_i for i < 3
void mat_scale(float* m, const float v[3]) {
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 4; ++k) m[c * 4 + k] = m[c * 4 + k] * v[c];
}

// a * b: column j = a0*b[j].x + a1*b[j].y + a2*b[j].z + a3*b[j].w
void mat_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k)
            r[j * 4 + k] = a[0 + k] * b[j * 4 + 0] + a[4 + k] * b[j * 4 + 1] + a[8 + k] * b[j * 4 + 2] +
                           a[12 + k] * b[j * 4 + 3];
    memcpy(out, r, sizeof(r));
}

}  // namespace

// ---------------------------------------------------------------------------
// GUIDs, .v and .p
// ---------------------------------------------------------------------------
extern "C" uint64_t vx_asset_guid(const char* path, size_t len) {
    // 64-bit FNV-1a: offset basis 14695981039346656037, prime 1099511628211.
    uint64_t h = 14695981039346656037ull;
    for (size_t i = 0; path && i < len; ++i) {
        h ^= (uint8_t)path[i];
        h *= 1099511628211ull;
    }
    return h;
}

extern "C" vx_status vx_vox_parse(const uint8_t* bytes, size_t nbytes, int32_t dims[3], size_t* data_offset) {
    if (!bytes || !dims) return failf(VX_ERR_INVALID_ARGUMENT, "vx_vox_parse: NULL argument");
    if (nbytes < 12) return failf(VX_ERR_INVALID_ARGUMENT, ".v data of %zu bytes has no header", nbytes);
    uint64_t d[3] = {le32(bytes), le32(bytes + 4), le32(bytes + 8)};
    for (int i = 0; i < 3; ++i)
        if (d[i] == 0 || d[i] > 4096)
            return failf(VX_ERR_RANGE, ".v size %llu on axis %d outside 1..4096", (unsigned long long)d[i], i);
    const uint64_t need = d[0] * d[1] * d[2];
    if (nbytes - 12 < need)
        return failf(VX_ERR_INVALID_ARGUMENT, ".v data holds %zu voxel bytes, %llux%llux%llu needs %llu",
                     nbytes - 12, (unsigned long long)d[0], (unsigned long long)d[1], (unsigned long long)d[2],
                     (unsigned long long)need);
    for (int i = 0; i < 3; ++i) dims[i] = (int32_t)d[i];
    if (data_offset) *data_offset = 12;
    return VX_OK;
}

extern "C" vx_status vx_model_load(vx_ctx* ctx, const char* path, vx_model** out, void* stream) {
    if (!ctx || !path || !out) return failf(VX_ERR_INVALID_ARGUMENT, "vx_model_load: NULL argument");
    *out = nullptr;
    return guarded("vx_model_load", [&]() -> vx_status {
        std::vector<uint8_t> b;
        if (!read_file(path, b)) return failf(VX_ERR_INVALID_ARGUMENT, "cannot read %s", path);
        int32_t d[3];
        size_t off = 0;
        vx_status st = vx_vox_parse(b.data(), b.size(), d, &off);
        if (st != VX_OK) return st;
        if ((d[0] | d[1] | d[2]) & 3)
            return failf(VX_ERR_UNSUPPORTED, "%s: sizes %dx%dx%d are not multiples of 4", path, d[0], d[1], d[2]);
        vx_model* m = nullptr;
        if ((st = vx_model_create(ctx, d[0], d[1], d[2], &m)) != VX_OK) return st;
        const size_t n = (size_t)d[0] * d[1] * d[2];
        if ((st = vx_model_upload(m, b.data() + off, n, stream)) != VX_OK || (st = vx_model_build_mips(m, stream)) != VX_OK) {
            vx_model_destroy(m);
            return st;
        }
        *out = m;
        return VX_OK;
    });
}

extern "C" vx_status vx_palette_parse(const uint8_t* bytes, size_t nbytes, float* palette) {
    if (!bytes || !palette) return failf(VX_ERR_INVALID_ARGUMENT, "vx_palette_parse: NULL argument");
    if (nbytes < 256 * 7) return failf(VX_ERR_INVALID_ARGUMENT, ".p data of %zu bytes, need 1792", nbytes);
    for (int i = 0; i < 256; ++i) {
        const uint8_t* m = bytes + 7 * i;  // r g b a roughness metallic emit
        const uint8_t src[6] = {m[0], m[1], m[2], m[4], m[5], m[6]};
        for (int k = 0; k < 6; ++k) palette[6 * i + k] = (float)src[k] / 255.0f;
    }
    return VX_OK;
}

extern "C" vx_status vx_palette_load(const char* path, float* palette) {
    if (!path || !palette) return failf(VX_ERR_INVALID_ARGUMENT, "vx_palette_load: NULL argument");
    return guarded("vx_palette_load", [&]() -> vx_status {
        std::vector<uint8_t> b;
        if (!read_file(path, b)) return failf(VX_ERR_INVALID_ARGUMENT, "cannot read %s", path);
        return vx_palette_parse(b.data(), b.size(), palette);
    });
}

// ---------------------------------------------------------------------------
// Prefabs
// ---------------------------------------------------------------------------
struct vx_prefab {
    std::string mods_dir;
    std::unordered_map<uint64_t, std::string> assets;  // GUID -> mod-relative path
    std::vector<vx_prefab_entity> ents;
    std::vector<std::string> files;                     // mod-relative paths of referenced .v/.p
    std::unordered_map<uint64_t, int32_t> file_index;
    std::vector<int32_t> vox_pool;                      // entity indices in VoxRenderer creation order
    std::vector<int32_t> light_pool;                    // entity indices in Light creation order
};

namespace {

void scan_dir(const std::string& root, const std::string& rel, std::unordered_map<uint64_t, std::string>& out,
              int depth) {
    if (depth > 32) return;
    std::string dir = rel.empty() ? root : root + "/" + rel;
    DIR* d = opendir(dir.c_str());
    if (!d) return;
    while (dirent* e = readdir(d)) {
        std::string name = e->d_name;
        if (name == "." || name == "..") continue;
        std::string r = rel.empty() ? name : rel + "/" + name;
        struct stat st;
        if (stat((root + "/" + r).c_str(), &st) != 0) continue;
        if (S_ISDIR(st.st_mode)) scan_dir(root, r, out, depth + 1);
        else out.emplace(vx_asset_guid(r.data(), r.size()), r);
    }
    closedir(d);
}

int32_t file_ref(vx_prefab* P, uint64_t guid) {
    auto it = P->file_index.find(guid);
    if (it != P->file_index.end()) return it->second;
    auto a = P->assets.find(guid);
    if (a == P->assets.end()) return -1;
    int32_t i = (int32_t)P->files.size();
    P->files.push_back(a->second);
    P->file_index.emplace(guid, i);
    return i;
}

struct SpawnErr {
    std::string msg;
};

void set_parent(vx_prefab* P, int32_t child, int32_t parent) { P->ents[child].parent = parent; }

// PrefabAsset::Spawn (PrefabAsset.cpp:30-142): returns the root entity.
int32_t spawn(vx_prefab* P, const std::string& rel, int32_t parent_entity, int depth) {
    if (depth > 16) throw SpawnErr{"prefab instances nest deeper than 16 (cycle?) at " + rel};
    std::vector<uint8_t> text;
    if (!read_file(P->mods_dir + "/" + rel, text)) throw SpawnErr{"cannot read " + P->mods_dir + "/" + rel};
    Json J;
    std::string err;
    if (!parse_json(text, J, err)) throw SpawnErr{rel + ": JSON " + err};
    if (J.kind != Json::ARR) throw SpawnErr{rel + ": top level is not an array"};
    std::unordered_map<int, int32_t> id_map;
    int32_t root = -1;
    for (const Json& E : J.arr) {
        int32_t parent = -1;
        if (!E["Parent"].is_null()) {
            auto it = id_map.find(E["Parent"].int_value());
            if (it == id_map.end()) throw SpawnErr{rel + ": Parent refers to an entity not yet spawned"};
            parent = it->second;
        }
        const bool is_root = parent < 0;
        int32_t e;
        if (E["Instance"].is_null()) {
            e = (int32_t)P->ents.size();
            vx_prefab_entity z;
            memset(&z, 0, sizeof(z));
            z.parent = -1;
            z.model = z.palette = -1;
            z.light_type = -1;
            P->ents.push_back(z);
        } else {
            const uint64_t g = read_guid(E["Instance"].string_value());
            auto a = P->assets.find(g);
            if (g == 0 || a == P->assets.end()) throw SpawnErr{rel + ": Instance " + E["Instance"].string_value() + " is not in the mod folder"};
            e = spawn(P, a->second, parent, depth + 1);
            P->ents[e].flags |= VX_ENT_INSTANCE;
        }
        P->ents[e].id = E["Id"].int_value();
        id_map[E["Id"].int_value()] = e;
        if (is_root) {
            root = e;
            set_parent(P, e, parent_entity);
        } else {
            set_parent(P, e, parent);
        }
        // components in the (sorted) key order json11 iterates
        for (const auto& kv : E.obj) {
            const Json& C = kv.second;
            vx_prefab_entity& x = P->ents[e];
            if (kv.first == "Transform") {
                if (!read_vec3(C["Position"].string_value(), x.position) ||
                    !read_vec3(C["Rotation"].string_value(), x.rotation) ||
                    !read_vec3(C["Scale"].string_value(), x.scale))
                    throw SpawnErr{rel + ": malformed Transform vector"};
                x.flags |= VX_ENT_TRANSFORM;
            } else if (kv.first == "VoxRenderer") {
                x.palette_guid = read_guid(C["Pallete"].string_value());
                x.vox_guid = read_guid(C["Vox"].string_value());
                if (!C["Pivot"].is_null() && !read_vec3(C["Pivot"].string_value(), x.pivot))
                    throw SpawnErr{rel + ": malformed Pivot"};
                x.model = file_ref(P, x.vox_guid);
                x.palette = file_ref(P, x.palette_guid);
                if (!(x.flags & VX_ENT_VOX)) P->vox_pool.push_back(e);
                x.flags |= VX_ENT_VOX;
            } else if (kv.first == "Light") {
                x.light_type = C["LightType"].int_value();
                x.light_intensity = (float)C["Intensity"].number_value();
                if (!read_vec3(C["Color"].string_value(), x.light_color)) throw SpawnErr{rel + ": malformed Light Color"};
                x.light_attenuation = (float)C["Attenuation"].number_value();
                x.light_range = (float)C["Range"].number_value();
                x.light_angle = (float)C["Angle"].number_value();
                x.light_angle_attenuation = (float)C["AngleAttenuation"].number_value();
                if (!(x.flags & VX_ENT_LIGHT)) P->light_pool.push_back(e);
                x.flags |= VX_ENT_LIGHT;
            }
        }
    }
    if (root < 0) throw SpawnErr{rel + ": no root entity"};
    return root;
}

}  // namespace

extern "C" vx_status vx_prefab_load(const char* mods_dir, const char* path, vx_prefab** out) {
    if (!mods_dir || !path || !out) return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_load: NULL argument");
    *out = nullptr;
    return guarded("vx_prefab_load", [&]() -> vx_status {
        std::unique_ptr<vx_prefab> P(new vx_prefab());
        P->mods_dir = mods_dir;
        scan_dir(P->mods_dir, "", P->assets, 0);
        try {
            spawn(P.get(), path, -1, 0);
        } catch (const SpawnErr& e) {
            return failf(VX_ERR_INVALID_ARGUMENT, "%s", e.msg.c_str());
        }
        // TransformSystem::OnUpdate (TransformSystem.cpp:36-70): entities in
        // creation order (the Dirty sort by entity id); RealculateMatrix
        // (TransformSystem.cpp:124-135) with the parent's world matrix, or the
        // identity for an entity without a parent transform.
        for (size_t i = 0; i < P->ents.size(); ++i) {
            vx_prefab_entity& x = P->ents[i];
            if (!(x.flags & VX_ENT_TRANSFORM)) {
                // World::Create adds no Transform; such entities keep identity matrices
                mat_identity(x.matrix);
                mat_identity(x.world);
                continue;
            }
            static const float kZ[3] = {0, 0, 1}, kY[3] = {0, 1, 0}, kX[3] = {1, 0, 0};
            mat_identity(x.matrix);
            mat_translate(x.matrix, x.position);
            mat_rotate(x.matrix, x.rotation[2], kZ);
            mat_rotate(x.matrix, x.rotation[1], kY);
            mat_rotate(x.matrix, x.rotation[0], kX);
            mat_scale(x.matrix, x.scale);
            float parent[16];
            if (x.parent >= 0 && (P->ents[x.parent].flags & VX_ENT_TRANSFORM)) memcpy(parent, P->ents[x.parent].world, sizeof(parent));
            else mat_identity(parent);
            mat_mul(parent, x.matrix, x.world);
        }
        *out = P.release();
        return VX_OK;
    });
}

extern "C" vx_status vx_prefab_destroy(vx_prefab* p) {
    delete p;
    return VX_OK;
}

extern "C" vx_status vx_prefab_counts(const vx_prefab* p, uint32_t* nentities, uint32_t* nfiles) {
    if (!p) return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_counts: NULL prefab");
    if (nentities) *nentities = (uint32_t)p->ents.size();
    if (nfiles) *nfiles = (uint32_t)p->files.size();
    return VX_OK;
}

extern "C" vx_status vx_prefab_entities(const vx_prefab* p, vx_prefab_entity* out, uint32_t cap) {
    if (!p || (!out && cap)) return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_entities: NULL argument");
    if (cap < p->ents.size()) return failf(VX_ERR_RANGE, "capacity %u below %zu entities", cap, p->ents.size());
    if (!p->ents.empty()) memcpy(out, p->ents.data(), p->ents.size() * sizeof(vx_prefab_entity));
    return VX_OK;
}

extern "C" vx_status vx_prefab_file(const vx_prefab* p, uint32_t i, char* buf, size_t cap) {
    if (!p || !buf) return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_file: NULL argument");
    if (i >= p->files.size()) return failf(VX_ERR_RANGE, "file index %u of %zu", i, p->files.size());
    const std::string full = p->mods_dir + "/" + p->files[i];
    if (full.size() + 1 > cap) return failf(VX_ERR_RANGE, "buffer of %zu bytes below %zu", cap, full.size() + 1);
    memcpy(buf, full.c_str(), full.size() + 1);
    return VX_OK;
}

// EnTT 3.6 views walk a pool's packed array from its last element to its
// first (sparse_set::begin starts at packed.size()), led by the smallest pool.
extern "C" vx_status vx_prefab_rebuild_order(const vx_prefab* p, uint32_t* idx, uint32_t cap, uint32_t* n) {
    if (!p || !n || (!idx && cap)) return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_rebuild_order: NULL argument");
    uint32_t k = 0;
    for (size_t j = p->vox_pool.size(); j-- > 0;) {
        const vx_prefab_entity& x = p->ents[p->vox_pool[j]];
        if (!(x.flags & VX_ENT_TRANSFORM)) continue;  // the view needs Transform too
        if (k < cap) idx[k] = (uint32_t)p->vox_pool[j];
        ++k;
    }
    *n = k;
    if (k > cap) return failf(VX_ERR_RANGE, "capacity %u below %u instances", cap, k);
    return VX_OK;
}

extern "C" vx_status vx_prefab_lights(const vx_prefab* p, vx_light* out, uint32_t cap, uint32_t* n) {
    if (!p || !n || (!out && cap)) return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_lights: NULL argument");
    uint32_t k = 0;
    // the point pipeline draws first, then the spot pipeline (WorldRenderer.cpp:243-257)
    for (int type = VX_LIGHT_POINT; type <= VX_LIGHT_SPOT; ++type) {
        for (size_t j = p->light_pool.size(); j-- > 0;) {
            const vx_prefab_entity& x = p->ents[p->light_pool[j]];
            if (!(x.flags & VX_ENT_TRANSFORM) || x.light_type != type) continue;
            if (k < cap) {
                vx_light& L = out[k];
                memset(&L, 0, sizeof(L));
                L.type = type;
                for (int c = 0; c < 3; ++c) {
                    L.position[c] = x.world[12 + c];                       // WorldMatrix[3]
                    L.color[c] = x.light_color[c] * x.light_intensity;     // Color * Intensity
                    L.direction[c] = x.world[8 + c];                       // WorldMatrix[2]
                }
                L.range = x.light_range;
                L.attenuation = x.light_attenuation;
                L.angle = x.light_angle;
                L.angle_attenuation = x.light_angle_attenuation;
            }
            ++k;
        }
    }
    *n = k;
    if (k > cap) return failf(VX_ERR_RANGE, "capacity %u below %u lights", cap, k);
    return VX_OK;
}

// WorldRenderer.cpp:210-220: the owning group<VoxRenderer>(get<Transform>)
// walks the VoxRenderer pool from its last element; each draw's matrices are
// translated by -Pivot (glm::translate, fp32).
extern "C" vx_status vx_prefab_vox_cmds(const vx_prefab* p, const vx_model* const* models, uint32_t nmodels,
                                        const float* prev, vx_vox_cmd* out, uint32_t cap, uint32_t* n) {
    if (!p || !n || (!out && cap) || (!models && nmodels))
        return failf(VX_ERR_INVALID_ARGUMENT, "vx_prefab_vox_cmds: NULL argument");
    uint32_t k = 0;
    for (size_t j = p->vox_pool.size(); j-- > 0;) {
        const int32_t e = p->vox_pool[j];
        const vx_prefab_entity& x = p->ents[e];
        if (!(x.flags & VX_ENT_TRANSFORM) || x.model < 0 || x.palette < 0) continue;
        if ((uint32_t)x.model >= nmodels || !models[x.model])
            return failf(VX_ERR_INVALID_ARGUMENT, "model of file %d (%s) is not given", x.model,
                         p->files[x.model].c_str());
        if (k < cap) {
            vx_vox_cmd& c = out[k];
            memset(&c, 0, sizeof(c));
            c.model = models[x.model];
            const float neg[3] = {-x.pivot[0], -x.pivot[1], -x.pivot[2]};
            memcpy(c.world, x.world, sizeof(c.world));
            mat_translate(c.world, neg);
            if (prev) memcpy(c.last_world, prev + 16 * (size_t)e, sizeof(c.last_world));
            else mat_identity(c.last_world);
            mat_translate(c.last_world, neg);
            c.volume_rid = x.model + 1;
            c.palette = x.palette;
        }
        ++k;
    }
    *n = k;
    if (k > cap) return failf(VX_ERR_RANGE, "capacity %u below %u draws", cap, k);
    return VX_OK;
}

static_assert(sizeof(vx_prefab_entity) == 256, "vx_prefab_entity is 256 bytes");
static_assert(sizeof(vx_vox_cmd) == 144, "vx_vox_cmd is 144 bytes");
