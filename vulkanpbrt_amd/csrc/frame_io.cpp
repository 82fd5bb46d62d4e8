// Offline frames (include/pbrt_io.h): camera matrices (JSON and plain lists),
// PFM float images and the G-buffer conversions of source/io/RenderIO.cpp.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pbrt_io.h"
#include "capi_common.hpp"

// the conversions follow the reference's FP32 expressions operation by operation
#pragma clang fp contract(off)

namespace pbrt {
namespace {

std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(PBRT_ERR_INVALID_ARG, "cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// ---------------------------------------------------------------------------
// a small JSON reader (RFC 8259 values; numbers as double)
// ---------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> fields;
    const Json* get(const std::string& k) const {
        for (const auto& f : fields)
            if (f.first == k) return &f.second;
        return nullptr;
    }
};

struct JsonParser {
    const char* p;
    const char* e;
    std::string where;
    [[noreturn]] void bad(const char* what) { fail(PBRT_ERR_INVALID_ARG, where + ": JSON " + what); }
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(e - p) >= n && std::memcmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    std::string string() {
        if (p >= e || *p != '"') bad("expected a string");
        ++p;
        std::string out;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) bad("bad escape");
                const char c = *p++;
                switch (c) {
                    case 'n': out.push_back('\n'); break;
                    case 't': out.push_back('\t'); break;
                    case 'r': out.push_back('\r'); break;
                    case 'b': out.push_back('\b'); break;
                    case 'f': out.push_back('\f'); break;
                    case 'u': {
                        if (e - p < 4) bad("bad escape");
                        unsigned v = 0;
                        std::from_chars(p, p + 4, v, 16);
                        p += 4;
                        if (v < 0x80) out.push_back((char)v);
                        else out.push_back('?');  // names and types here are ASCII
                        break;
                    }
                    default: out.push_back(c);
                }
            } else {
                out.push_back(*p++);
            }
        }
        if (p >= e) bad("unterminated string");
        ++p;
        return out;
    }
    Json value(int depth = 0) {
        if (depth > 64) bad("nested too deeply");
        ws();
        Json v;
        if (p >= e) bad("unexpected end");
        if (*p == '{') {
            ++p;
            v.kind = Json::Object;
            ws();
            if (p < e && *p == '}') {
                ++p;
                return v;
            }
            while (true) {
                ws();
                std::string k = string();
                ws();
                if (p >= e || *p != ':') bad("expected ':'");
                ++p;
                v.fields.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == '}') {
                    ++p;
                    return v;
                }
                bad("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = Json::Array;
            ws();
            if (p < e && *p == ']') {
                ++p;
                return v;
            }
            while (true) {
                v.items.push_back(value(depth + 1));
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == ']') {
                    ++p;
                    return v;
                }
                bad("expected ',' or ']'");
            }
        }
        if (*p == '"') {
            v.kind = Json::String;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Json::Bool;
            return v;
        }
        if (lit("null")) return v;
        v.kind = Json::Number;
        auto r = std::from_chars(p, e, v.num);
        if (r.ec != std::errc()) bad("bad number");
        p = r.ptr;
        return v;
    }
};

// vsg::mat4 from 16 numbers in storage order (RenderIO.cpp:604-616)
void mat_from_json(const Json* j, float* m, const std::string& what) {
    if (!j || j->kind != Json::Array || j->items.size() != 16)
        fail(PBRT_ERR_INVALID_ARG, "matrices: '" + what + "' is not an array of 16 numbers");
    for (int i = 0; i < 16; ++i) {
        if (j->items[(size_t)i].kind != Json::Number) fail(PBRT_ERR_INVALID_ARG, "matrices: non-number in " + what);
        m[i] = (float)j->items[(size_t)i].num;
    }
}

// general 4x4 inverse (column-major) by cofactors in double, rounded to float
void invert(const float* a, float* out) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = a[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) fail(PBRT_ERR_INVALID_ARG, "matrices: singular matrix");
    for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
}

std::vector<pbrt_camera_matrices> read_matrices(const std::string& path) {
    const std::string text = slurp(path);
    std::vector<pbrt_camera_matrices> out;
    const size_t dot = path.find_last_of('.');
    if (dot != std::string::npos && path.substr(dot) == ".json") {
        JsonParser jp{text.data(), text.data() + text.size(), path};
        const Json root = jp.value();
        const Json* n = root.get("amtOfFrames");
        const Json* list = root.get("matrices");
        if (!n || n->kind != Json::Number || !list || list->kind != Json::Array)
            fail(PBRT_ERR_INVALID_ARG, path + ": expected {\"amtOfFrames\": n, \"matrices\": [...]}");
        const size_t frames = (size_t)n->num;
        if ((double)frames != n->num || frames > list->items.size())
            fail(PBRT_ERR_INVALID_ARG, path + ": amtOfFrames does not match the matrices");
        out.resize(frames);
        for (size_t i = 0; i < frames; ++i) {
            const Json& m = list->items[i];
            pbrt_camera_matrices& c = out[i];
            std::memset(&c, 0, sizeof(c));
            mat_from_json(m.get("view"), c.view, "view");
            mat_from_json(m.get("invView"), c.inv_view, "invView");
            const Json* type = m.get("type");
            if (type && type->kind == Json::String && type->str == "ModelView+Projection") {
                mat_from_json(m.get("proj"), c.proj, "proj");
                mat_from_json(m.get("invProj"), c.inv_proj, "invProj");
                c.has_proj = 1;
            }
        }
        return out;
    }
    // plain list: tokens that start with a digit or '-' after dropping a trailing
    // ',' and a leading '{' (RenderIO.cpp:636-662)
    std::istringstream in(text);
    std::string tok;
    float tmp[16];
    int count = 0;
    while (in >> tok) {
        if (!tok.empty() && tok.back() == ',') tok.pop_back();
        if (!tok.empty() && tok.front() == '{') tok = tok.substr(1);
        if (tok.empty() || !(std::isdigit((unsigned char)tok[0]) || tok[0] == '-')) continue;
        float v = 0.0f;
        const char* s = tok.data();
        auto r = std::from_chars(s, s + tok.size(), v);
        if (r.ec != std::errc()) fail(PBRT_ERR_INVALID_ARG, path + ": bad number '" + tok + "'");
        tmp[count++] = v;
        if (count == 16) {
            pbrt_camera_matrices c;
            std::memset(&c, 0, sizeof(c));
            std::memcpy(c.view, tmp, sizeof(tmp));
            invert(tmp, c.inv_view);
            out.push_back(c);
            count = 0;
        }
    }
    return out;
}

// a float as the shortest decimal that reads back as the same double
std::string num(float f) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), (double)f);
    std::string s(buf, r.ptr);
    if (s.find_first_of(".eE") == std::string::npos && s.find("inf") == std::string::npos &&
        s.find("nan") == std::string::npos)
        s += ".0";
    return s;
}

void write_matrices(const std::string& path, const pbrt_camera_matrices* m, uint32_t count) {
    std::ostringstream o;
    auto arr = [&](const float* a, const char* ind) {
        o << "[\n";
        for (int i = 0; i < 16; ++i) o << ind << "    " << num(a[i]) << (i < 15 ? ",\n" : "\n");
        o << ind << "]";
    };
    o << "{\n    \"amtOfFrames\": " << count << ",\n    \"matrices\": [";
    for (uint32_t i = 0; i < count; ++i) {
        const pbrt_camera_matrices& c = m[i];
        o << (i ? ",\n" : "\n") << "        {\n";
        if (c.has_proj) {
            o << "            \"invProj\": ";
            arr(c.inv_proj, "            ");
            o << ",\n";
        }
        o << "            \"invView\": ";
        arr(c.inv_view, "            ");
        o << ",\n";
        if (c.has_proj) {
            o << "            \"proj\": ";
            arr(c.proj, "            ");
            o << ",\n";
        }
        o << "            \"storageType\": \"ColumnMajor\",\n";
        o << "            \"type\": \"" << (c.has_proj ? "ModelView+Projection" : "ModelViewProjection") << "\",\n";
        o << "            \"view\": ";
        arr(c.view, "            ");
        o << "\n        }";
    }
    o << (count ? "\n    ]\n}" : "]\n}");
    std::ofstream f(path, std::ios::binary);
    const std::string s = o.str();
    if (!f || !f.write(s.data(), (std::streamsize)s.size())) fail(PBRT_ERR_INVALID_ARG, "cannot write " + path);
}

// ---------------------------------------------------------------------------
// PFM
// ---------------------------------------------------------------------------
void write_pfm(const std::string& path, uint32_t w, uint32_t h, uint32_t ch, const float* px, uint32_t stride) {
    std::ofstream f(path, std::ios::binary);
    if (!f) fail(PBRT_ERR_INVALID_ARG, "cannot write " + path);
    const std::string head = std::string(ch == 3 ? "PF" : "Pf") + "\n" + std::to_string(w) + " " + std::to_string(h) +
                             "\n-1.0\n";
    f.write(head.data(), (std::streamsize)head.size());
    std::vector<float> row((size_t)w * ch);
    for (uint32_t y = 0; y < h; ++y) {
        const float* src = px + (size_t)(h - 1 - y) * w * stride;  // bottom row first
        for (uint32_t x = 0; x < w; ++x)
            for (uint32_t c = 0; c < ch; ++c) row[(size_t)x * ch + c] = src[(size_t)x * stride + c];
        f.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)(row.size() * sizeof(float)));
    }
    if (!f) fail(PBRT_ERR_INVALID_ARG, "cannot write " + path);
}

std::vector<float> read_pfm(const std::string& path, uint32_t& w, uint32_t& h, uint32_t& ch) {
    const std::string s = slurp(path);
    std::istringstream in(s);
    std::string magic;
    double scale = 0.0;
    in >> magic >> w >> h >> scale;
    if (!in || (magic != "PF" && magic != "Pf") || w == 0 || h == 0 || scale == 0.0)
        fail(PBRT_ERR_INVALID_ARG, path + ": not a PFM file");
    ch = magic == "PF" ? 3 : 1;
    const size_t off = (size_t)in.tellg() + 1;  // one whitespace byte after the scale
    const size_t n = (size_t)w * h * ch;
    if (off + n * 4 > s.size()) fail(PBRT_ERR_INVALID_ARG, path + ": truncated PFM data");
    std::vector<float> out(n);
    const bool big = scale > 0.0;
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* src = reinterpret_cast<const uint8_t*>(s.data()) + off + (size_t)(h - 1 - y) * w * ch * 4;
        for (size_t i = 0; i < (size_t)w * ch; ++i) {
            uint8_t b[4];
            std::memcpy(b, src + 4 * i, 4);
            if (big) std::swap(b[0], b[3]), std::swap(b[1], b[2]);
            std::memcpy(&out[(size_t)y * w * ch + i], b, 4);
        }
    }
    return out;
}

// vsg::mat4 * vec4 as the reference evaluates it: left-to-right sums over columns
void mul(const float* m, const float* v, float* out) {
    for (int r = 0; r < 4; ++r) out[r] = m[0 + r] * v[0] + m[4 + r] * v[1] + m[8 + r] * v[2] + m[12 + r] * v[3];
}

template <class T>
T* alloc(size_t n) {
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(1, n) * sizeof(T)));
    if (!p) fail(PBRT_ERR_OUT_OF_MEMORY, "host allocation failed");
    return p;
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_matrices_read(const char* path, pbrt_camera_matrices** out, uint32_t* count) {
    return guard([&] {
        PBRT_REQUIRE(path && out && count, "pbrt_matrices_read: NULL argument");
        *out = nullptr;
        *count = 0;
        const auto m = read_matrices(path);
        pbrt_camera_matrices* p = alloc<pbrt_camera_matrices>(m.size());
        if (!m.empty()) std::memcpy(p, m.data(), m.size() * sizeof(pbrt_camera_matrices));
        *out = p;
        *count = (uint32_t)m.size();
    });
}

extern "C" int32_t pbrt_matrices_write(const char* path, const pbrt_camera_matrices* m, uint32_t count) {
    return guard([&] {
        PBRT_REQUIRE(path && (m || count == 0), "pbrt_matrices_write: NULL argument");
        write_matrices(path, m, count);
    });
}

extern "C" int32_t pbrt_pfm_write(const char* path, uint32_t w, uint32_t h, uint32_t ch, const float* px,
                                  uint32_t stride) {
    return guard([&] {
        PBRT_REQUIRE(path && px && w && h, "pbrt_pfm_write: bad argument");
        PBRT_REQUIRE((ch == 1 || ch == 3) && stride >= ch, "pbrt_pfm_write: channels must be 1 or 3, stride >= channels");
        write_pfm(path, w, h, ch, px, stride);
    });
}

extern "C" int32_t pbrt_pfm_read(const char* path, uint32_t* w, uint32_t* h, uint32_t* ch, float** px) {
    return guard([&] {
        PBRT_REQUIRE(path && w && h && ch && px, "pbrt_pfm_read: NULL argument");
        *px = nullptr;
        const std::vector<float> v = read_pfm(path, *w, *h, *ch);
        float* p = alloc<float>(v.size());
        std::memcpy(p, v.data(), v.size() * sizeof(float));
        *px = p;
    });
}

extern "C" int32_t pbrt_normals_to_spherical(const float* xyzw, uint64_t n, float* tp) {
    return guard([&] {
        PBRT_REQUIRE((xyzw && tp) || n == 0, "pbrt_normals_to_spherical: NULL argument");
        for (uint64_t i = 0; i < n; ++i) {
            tp[2 * i] = std::acos(xyzw[4 * i + 2]);
            tp[2 * i + 1] = std::atan2(xyzw[4 * i + 1], xyzw[4 * i]);
        }
    });
}

extern "C" int32_t pbrt_spherical_to_normals(const float* tp, uint64_t n, float* xyzw) {
    return guard([&] {
        PBRT_REQUIRE((xyzw && tp) || n == 0, "pbrt_spherical_to_normals: NULL argument");
        for (uint64_t i = 0; i < n; ++i) {
            const float t = tp[2 * i], p = tp[2 * i + 1];
            xyzw[4 * i] = std::cos(p) * std::sin(t);
            xyzw[4 * i + 1] = std::sin(p) * std::sin(t);
            xyzw[4 * i + 2] = std::cos(t);
            xyzw[4 * i + 3] = 1.0f;
        }
    });
}

extern "C" int32_t pbrt_albedo_from_float(const float* rgba, uint64_t n, uint8_t* out) {
    return guard([&] {
        PBRT_REQUIRE((rgba && out) || n == 0, "pbrt_albedo_from_float: NULL argument");
        for (uint64_t i = 0; i < 4 * n; ++i) {
            const float v = rgba[i] * 255.0f;
            out[i] = (uint8_t)(v >= 255.0f ? 255.0f : (v > 0.0f ? v : 0.0f));  // truncation, as static_cast
        }
    });
}

extern "C" int32_t pbrt_albedo_from_half(const uint16_t* rgba, uint64_t n, uint8_t* out) {
    return guard([&] {
        PBRT_REQUIRE((rgba && out) || n == 0, "pbrt_albedo_from_half: NULL argument");
        for (uint64_t i = 0; i < 4 * n; ++i) {
            const uint32_t h = rgba[i];
            // sign, exponent rebias by 112 (0x1C000 << 13), mantissa: RenderIO.cpp:199-203
            const uint32_t t = ((h & 0x8000u) << 16) | (((h & 0x7C00u) + 0x1C000u) << 13) | ((h & 0x03FFu) << 13);
            float f;
            std::memcpy(&f, &t, 4);
            const float v = f * 255.0f;
            out[i] = (uint8_t)(v >= 255.0f ? 255.0f : (v > 0.0f ? v : 0.0f));
        }
    });
}

extern "C" int32_t pbrt_unorm_to_float(const uint8_t* in, uint64_t n, float* out) {
    return guard([&] {
        PBRT_REQUIRE((in && out) || n == 0, "pbrt_unorm_to_float: NULL argument");
        for (uint64_t i = 0; i < 4 * n; ++i) out[i] = (float)in[i] / 255.0f;
    });
}

extern "C" int32_t pbrt_depth_to_position(const float* depth, uint32_t w, uint32_t h, const pbrt_camera_matrices* m,
                                          float* out) {
    return guard([&] {
        PBRT_REQUIRE(depth && m && out && w && h, "pbrt_depth_to_position: bad argument");
        PBRT_REQUIRE(m->has_proj, "pbrt_depth_to_position: the matrices need a separate projection matrix");
        const float* cam = m->inv_view + 12;  // inv_view[3]
        for (uint32_t i = 0; i < w * h; ++i) {
            const uint32_t x = i % w, y = i / w;
            const float p[4] = {(x + 0.5f) / w * 2 - 1, (y + 0.5f) / h * 2 - 1, 1.0f, 1.0f};
            float dir[4];
            mul(m->inv_proj, p, dir);
            dir[3] = 0.0f;
            const float len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] + dir[3] * dir[3]);
            const float nd[4] = {dir[0] / len, dir[1] / len, dir[2] / len, dir[3] / len};
            float wd[4];
            mul(m->inv_view, nd, wd);
            for (int k = 0; k < 3; ++k) out[4 * (size_t)i + k] = cam[k] + wd[k] * depth[i];
            out[4 * (size_t)i + 3] = 1.0f;
        }
    });
}

extern "C" int32_t pbrt_position_to_depth(const float* xyzw, uint64_t n, const pbrt_camera_matrices* m, float* depth) {
    return guard([&] {
        PBRT_REQUIRE((xyzw && depth && m) || n == 0, "pbrt_position_to_depth: NULL argument");
        float cam[3];
        const float wc = m->inv_view[15];
        PBRT_REQUIRE(wc != 0.0f, "pbrt_position_to_depth: inv_view has no camera position");
        for (int k = 0; k < 3; ++k) cam[k] = m->inv_view[12 + k] / wc;
        for (uint64_t i = 0; i < n; ++i) {
            const float d[3] = {cam[0] - xyzw[4 * i], cam[1] - xyzw[4 * i + 1], cam[2] - xyzw[4 * i + 2]};
            depth[i] = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
    });
}
