// Scene front-end and image files (include/pbrt_io.h): Wavefront OBJ/MTL read
// with the reference's assimp import flags and vsgXchange material conversion,
// plus the PNG / TGA / PNM codecs its textures and the frame dumps need.
#include <zlib.h>

#include <algorithm>
#include <array>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pbrt_io.h"
#include "capi_common.hpp"

namespace pbrt {
namespace {

// ---------------------------------------------------------------------------
// files
// ---------------------------------------------------------------------------
std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(PBRT_ERR_INVALID_ARG, "cannot open " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    std::vector<uint8_t> data((size_t)std::max<std::streamoff>(n, 0));
    if (n > 0 && !f.read(reinterpret_cast<char*>(data.data()), n)) fail(PBRT_ERR_INVALID_ARG, "cannot read " + path);
    return data;
}

std::string dir_of(const std::string& path) {
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct Image {
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

// ---------------------------------------------------------------------------
// PNG (ISO/IEC 15948): chunks, zlib stream, per-row filters, colour types
// ---------------------------------------------------------------------------
uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

Image decode_png(const std::vector<uint8_t>& f, const std::string& name) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) fail(PBRT_ERR_INVALID_ARG, name + ": not a PNG file");
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    bool have_ihdr = false;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (pos + 12 + (size_t)len > f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated PNG chunk");
        const char* type = reinterpret_cast<const char*>(&f[pos + 4]);
        const uint8_t* data = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (!have_ihdr || w == 0 || h == 0) fail(PBRT_ERR_INVALID_ARG, name + ": PNG without a valid IHDR");
    if (interlace != 0) fail(PBRT_ERR_UNSUPPORTED, name + ": interlaced PNG is not supported");
    int channels = 0;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: fail(PBRT_ERR_INVALID_ARG, name + ": bad PNG colour type");
    }
    const bool depth_ok = depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4));
    if (!depth_ok || (ctype == 3 && depth == 16)) fail(PBRT_ERR_INVALID_ARG, name + ": bad PNG bit depth");
    if (ctype == 3 && plte.size() < 3) fail(PBRT_ERR_INVALID_ARG, name + ": palette PNG without PLTE");
    if ((uint64_t)w * h > (1ull << 28)) fail(PBRT_ERR_UNSUPPORTED, name + ": PNG too large");
    const size_t bits = (size_t)channels * depth, stride = ((size_t)w * bits + 7) / 8, bpp = std::max<size_t>(1, bits / 8);
    std::vector<uint8_t> raw((stride + 1) * h);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) fail(PBRT_ERR_INTERNAL, "inflateInit failed");
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t produced = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || produced != raw.size())
        fail(PBRT_ERR_INVALID_ARG, name + ": corrupt PNG image data");
    // undo the row filters in place (filter byte, then the row)
    std::vector<uint8_t> prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* row = &raw[y * (stride + 1)];
        const int ft = row[0];
        uint8_t* cur = row + 1;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = cur[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: fail(PBRT_ERR_INVALID_ARG, name + ": bad PNG filter type");
            }
            cur[i] = (uint8_t)v;
        }
        std::memcpy(prev.data(), cur, stride);
    }
    Image img;
    img.w = w;
    img.h = h;
    img.rgba.resize((size_t)w * h * 4);
    // sample k of a row at the file's depth
    auto sample = [&](const uint8_t* row, size_t k) -> uint32_t {
        if (depth == 8) return row[k];
        if (depth == 16) return (uint32_t)row[2 * k] << 8 | row[2 * k + 1];
        const size_t bit = k * depth;
        return (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
    };
    auto to8 = [&](uint32_t v) -> uint8_t {
        if (depth == 8) return (uint8_t)v;
        if (depth == 16) return (uint8_t)(v >> 8);
        return (uint8_t)(v * 255 / ((1u << depth) - 1));
    };
    const bool key = !trns.empty() && (ctype == 0 || ctype == 2);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = &raw[y * (stride + 1) + 1];
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t* o = &img.rgba[((size_t)y * w + x) * 4];
            const size_t k = (size_t)x * channels;
            if (ctype == 3) {
                const uint32_t i = sample(row, k);
                if ((size_t)3 * i + 2 >= plte.size()) fail(PBRT_ERR_INVALID_ARG, name + ": PNG palette index out of range");
                o[0] = plte[3 * i];
                o[1] = plte[3 * i + 1];
                o[2] = plte[3 * i + 2];
                o[3] = i < trns.size() ? trns[i] : 255;
                continue;
            }
            if (channels <= 2) {
                const uint32_t g = sample(row, k);
                o[0] = o[1] = o[2] = to8(g);
                o[3] = channels == 2 ? to8(sample(row, k + 1)) : 255;
                if (key && trns.size() >= 2 && g == ((uint32_t)trns[0] << 8 | trns[1])) o[3] = 0;
            } else {
                const uint32_t r = sample(row, k), g = sample(row, k + 1), b = sample(row, k + 2);
                o[0] = to8(r);
                o[1] = to8(g);
                o[2] = to8(b);
                o[3] = channels == 4 ? to8(sample(row, k + 3)) : 255;
                if (key && trns.size() >= 6 && r == ((uint32_t)trns[0] << 8 | trns[1]) &&
                    g == ((uint32_t)trns[2] << 8 | trns[3]) && b == ((uint32_t)trns[4] << 8 | trns[5]))
                    o[3] = 0;
            }
        }
    }
    return img;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void put_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, (uint32_t)crc32(0L, &out[start], (uInt)(out.size() - start)));
}

std::vector<uint8_t> encode_png(uint32_t w, uint32_t h, const uint8_t* px, bool bgra) {
    std::vector<uint8_t> raw((size_t)(4 * w + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* row = &raw[(size_t)y * (4 * w + 1)];
        row[0] = 0;  // filter: none
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t* s = &px[((size_t)y * w + x) * 4];
            uint8_t* d = row + 1 + 4 * (size_t)x;
            d[0] = bgra ? s[2] : s[0];
            d[1] = s[1];
            d[2] = bgra ? s[0] : s[2];
            d[3] = s[3];
        }
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) fail(PBRT_ERR_INTERNAL, "deflate failed");
    z.resize(zn);
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, adaptive filters, no interlace
    put_chunk(out, "IHDR", ihdr);
    put_chunk(out, "IDAT", z);
    put_chunk(out, "IEND", {});
    return out;
}

// ---------------------------------------------------------------------------
// TGA (Truevision TGA 2.0): types 2/3 raw, 10/11 run-length, 8/24/32 bits
// ---------------------------------------------------------------------------
Image decode_tga(const std::vector<uint8_t>& f, const std::string& name) {
    if (f.size() < 18) fail(PBRT_ERR_INVALID_ARG, name + ": truncated TGA header");
    const int idlen = f[0], cmap_type = f[1], type = f[2];
    const int cmap_len = f[5] | f[6] << 8, cmap_bits = f[7];
    const uint32_t w = f[12] | f[13] << 8, h = f[14] | f[15] << 8;
    const int bits = f[16], desc = f[17];
    const bool gray = type == 3 || type == 11, rle = type == 10 || type == 11;
    if (!(type == 2 || type == 3 || type == 10 || type == 11))
        fail(PBRT_ERR_UNSUPPORTED, name + ": TGA image type " + std::to_string(type) + " is not supported");
    if (gray ? bits != 8 : (bits != 24 && bits != 32)) fail(PBRT_ERR_UNSUPPORTED, name + ": TGA pixel depth not supported");
    if (w == 0 || h == 0) fail(PBRT_ERR_INVALID_ARG, name + ": empty TGA image");
    const size_t bpp = (size_t)bits / 8;
    size_t pos = 18 + (size_t)idlen + (cmap_type ? (size_t)cmap_len * ((cmap_bits + 7) / 8) : 0);
    std::vector<uint8_t> pix((size_t)w * h * bpp);
    if (!rle) {
        if (pos + pix.size() > f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated TGA data");
        std::memcpy(pix.data(), &f[pos], pix.size());
    } else {
        size_t o = 0;
        while (o < pix.size()) {
            if (pos >= f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated TGA run");
            const int head = f[pos++];
            const size_t count = (size_t)(head & 0x7F) + 1;
            if (o + count * bpp > pix.size()) fail(PBRT_ERR_INVALID_ARG, name + ": TGA run past the image");
            if (head & 0x80) {
                if (pos + bpp > f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated TGA run");
                for (size_t k = 0; k < count; ++k, o += bpp) std::memcpy(&pix[o], &f[pos], bpp);
                pos += bpp;
            } else {
                if (pos + count * bpp > f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated TGA run");
                std::memcpy(&pix[o], &f[pos], count * bpp);
                o += count * bpp;
                pos += count * bpp;
            }
        }
    }
    const bool top = desc & 0x20, right = desc & 0x10;  // origin: bottom-left unless these bits say otherwise
    Image img;
    img.w = w;
    img.h = h;
    img.rgba.resize((size_t)w * h * 4);
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t sy = top ? y : h - 1 - y;
        for (uint32_t x = 0; x < w; ++x) {
            const uint32_t sx = right ? w - 1 - x : x;
            const uint8_t* s = &pix[((size_t)sy * w + sx) * bpp];
            uint8_t* d = &img.rgba[((size_t)y * w + x) * 4];
            if (gray) {
                d[0] = d[1] = d[2] = s[0];
                d[3] = 255;
            } else {  // stored B, G, R(, A)
                d[0] = s[2];
                d[1] = s[1];
                d[2] = s[0];
                d[3] = bpp == 4 ? s[3] : 255;
            }
        }
    }
    return img;
}

// ---------------------------------------------------------------------------
// Netpbm: P5 (grey), P6 (RGB), P7 (PAM: GRAYSCALE(_ALPHA), RGB(_ALPHA))
// ---------------------------------------------------------------------------
Image decode_pnm(const std::vector<uint8_t>& f, const std::string& name) {
    size_t pos = 2;
    const char kind = (char)f[1];
    auto skip_space = [&] {
        while (pos < f.size()) {
            if (f[pos] == '#') {
                while (pos < f.size() && f[pos] != '\n') ++pos;
            } else if (std::isspace(f[pos])) {
                ++pos;
            } else {
                break;
            }
        }
    };
    auto number = [&]() -> uint32_t {
        skip_space();
        uint32_t v = 0;
        bool any = false;
        while (pos < f.size() && std::isdigit(f[pos])) {
            const uint32_t digit = (uint32_t)(f[pos++] - '0');
            if (v > (UINT32_MAX - digit) / 10) fail(PBRT_ERR_INVALID_ARG, name + ": PNM header number out of range");
            v = v * 10 + digit;
            any = true;
        }
        if (!any) fail(PBRT_ERR_INVALID_ARG, name + ": bad PNM header");
        return v;
    };
    uint32_t w = 0, h = 0, depth = 0, maxval = 0;
    if (kind == '5' || kind == '6') {
        w = number();
        h = number();
        maxval = number();
        depth = kind == '5' ? 1 : 3;
        ++pos;  // one whitespace byte before the raster
    } else if (kind == '7') {
        while (true) {
            skip_space();
            size_t e = pos;
            while (e < f.size() && !std::isspace(f[e])) ++e;
            const std::string tok(reinterpret_cast<const char*>(&f[pos]), e - pos);
            pos = e;
            if (tok == "ENDHDR") {
                while (pos < f.size() && f[pos] != '\n') ++pos;
                ++pos;
                break;
            }
            if (tok == "WIDTH") w = number();
            else if (tok == "HEIGHT") h = number();
            else if (tok == "DEPTH") depth = number();
            else if (tok == "MAXVAL") maxval = number();
            else if (tok.empty()) fail(PBRT_ERR_INVALID_ARG, name + ": PAM header without ENDHDR");
            else while (pos < f.size() && f[pos] != '\n') ++pos;  // TUPLTYPE and the like
        }
    } else {
        fail(PBRT_ERR_UNSUPPORTED, name + ": only binary PGM/PPM (P5/P6) and PAM (P7) are read");
    }
    if (w == 0 || h == 0 || depth < 1 || depth > 4 || maxval < 1 || maxval > 65535)
        fail(PBRT_ERR_INVALID_ARG, name + ": bad PNM header");
    // the same cap as PNG: w * h * 4 and the raster size then fit size_t with room to spare
    if ((uint64_t)w * h > (1ull << 28)) fail(PBRT_ERR_INVALID_ARG, name + ": PNM too large");
    if (pos > f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated PNM header");
    const size_t bps = maxval > 255 ? 2 : 1;
    if (pos + (size_t)w * h * depth * bps > f.size()) fail(PBRT_ERR_INVALID_ARG, name + ": truncated PNM data");
    Image img;
    img.w = w;
    img.h = h;
    img.rgba.resize((size_t)w * h * 4);
    for (size_t p = 0; p < (size_t)w * h; ++p) {
        uint32_t c[4] = {0, 0, 0, maxval};
        for (uint32_t k = 0; k < depth; ++k) {
            const size_t o = pos + (p * depth + k) * bps;
            c[k] = bps == 2 ? ((uint32_t)f[o] << 8 | f[o + 1]) : f[o];
        }
        uint8_t* d = &img.rgba[p * 4];
        auto s8 = [&](uint32_t v) { return (uint8_t)((std::min(v, maxval) * 255u + maxval / 2) / maxval); };
        if (depth <= 2) {
            d[0] = d[1] = d[2] = s8(c[0]);
            d[3] = depth == 2 ? s8(c[1]) : 255;
        } else {
            d[0] = s8(c[0]);
            d[1] = s8(c[1]);
            d[2] = s8(c[2]);
            d[3] = depth == 4 ? s8(c[3]) : 255;
        }
    }
    return img;
}

Image load_image(const std::string& path, uint64_t* bytes_read = nullptr) {
    const std::vector<uint8_t> f = read_file(path);
    if (bytes_read) *bytes_read += f.size();
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') return decode_png(f, path);
    if (f.size() >= 3 && f[0] == 'P' && (f[1] == '5' || f[1] == '6' || f[1] == '7')) return decode_pnm(f, path);
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8)
        fail(PBRT_ERR_UNSUPPORTED, path + ": JPEG is not decoded by this build (convert the texture to PNG or TGA)");
    const std::string ext = lower(path.substr(path.find_last_of('.') + 1));
    if (ext == "tga") return decode_tga(f, path);
    fail(PBRT_ERR_UNSUPPORTED, path + ": unknown image format (PNG, TGA, PPM/PGM and PAM are read)");
}

// ---------------------------------------------------------------------------
// materials: vsgXchange (assimp.cpp:672-745) then RayTracingVisitor.cpp:253-298
// ---------------------------------------------------------------------------
bool transmissive_not_one(const float* t) { return t[0] != 1.0f || t[1] != 1.0f || t[2] != 1.0f; }

void pack_phong(const pbrt_phong_material& m, pbrt_material& out, uint32_t& emissive) {
    std::memset(&out, 0, sizeof(out));
    std::memcpy(out.ambient_roughness, m.ambient, 16);
    std::memcpy(out.specular_dissolve, m.specular, 16);
    std::memcpy(out.diffuse_ior, m.diffuse, 16);
    std::memcpy(out.emission_texture_id, m.emissive, 16);
    std::memcpy(out.transmittance_illum, m.transmissive, 12);
    emissive = (m.emissive[0] + m.emissive[1] + m.emissive[2]) != 0.0f;
    // shininess to roughness: sqrt(2 / (shininess + 2))
    out.ambient_roughness[3] = std::sqrt(2.0f / (m.shininess + 2.0f));
    out.diffuse_ior[3] = m.index_of_refraction;
    out.specular_dissolve[3] = m.alpha_mask;
    out.emission_texture_id[3] = m.alpha_mask_cutoff;
    out.category_id = m.category_id;
    if (transmissive_not_one(m.transmissive)) out.transmittance_illum[3] = 7.0f;  // refraction + reflection
}

void pack_pbr(const pbrt_pbr_material& m, pbrt_material& out, uint32_t& emissive) {
    std::memset(&out, 0, sizeof(out));
    std::memcpy(out.ambient_roughness, m.base_color, 16);
    std::memcpy(out.specular_dissolve, m.specular, 16);
    std::memcpy(out.diffuse_ior, m.diffuse, 16);
    std::memcpy(out.emission_texture_id, m.emissive, 16);
    std::memcpy(out.transmittance_illum, m.transmission, 12);
    emissive = (m.emissive[0] + m.emissive[1] + m.emissive[2]) != 0.0f;
    out.ambient_roughness[3] = m.roughness;
    out.diffuse_ior[3] = m.index_of_refraction;
    out.specular_dissolve[3] = m.alpha_mask;
    out.emission_texture_id[3] = m.alpha_mask_cutoff;
    out.category_id = m.category_id;
    if (transmissive_not_one(m.transmission)) out.transmittance_illum[3] = 7.0f;
}

// one MTL material as assimp's OBJ reader holds it (defaults: Kd 0.6, Tf 1, Ni 1)
struct MtlMaterial {
    float ka[3] = {0, 0, 0}, kd[3] = {0.6f, 0.6f, 0.6f}, ks[3] = {0, 0, 0}, ke[3] = {0, 0, 0}, tf[3] = {1, 1, 1};
    float ns = 0.0f, ni = 1.0f, d = 1.0f;
    int illum = 1;
    std::string map_kd, map_ks, map_ke, map_kn;
};

pbrt_phong_material phong_of(const MtlMaterial& m) {
    pbrt_phong_material p{};
    for (int k = 0; k < 3; ++k) {
        p.ambient[k] = m.ka[k];
        p.diffuse[k] = m.kd[k];
        p.specular[k] = m.ks[k];
        p.emissive[k] = m.ke[k];
        p.transmissive[k] = m.tf[k];
    }
    p.ambient[3] = p.diffuse[3] = p.specular[3] = p.emissive[3] = 1.0f;
    p.shininess = m.ns;  // no shininess strength in OBJ: factor 1
    if (p.shininess < 0.01f) {
        p.shininess = 0.0f;
        std::memset(p.specular, 0, sizeof(p.specular));
    }
    p.alpha_mask = 1.0f;
    p.alpha_mask_cutoff = 0.5f;
    p.index_of_refraction = m.ni;
    p.category_id = 0;
    return p;
}

// ---------------------------------------------------------------------------
// tokenizer
// ---------------------------------------------------------------------------
struct Line {
    const char* p;
    const char* e;
    void skip() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    }
    bool done() {
        skip();
        return p >= e;
    }
    std::string word() {
        skip();
        const char* s = p;
        while (p < e && *p != ' ' && *p != '\t' && *p != '\r') ++p;
        return std::string(s, p);
    }
    std::string rest() {
        skip();
        const char* end = e;
        while (end > p && (end[-1] == ' ' || end[-1] == '\t' || end[-1] == '\r')) --end;
        return std::string(p, end);
    }
    float number(const std::string& where) {
        skip();
        float v = 0.0f;
        const char* s = p;
        if (s < e && *s == '+') ++s;
        auto r = std::from_chars(s, e, v);
        if (r.ec != std::errc()) fail(PBRT_ERR_INVALID_ARG, where + ": expected a number");
        p = r.ptr;
        return v;
    }
    bool try_number(float& v) {
        skip();
        if (p >= e) return false;
        const char* s = p;
        if (*s == '+') ++s;
        auto r = std::from_chars(s, e, v);
        if (r.ec != std::errc()) return false;
        p = r.ptr;
        return true;
    }
};

template <class F>
void for_each_line(const std::vector<uint8_t>& text, F&& fn) {
    const char* p = reinterpret_cast<const char*>(text.data());
    const char* end = p + text.size();
    std::string joined;
    uint32_t lineno = 0;
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        const char* e = nl ? nl : end;
        ++lineno;
        // a trailing backslash continues the line
        const char* t = e;
        while (t > p && (t[-1] == '\r' || t[-1] == ' ' || t[-1] == '\t')) --t;
        if (t > p && t[-1] == '\\') {
            joined.append(p, t - 1);
            joined.push_back(' ');
        } else if (!joined.empty()) {
            joined.append(p, e);
            Line l{joined.data(), joined.data() + joined.size()};
            fn(l, lineno);
            joined.clear();
        } else {
            Line l{p, e};
            fn(l, lineno);
        }
        p = nl ? nl + 1 : end;
    }
    if (!joined.empty()) {
        Line l{joined.data(), joined.data() + joined.size()};
        fn(l, lineno);
    }
}

// a texture statement: options (-o u v w, -clamp on, ...) then the file name
std::string texture_name(Line& l) {
    std::string last;
    while (!l.done()) {
        std::string w = l.word();
        if (!w.empty() && w[0] == '-') {
            // options: -imfchan / -type take a word, the others numbers or on / off
            if (w == "-imfchan" || w == "-type") {
                l.word();
                continue;
            }
            float v;
            while (l.try_number(v)) {
            }
            Line save = l;
            const std::string a = l.word();
            if (a != "on" && a != "off") l = save;
            continue;
        }
        last = w;
        const std::string r = l.rest();
        if (!r.empty()) last += " " + r;  // names with spaces run to the end of the line
        break;
    }
    return last;
}

void parse_mtl(const std::vector<uint8_t>& text, const std::string& name, std::map<std::string, MtlMaterial>& out) {
    MtlMaterial* cur = nullptr;
    for_each_line(text, [&](Line& l, uint32_t lineno) {
        const std::string key = l.word();
        if (key.empty() || key[0] == '#') return;
        const std::string where = name + ":" + std::to_string(lineno);
        if (key == "newmtl") {
            cur = &out[l.rest()];
            *cur = MtlMaterial();
            return;
        }
        if (!cur) return;
        auto color = [&](float* c) {
            const std::string first = l.word();
            if (first == "spectral" || first == "xyz") return;  // not read by assimp either
            Line f{first.data(), first.data() + first.size()};
            c[0] = f.number(where);
            float v;
            if (l.try_number(v)) {
                c[1] = v;
                c[2] = l.try_number(v) ? v : c[1];
            } else {
                c[1] = c[2] = c[0];
            }
        };
        if (key == "Ka") color(cur->ka);
        else if (key == "Kd") color(cur->kd);
        else if (key == "Ks") color(cur->ks);
        else if (key == "Ke") color(cur->ke);
        else if (key == "Tf") color(cur->tf);
        else if (key == "Ns") cur->ns = l.number(where);
        else if (key == "Ni") cur->ni = l.number(where);
        else if (key == "d") cur->d = l.number(where);
        else if (key == "Tr") cur->d = 1.0f - l.number(where);
        else if (key == "illum") cur->illum = (int)l.number(where);
        else if (key == "map_Kd") cur->map_kd = texture_name(l);
        else if (key == "map_Ks") cur->map_ks = texture_name(l);
        else if (key == "map_Ke" || key == "map_emissive") cur->map_ke = texture_name(l);
        else if (key == "map_Kn" || key == "norm") cur->map_kn = texture_name(l);
    });
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------
struct MeshBuild {
    std::vector<float> pos, nor, uv;
};

struct Corner {
    int v, t, n;  // 0-based, -1 when absent
};

}  // namespace
}  // namespace pbrt

struct pbrt_loaded_scene {
    struct Mesh {
        std::vector<float> pos, nor, uv;
        std::vector<uint16_t> idx16;
        std::vector<uint32_t> idx32;
        bool emissive = false;
    };
    std::vector<Mesh> meshes;
    std::vector<pbrt_mesh_desc> mesh_descs;
    std::vector<pbrt_instance_desc> instances;
    std::vector<std::vector<uint8_t>> texture_pixels;
    std::vector<pbrt_texture_desc> textures;
    pbrt_loaded_scene_stats stats{};
    uint32_t bvh_builder = PBRT_BVH_SAH_HOST;
};

namespace pbrt {
namespace {

void load_obj(const std::string& path, const pbrt_load_options* opt, pbrt_loaded_scene& sc) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::string base = dir_of(path);
    const std::vector<uint8_t> text = read_file(path);
    uint64_t bytes = text.size();
    std::vector<float> v, vt, vn;
    std::map<std::string, MtlMaterial> mtl;
    // materials in order of first use; index 0 is reserved for assimp's default material
    std::vector<std::string> mat_names = {std::string()};
    std::unordered_map<std::string, int> mat_index;
    std::vector<MeshBuild> builds(1);
    std::vector<bool> used(1, false);
    std::vector<int> order;  // materials in order of their first face
    int cur_mat = 0;
    std::vector<Corner> face;
    auto resolve = [&](int i, size_t count, const std::string& where) -> int {
        const long r = i > 0 ? (long)i - 1 : (long)count + i;
        if (i == 0 || r < 0 || r >= (long)count) fail(PBRT_ERR_INVALID_ARG, where + ": index out of range");
        return (int)r;
    };
    auto emit = [&](MeshBuild& b, const Corner& c) {
        for (int k = 0; k < 3; ++k) b.pos.push_back(v[3 * (size_t)c.v + k]);
        for (int k = 0; k < 3; ++k) b.nor.push_back(c.n >= 0 ? vn[3 * (size_t)c.n + k] : 0.0f);
        b.uv.push_back(c.t >= 0 ? vt[2 * (size_t)c.t] : 0.0f);
        b.uv.push_back(c.t >= 0 ? 1.0f - vt[2 * (size_t)c.t + 1] : 0.0f);  // aiProcess_FlipUVs
    };
    for_each_line(text, [&](Line& l, uint32_t lineno) {
        const std::string key = l.word();
        if (key.empty() || key[0] == '#') return;
        const std::string where = path + ":" + std::to_string(lineno);
        if (key == "v") {
            for (int k = 0; k < 3; ++k) v.push_back(l.number(where));
        } else if (key == "vt") {
            vt.push_back(l.number(where));
            float t = 0.0f;
            l.try_number(t);
            vt.push_back(t);
        } else if (key == "vn") {
            for (int k = 0; k < 3; ++k) vn.push_back(l.number(where));
        } else if (key == "f") {
            face.clear();
            while (!l.done()) {
                const std::string tok = l.word();
                Corner c{-1, -1, -1};
                int part = 0;
                const char* s = tok.data();
                const char* e = s + tok.size();
                while (s <= e && part < 3) {
                    const char* slash = std::find(s, e, '/');
                    if (slash > s) {
                        int idx = 0;
                        auto r = std::from_chars(s, slash, idx);
                        if (r.ec != std::errc() || r.ptr != slash) fail(PBRT_ERR_INVALID_ARG, where + ": bad face index");
                        const size_t count = part == 0 ? v.size() / 3 : part == 1 ? vt.size() / 2 : vn.size() / 3;
                        const int ri = resolve(idx, count, where);
                        (part == 0 ? c.v : part == 1 ? c.t : c.n) = ri;
                    }
                    ++part;
                    s = slash + 1;
                }
                if (c.v < 0) fail(PBRT_ERR_INVALID_ARG, where + ": face corner without a position");
                face.push_back(c);
            }
            if (face.size() < 3) return;  // points and lines are not traced
            MeshBuild& b = builds[cur_mat];
            if (!used[cur_mat]) {
                used[cur_mat] = true;
                order.push_back(cur_mat);
            }
            if (face.size() == 4) {
                // split at the reflex corner: the angles between the diagonal and the
                // two edges at a corner add up to more than pi only there
                int start = 0;
                for (int i = 0; i < 4; ++i) {
                    auto p = [&](int k) {
                        const float* q = &v[3 * (size_t)face[(size_t)k % 4].v];
                        return std::array<float, 3>{q[0], q[1], q[2]};
                    };
                    const auto o = p(i), a = p(i + 3), dg = p(i + 2), b2 = p(i + 1);
                    auto unit = [&](const std::array<float, 3>& q) {
                        std::array<float, 3> d{q[0] - o[0], q[1] - o[1], q[2] - o[2]};
                        const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                        for (float& x : d) x /= len;
                        return d;
                    };
                    const auto left = unit(a), diag = unit(dg), right = unit(b2);
                    const float angle = std::acos(left[0] * diag[0] + left[1] * diag[1] + left[2] * diag[2]) +
                                        std::acos(right[0] * diag[0] + right[1] * diag[1] + right[2] * diag[2]);
                    if (angle > 3.14159265f) {
                        start = i;
                        break;
                    }
                }
                const int tri[6] = {start, start + 1, start + 2, start, start + 2, start + 3};
                for (int k = 0; k < 6; ++k) emit(b, face[(size_t)tri[k] % 4]);
            } else {
                for (size_t k = 1; k + 1 < face.size(); ++k) {
                    emit(b, face[0]);
                    emit(b, face[k]);
                    emit(b, face[k + 1]);
                }
            }
        } else if (key == "usemtl") {
            const std::string name = l.rest();
            if (!mtl.count(name)) {
                cur_mat = 0;  // unknown material: the default
                return;
            }
            auto it = mat_index.find(name);
            if (it == mat_index.end()) {
                it = mat_index.emplace(name, (int)mat_names.size()).first;
                mat_names.push_back(name);
                builds.emplace_back();
                used.push_back(false);
            }
            cur_mat = it->second;
        } else if (key == "mtllib") {
            // one or more library files
            while (!l.done()) {
                const std::string lib = l.word();
                const std::string full = base + lib;
                const std::vector<uint8_t> mt = read_file(full);
                bytes += mt.size();
                parse_mtl(mt, full, mtl);
            }
        }
        // o, g, s, l, p and the rest: no effect on the traced geometry
    });
    std::unordered_map<std::string, int32_t> tex_index;
    auto texture = [&](const std::string& name) -> int32_t {
        if (name.empty()) return -1;
        const std::string full = name[0] == '/' ? name : base + name;
        auto it = tex_index.find(full);
        if (it != tex_index.end()) return it->second;
        Image img = load_image(full, &bytes);
        const int32_t id = (int32_t)sc.texture_pixels.size();
        sc.texture_pixels.push_back(std::move(img.rgba));
        pbrt_texture_desc td{};
        td.width = img.w;
        td.height = img.h;
        sc.textures.push_back(td);
        tex_index.emplace(full, id);
        return id;
    };
    float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int m : order) {
        MeshBuild& b = builds[(size_t)m];
        const MtlMaterial mm = m == 0 ? MtlMaterial() : mtl.at(mat_names[(size_t)m]);
        pbrt_loaded_scene::Mesh mesh;
        const size_t nv = b.pos.size() / 3;
        mesh.pos = std::move(b.pos);
        mesh.nor = std::move(b.nor);
        mesh.uv = std::move(b.uv);
        for (size_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) {
                bmin[k] = std::min(bmin[k], mesh.pos[3 * i + k]);
                bmax[k] = std::max(bmax[k], mesh.pos[3 * i + k]);
            }
        if (nv < 65535) {
            mesh.idx16.resize(nv);
            for (size_t i = 0; i < nv; ++i) mesh.idx16[i] = (uint16_t)i;
        } else {
            mesh.idx32.resize(nv);
            for (size_t i = 0; i < nv; ++i) mesh.idx32[i] = (uint32_t)i;
        }
        pbrt_mesh_desc md{};
        uint32_t emissive = 0;
        const pbrt_phong_material ph = phong_of(mm);
        pack_phong(ph, md.material, emissive);
        mesh.emissive = emissive != 0;
        md.diffuse_texture = texture(mm.map_kd);
        md.normal_texture = texture(mm.map_kn);
        md.specular_texture = texture(mm.map_ks);
        md.emissive_texture = texture(mm.map_ke);
        sc.meshes.push_back(std::move(mesh));
        sc.mesh_descs.push_back(md);
        sc.stats.triangle_count += (uint32_t)(nv / 3);
        sc.stats.vertex_count += (uint32_t)nv;
        sc.stats.emissive_mesh_count += emissive ? 1u : 0u;
    }
    if (sc.meshes.empty()) fail(PBRT_ERR_INVALID_ARG, path + ": no triangles");
    // descriptors point into the owned arrays (stable from here on)
    for (size_t i = 0; i < sc.meshes.size(); ++i) {
        pbrt_mesh_desc& md = sc.mesh_descs[i];
        const auto& m = sc.meshes[i];
        md.vertex_count = (uint32_t)(m.pos.size() / 3);
        md.positions = m.pos.data();
        md.normals = m.nor.data();
        md.uvs = m.uv.data();
        if (!m.idx16.empty()) {
            md.index_count = (uint32_t)m.idx16.size();
            md.index_stride = 2;
            md.indices = m.idx16.data();
        } else {
            md.index_count = (uint32_t)m.idx32.size();
            md.index_stride = 4;
            md.indices = m.idx32.data();
        }
    }
    for (size_t t = 0; t < sc.textures.size(); ++t) sc.textures[t].rgba8 = sc.texture_pixels[t].data();
    // instances: the file's meshes at each transform
    const uint32_t ni = opt && opt->instance_count ? opt->instance_count : 1;
    PBRT_REQUIRE(!(opt && opt->instance_count) || opt->instance_transforms,
                 "pbrt_load_obj: instance_count without instance_transforms");
    for (uint32_t i = 0; i < ni; ++i)
        for (uint32_t m = 0; m < (uint32_t)sc.meshes.size(); ++m) {
            pbrt_instance_desc in{};
            if (opt && opt->instance_count) {
                std::memcpy(in.transform, opt->instance_transforms + 16 * (size_t)i, sizeof(in.transform));
            } else {
                in.transform[0] = in.transform[5] = in.transform[10] = in.transform[15] = 1.0f;
            }
            in.mesh = m;
            sc.instances.push_back(in);
        }
    sc.bvh_builder = opt ? opt->bvh_builder : (uint32_t)PBRT_BVH_SAH_HOST;
    sc.stats.mesh_count = (uint32_t)sc.meshes.size();
    sc.stats.texture_count = (uint32_t)sc.textures.size();
    sc.stats.instance_count = (uint32_t)sc.instances.size();
    for (int k = 0; k < 3; ++k) {
        sc.stats.bounds_min[k] = bmin[k];
        sc.stats.bounds_max[k] = bmax[k];
    }
    sc.stats.file_bytes = bytes;
    sc.stats.parse_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

extern "C" int32_t pbrt_pack_phong_material(const pbrt_phong_material* m, pbrt_material* out, uint32_t* emissive) {
    return guard([&] {
        PBRT_REQUIRE(m && out, "pbrt_pack_phong_material: NULL argument");
        uint32_t e = 0;
        pack_phong(*m, *out, e);
        if (emissive) *emissive = e;
    });
}

extern "C" int32_t pbrt_pack_pbr_material(const pbrt_pbr_material* m, pbrt_material* out, uint32_t* emissive) {
    return guard([&] {
        PBRT_REQUIRE(m && out, "pbrt_pack_pbr_material: NULL argument");
        uint32_t e = 0;
        pack_pbr(*m, *out, e);
        if (emissive) *emissive = e;
    });
}

extern "C" int32_t pbrt_load_obj(const char* path, const pbrt_load_options* options, pbrt_loaded_scene** out) {
    return guard([&] {
        PBRT_REQUIRE(path && out, "pbrt_load_obj: NULL argument");
        *out = nullptr;
        PBRT_REQUIRE(!options || options->bvh_builder <= PBRT_BVH_PLOC_GPU, "pbrt_load_obj: unknown bvh_builder");
        auto sc = std::make_unique<pbrt_loaded_scene>();
        load_obj(path, options, *sc);
        *out = sc.release();
    });
}

extern "C" int32_t pbrt_loaded_scene_desc(const pbrt_loaded_scene* sc, pbrt_scene_desc* d) {
    return guard([&] {
        PBRT_REQUIRE(sc && d, "pbrt_loaded_scene_desc: NULL argument");
        std::memset(d, 0, sizeof(*d));
        d->mesh_count = (uint32_t)sc->mesh_descs.size();
        d->meshes = sc->mesh_descs.data();
        d->instance_count = (uint32_t)sc->instances.size();
        d->instances = sc->instances.data();
        d->texture_count = (uint32_t)sc->textures.size();
        d->textures = sc->textures.empty() ? nullptr : sc->textures.data();
        d->light_count = 0;
        d->lights = nullptr;
        d->bvh_builder = sc->bvh_builder;
    });
}

extern "C" int32_t pbrt_loaded_scene_stats_get(const pbrt_loaded_scene* sc, pbrt_loaded_scene_stats* s) {
    return guard([&] {
        PBRT_REQUIRE(sc && s, "pbrt_loaded_scene_stats_get: NULL argument");
        *s = sc->stats;
    });
}

extern "C" int32_t pbrt_loaded_scene_destroy(pbrt_loaded_scene* sc) {
    return guard([&] { delete sc; });
}

extern "C" int32_t pbrt_image_load_rgba8(const char* path, uint32_t* w, uint32_t* h, uint8_t** pixels) {
    return guard([&] {
        PBRT_REQUIRE(path && w && h && pixels, "pbrt_image_load_rgba8: NULL argument");
        *pixels = nullptr;
        Image img = load_image(path);
        uint8_t* p = static_cast<uint8_t*>(std::malloc(img.rgba.size()));
        if (!p) fail(PBRT_ERR_OUT_OF_MEMORY, "pbrt_image_load_rgba8: host allocation failed");
        std::memcpy(p, img.rgba.data(), img.rgba.size());
        *w = img.w;
        *h = img.h;
        *pixels = p;
    });
}

extern "C" void pbrt_io_free(void* memory) { std::free(memory); }

extern "C" int32_t pbrt_image_write_png(const char* path, uint32_t w, uint32_t h, const uint8_t* px, uint32_t bgra) {
    return guard([&] {
        PBRT_REQUIRE(path && px && w > 0 && h > 0, "pbrt_image_write_png: bad argument");
        const std::vector<uint8_t> png = encode_png(w, h, px, bgra != 0);
        std::ofstream f(path, std::ios::binary);
        if (!f || !f.write(reinterpret_cast<const char*>(png.data()), (std::streamsize)png.size()))
            fail(PBRT_ERR_INVALID_ARG, std::string("cannot write ") + path);
    });
}
