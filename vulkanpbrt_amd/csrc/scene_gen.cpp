// Seed-driven synthetic scenes (see include/pbrt_scenegen.h).
//
// Material convention: the reference shades spec-gloss materials and converts
// them to metallic at hit time (ptClosesthit.rchit:59-75, color.glsl:30-47).
// A generated dielectric therefore has diffuse = albedo and specular = 0.04,
// a generated metal has diffuse = 0 and specular = albedo; both yield the
// intended metallic value through convertMetallic.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pbrt_scenegen.h"
#include "capi_common.hpp"

namespace pbrt {
namespace {

// ---------------------------------------------------------------------------
// PRNG: splitmix64 seeding a PCG32 (O'Neill, "PCG: A Family of Simple Fast
// Space-Efficient Statistically Good Algorithms", XSH-RR output).
// ---------------------------------------------------------------------------
uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    explicit Pcg32(uint64_t seed) {
        uint64_t s = seed;
        uint64_t init_state = splitmix64(s);
        uint64_t init_seq = splitmix64(s);
        state = 0;
        inc = (init_seq << 1u) | 1u;
        next();
        state += init_state;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    float log_uniform(float a, float b) { return std::exp(uniform(std::log(a), std::log(b))); }
};

struct Mesh {
    std::vector<float> pos, nor, uv;
    std::vector<uint32_t> idx32;
    std::vector<uint16_t> idx16;
    bool use16 = false;
    bool zero_normals = false;
    pbrt_material mat{};
    int32_t tex[4] = {-1, -1, -1, -1};  // diffuse, normal, specular, emissive

    uint32_t add_vertex(float x, float y, float z, float nx, float ny, float nz, float u, float v) {
        uint32_t i = (uint32_t)(pos.size() / 3);
        pos.insert(pos.end(), {x, y, z});
        nor.insert(nor.end(), {nx, ny, nz});
        uv.insert(uv.end(), {u, v});
        return i;
    }
    void tri(uint32_t a, uint32_t b, uint32_t c) {
        if (use16) {
            idx16.insert(idx16.end(), {(uint16_t)a, (uint16_t)b, (uint16_t)c});
        } else {
            idx32.insert(idx32.end(), {a, b, c});
        }
    }
    uint32_t tri_count() const { return (uint32_t)((use16 ? idx16.size() : idx32.size()) / 3); }
};

pbrt_material make_material(float r, float g, float b, float roughness, bool metal) {
    pbrt_material m{};
    if (metal) {
        m.diffuse_ior[0] = m.diffuse_ior[1] = m.diffuse_ior[2] = 0.0f;
        m.specular_dissolve[0] = r;
        m.specular_dissolve[1] = g;
        m.specular_dissolve[2] = b;
    } else {
        m.diffuse_ior[0] = r;
        m.diffuse_ior[1] = g;
        m.diffuse_ior[2] = b;
        m.specular_dissolve[0] = m.specular_dissolve[1] = m.specular_dissolve[2] = 0.04f;
    }
    m.ambient_roughness[3] = roughness;
    m.diffuse_ior[3] = 1.5f;        // index of refraction
    m.specular_dissolve[3] = 1.0f;  // alpha mask flag / dissolve
    m.transmittance_illum[0] = m.transmittance_illum[1] = m.transmittance_illum[2] = 1.0f;
    m.transmittance_illum[3] = 0.0f;  // illum model (7 = refractive)
    m.emission_texture_id[3] = 0.5f;  // alpha cutoff
    m.category_id = 0;
    return m;
}

// Quad p0..p3 (counter-clockwise seen from the front), subdivided nu x nv.
void add_quad(Mesh& m, const float* p0, const float* p1, const float* p3, int nu, int nv, float uscale,
              float vscale, bool with_normals) {
    float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    float e2[3] = {p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]};
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (float& c : n) c = with_normals ? c / len : 0.0f;
    uint32_t base = (uint32_t)(m.pos.size() / 3);
    for (int j = 0; j <= nv; ++j)
        for (int i = 0; i <= nu; ++i) {
            float a = (float)i / nu, b = (float)j / nv;
            m.add_vertex(p0[0] + a * e1[0] + b * e2[0], p0[1] + a * e1[1] + b * e2[1], p0[2] + a * e1[2] + b * e2[2],
                         n[0], n[1], n[2], a * uscale, b * vscale);
        }
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nu; ++i) {
            uint32_t v00 = base + j * (nu + 1) + i, v10 = v00 + 1, v01 = v00 + (nu + 1), v11 = v01 + 1;
            m.tri(v00, v10, v11);
            m.tri(v00, v11, v01);
        }
}

// Axis-aligned unit box [-0.5, 0.5]^3 with outward faces (24 vertices).
void add_unit_box(Mesh& m, float uvscale) {
    const float h = 0.5f;
    const float faces[6][4][3] = {
        {{-h, -h, -h}, {h, -h, -h}, {h, -h, h}, {-h, -h, h}},  // -y
        {{h, h, -h}, {-h, h, -h}, {-h, h, h}, {h, h, h}},      // +y
        {{h, -h, -h}, {h, h, -h}, {h, h, h}, {h, -h, h}},      // +x
        {{-h, h, -h}, {-h, -h, -h}, {-h, -h, h}, {-h, h, h}},  // -x
        {{-h, -h, h}, {h, -h, h}, {h, h, h}, {-h, h, h}},      // +z
        {{-h, h, -h}, {h, h, -h}, {h, -h, -h}, {-h, -h, -h}},  // -z
    };
    for (auto& f : faces) add_quad(m, f[0], f[1], f[3], 1, 1, uvscale, uvscale, true);
}

// Icosphere of the given subdivision level on the unit sphere.  Vertices are
// duplicated per face so each face gets a planar uv projection along its
// dominant axis (well-conditioned tangent frames everywhere).
void add_icosphere(Mesh& m, int subdiv) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    std::vector<float> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                            0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<uint32_t> f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                               4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9,  4, 3,  4,  2, 3, 2, 6, 3,
                               6, 8, 3,  8, 9,  4, 9, 5, 2, 4, 11, 6,  2, 10, 8,  6, 7, 9, 8, 1};
    auto norm = [&](uint32_t i) {
        float* p = &v[3 * i];
        float l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        p[0] /= l;
        p[1] /= l;
        p[2] /= l;
    };
    for (uint32_t i = 0; i < v.size() / 3; ++i) norm(i);
    for (int s = 0; s < subdiv; ++s) {
        std::vector<uint32_t> nf;
        for (size_t k = 0; k < f.size(); k += 3) {
            uint32_t a = f[k], b = f[k + 1], c = f[k + 2];
            auto mid = [&](uint32_t x, uint32_t y) {
                uint32_t i = (uint32_t)(v.size() / 3);
                v.insert(v.end(), {(v[3 * x] + v[3 * y]) * 0.5f, (v[3 * x + 1] + v[3 * y + 1]) * 0.5f,
                                   (v[3 * x + 2] + v[3 * y + 2]) * 0.5f});
                norm(i);
                return i;
            };
            uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        f.swap(nf);
    }
    for (size_t k = 0; k < f.size(); k += 3) {
        const float* p[3] = {&v[3 * f[k]], &v[3 * f[k + 1]], &v[3 * f[k + 2]]};
        float e1[3], e2[3], n[3];
        for (int a = 0; a < 3; ++a) {
            e1[a] = p[1][a] - p[0][a];
            e2[a] = p[2][a] - p[0][a];
        }
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        int ax = 0;
        if (std::fabs(n[1]) > std::fabs(n[ax])) ax = 1;
        if (std::fabs(n[2]) > std::fabs(n[ax])) ax = 2;
        int ua = (ax + 1) % 3, va = (ax + 2) % 3;
        uint32_t ids[3];
        for (int q = 0; q < 3; ++q)
            ids[q] = m.add_vertex(p[q][0], p[q][1], p[q][2], p[q][0], p[q][1], p[q][2], 0.5f + 0.5f * p[q][ua],
                                  0.5f + 0.5f * p[q][va]);
        m.tri(ids[0], ids[1], ids[2]);
    }
}

// Capped cylinder, radius 1, z in [0, 1], seg x rings side quads.
void add_cylinder(Mesh& m, int seg, int rings) {
    const float two_pi = 6.28318530718f;
    for (int r = 0; r <= rings; ++r)
        for (int s = 0; s <= seg; ++s) {
            float a = two_pi * s / seg;
            float z = (float)r / rings;
            m.add_vertex(std::cos(a), std::sin(a), z, std::cos(a), std::sin(a), 0.0f, 2.0f * s / seg, 4.0f * z);
        }
    for (int r = 0; r < rings; ++r)
        for (int s = 0; s < seg; ++s) {
            uint32_t v00 = r * (seg + 1) + s, v10 = v00 + 1, v01 = v00 + seg + 1, v11 = v01 + 1;
            m.tri(v00, v10, v11);
            m.tri(v00, v11, v01);
        }
    for (int cap = 0; cap < 2; ++cap) {
        float z = (float)cap;
        float nz = cap ? 1.0f : -1.0f;
        uint32_t c = m.add_vertex(0, 0, z, 0, 0, nz, 0.5f, 0.5f);
        uint32_t first = (uint32_t)(m.pos.size() / 3);
        for (int s = 0; s <= seg; ++s) {
            float a = two_pi * s / seg;
            m.add_vertex(std::cos(a), std::sin(a), z, 0, 0, nz, 0.5f + 0.5f * std::cos(a), 0.5f + 0.5f * std::sin(a));
        }
        for (int s = 0; s < seg; ++s) {
            if (cap)
                m.tri(c, first + s, first + s + 1);
            else
                m.tri(c, first + s + 1, first + s);
        }
    }
}

// Half torus (arch) in the xz plane spanning x in [-1, 1], tube radius r.
void add_arch(Mesh& m, int seg, int tube, float r) {
    const float pi = 3.14159265359f;
    for (int s = 0; s <= seg; ++s) {
        float a = pi * s / seg;
        float cx = std::cos(a), cz = std::sin(a);
        for (int k = 0; k <= tube; ++k) {
            float b = 2.0f * pi * k / tube;
            float nx = std::cos(b) * cx, nz = std::cos(b) * cz, ny = std::sin(b);
            m.add_vertex(cx + r * nx, r * ny, cz + r * nz, nx, ny, nz, 4.0f * s / seg, (float)k / tube);
        }
    }
    for (int s = 0; s < seg; ++s)
        for (int k = 0; k < tube; ++k) {
            uint32_t v00 = s * (tube + 1) + k, v01 = v00 + 1, v10 = v00 + tube + 1, v11 = v10 + 1;
            m.tri(v00, v10, v11);
            m.tri(v00, v11, v01);
        }
}

void identity(float* t) {
    std::memset(t, 0, 16 * sizeof(float));
    t[0] = t[5] = t[10] = t[15] = 1.0f;
}
// T * Rz(angle) * S, column-major.
void trs(float* t, float tx, float ty, float tz, float angle_z, float sx, float sy, float sz) {
    identity(t);
    float c = std::cos(angle_z), s = std::sin(angle_z);
    t[0] = c * sx;
    t[1] = s * sx;
    t[4] = -s * sy;
    t[5] = c * sy;
    t[10] = sz;
    t[12] = tx;
    t[13] = ty;
    t[14] = tz;
}

struct Texture {
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

float hash01(uint32_t x, uint32_t y, uint32_t s) {
    uint32_t h = x * 374761393u + y * 668265263u + s * 2246822519u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return (float)(h & 0xFFFFFF) / 16777216.0f;
}
float value_noise(float x, float y, uint32_t s) {
    int xi = (int)std::floor(x), yi = (int)std::floor(y);
    float fx = x - xi, fy = y - yi;
    fx = fx * fx * (3 - 2 * fx);
    fy = fy * fy * (3 - 2 * fy);
    float a = hash01(xi, yi, s), b = hash01(xi + 1, yi, s), c = hash01(xi, yi + 1, s), d = hash01(xi + 1, yi + 1, s);
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
}
uint8_t to8(float x) { return (uint8_t)std::lround(std::fmin(std::fmax(x, 0.0f), 1.0f) * 255.0f); }

Texture make_texture(int kind, uint32_t size, uint32_t seed) {
    Texture t;
    t.w = t.h = size;
    t.rgba.resize((size_t)size * size * 4);
    for (uint32_t y = 0; y < size; ++y)
        for (uint32_t x = 0; x < size; ++x) {
            float u = (float)x / size, v = (float)y / size;
            float r = 1, g = 1, b = 1, a = 1;
            float n = value_noise(u * 32, v * 32, seed) * 0.5f + value_noise(u * 128, v * 128, seed + 1) * 0.25f;
            switch (kind) {
                case 0: {  // floor tiles with grout
                    int cx = (int)(u * 8), cy = (int)(v * 8);
                    float fx = u * 8 - cx, fy = v * 8 - cy;
                    bool grout = fx < 0.04f || fy < 0.04f;
                    float base = ((cx + cy) & 1) ? 0.82f : 0.35f;
                    r = g = b = grout ? 0.15f : base * (0.85f + 0.3f * n);
                    b *= 0.92f;
                    break;
                }
                case 1: {  // bricks
                    float by = v * 16;
                    int row = (int)by;
                    float bx = u * 8 + ((row & 1) ? 0.5f : 0.0f);
                    float fx = bx - std::floor(bx), fy = by - row;
                    bool mortar = fx < 0.05f || fy < 0.08f;
                    float k = 0.8f + 0.4f * hash01((uint32_t)std::floor(bx), row, seed);
                    r = mortar ? 0.7f : 0.55f * k * (0.8f + 0.4f * n);
                    g = mortar ? 0.68f : 0.25f * k * (0.8f + 0.4f * n);
                    b = mortar ? 0.62f : 0.18f * k;
                    break;
                }
                case 2: {  // marble
                    float m = 0.5f + 0.5f * std::sin(12.0f * u + 8.0f * n + 3.0f * v);
                    r = 0.75f + 0.2f * m;
                    g = 0.72f + 0.2f * m;
                    b = 0.68f + 0.2f * m;
                    break;
                }
                case 3: {  // wood
                    float w = u * 40 + 6 * n;
                    float ring = w - std::floor(w);
                    r = 0.45f + 0.2f * ring;
                    g = 0.28f + 0.12f * ring;
                    b = 0.12f + 0.05f * ring;
                    break;
                }
                case 4: {  // leaves: alpha-masked ellipses on transparent background
                    float best = 1e9f;
                    int gx = (int)(u * 16), gy = (int)(v * 16);
                    for (int oy = -1; oy <= 1; ++oy)
                        for (int ox = -1; ox <= 1; ++ox) {
                            int cx = gx + ox, cy = gy + oy;
                            float px = (cx + hash01(cx & 15, cy & 15, seed)) / 16.0f;
                            float py = (cy + hash01(cx & 15, cy & 15, seed + 7)) / 16.0f;
                            float dx = (u - px) * 16, dy = (v - py) * 16 * 1.8f;
                            best = std::fmin(best, dx * dx + dy * dy);
                        }
                    a = best < 0.35f ? 1.0f : 0.0f;
                    r = 0.12f + 0.1f * n;
                    g = 0.35f + 0.25f * n;
                    b = 0.08f;
                    break;
                }
                case 5: {  // grille: square holes
                    float fx = u * 24 - std::floor(u * 24), fy = v * 24 - std::floor(v * 24);
                    a = (fx > 0.2f && fy > 0.2f) ? 0.0f : 1.0f;
                    r = g = b = 0.25f + 0.1f * n;
                    break;
                }
                case 6: {  // tangent-space normal map (bumpy tiles)
                    float hx = value_noise(u * 64 + 0.01f, v * 64, seed) - value_noise(u * 64 - 0.01f, v * 64, seed);
                    float hy = value_noise(u * 64, v * 64 + 0.01f, seed) - value_noise(u * 64, v * 64 - 0.01f, seed);
                    float nx = -hx * 6, ny = -hy * 6, nz = 1;
                    float l = std::sqrt(nx * nx + ny * ny + nz * nz);
                    r = 0.5f + 0.5f * nx / l;
                    g = 0.5f + 0.5f * ny / l;
                    b = 0.5f + 0.5f * nz / l;
                    break;
                }
                default: {  // specular (rgb, sRGB) + roughness (a)
                    float s = 0.2f + 0.6f * n;
                    r = g = b = s;
                    a = 0.15f + 0.7f * value_noise(u * 16, v * 16, seed + 3);
                    break;
                }
            }
            uint8_t* px = &t.rgba[4 * ((size_t)y * size + x)];
            px[0] = to8(r);
            px[1] = to8(g);
            px[2] = to8(b);
            px[3] = to8(a);
        }
    return t;
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

struct pbrt_generated_scene {
    uint32_t config = 0;
    std::vector<Mesh> meshes;
    std::vector<pbrt_instance_desc> instances;
    std::vector<Texture> textures;
    std::vector<pbrt_light> lights;
    std::vector<pbrt_mesh_desc> mesh_descs;
    std::vector<pbrt_texture_desc> tex_descs;
    std::vector<pbrt_camera_pose> camera_keys;
    uint32_t path_frames = 64;  // frames of one loop of the camera path
    pbrt_camera_pose camera{};
};

namespace {

float catmull_rom(float p0, float p1, float p2, float p3, float t) {
    float t2 = t * t, t3 = t2 * t;
    return 0.5f * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t3);
}

void add_instance(pbrt_generated_scene& g, uint32_t mesh, const float* t) {
    pbrt_instance_desc in{};
    std::memcpy(in.transform, t, sizeof(in.transform));
    in.mesh = mesh;
    g.instances.push_back(in);
}

void build_cornell(pbrt_generated_scene& g, uint64_t seed) {
    Pcg32 rng(seed);
    const float wall_white[3] = {0.73f, 0.73f, 0.73f}, red[3] = {0.63f, 0.065f, 0.05f},
                green[3] = {0.14f, 0.45f, 0.09f};
    struct W {
        float p0[3], p1[3], p3[3];
        const float* col;
    } walls[5] = {
        {{-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, wall_white},  // floor (+z)
        {{-1, 1, 2}, {1, 1, 2}, {-1, -1, 2}, wall_white},   // ceiling (-z)
        {{-1, 1, 0}, {1, 1, 0}, {-1, 1, 2}, wall_white},    // back (-y)
        {{-1, 1, 0}, {-1, 1, 2}, {-1, -1, 0}, red},         // left (+x)
        {{1, -1, 0}, {1, -1, 2}, {1, 1, 0}, green},         // right (-x)
    };
    float ident[16];
    identity(ident);
    for (auto& w : walls) {
        Mesh m;
        m.zero_normals = true;  // exercises the area-weighted normal rule
        add_quad(m, w.p0, w.p1, w.p3, 1, 1, 1.0f, 1.0f, false);
        m.mat = make_material(w.col[0], w.col[1], w.col[2], 1.0f, false);
        g.meshes.push_back(m);
        add_instance(g, (uint32_t)g.meshes.size() - 1, ident);
    }
    // tall inner box (6 quads)
    {
        Mesh m;
        add_unit_box(m, 1.0f);
        float alb[3] = {rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f)};
        float rough = rng.uniform(0.05f, 1.0f);
        bool metal = rng.uniform() < 0.3f;
        m.mat = make_material(alb[0], alb[1], alb[2], rough, metal);
        g.meshes.push_back(m);
        float t[16];
        trs(t, -0.45f, 0.35f, 0.6f, 0.35f, 0.5f, 0.5f, 1.2f);
        add_instance(g, (uint32_t)g.meshes.size() - 1, t);
    }
    // 0.2 x 0.2 ceiling light facing down
    {
        Mesh m;
        float p0[3] = {-0.1f, 0.1f, 1.98f}, p1[3] = {0.1f, 0.1f, 1.98f}, p3[3] = {-0.1f, -0.1f, 1.98f};
        add_quad(m, p0, p1, p3, 1, 1, 1.0f, 1.0f, true);
        m.mat = make_material(0.78f, 0.78f, 0.78f, 1.0f, false);
        m.mat.emission_texture_id[0] = 8.0f;
        m.mat.emission_texture_id[1] = 7.0f;
        m.mat.emission_texture_id[2] = 5.5f;
        m.mat.category_id = 1;
        g.meshes.push_back(m);
        add_instance(g, (uint32_t)g.meshes.size() - 1, ident);
    }
    // 16 non-overlapping icospheres
    struct S {
        float x, y, z, r;
    };
    std::vector<S> placed;
    int attempts = 0;
    while (placed.size() < 16 && attempts < 100000) {
        ++attempts;
        float r = rng.uniform(0.1f, 0.3f);
        float x = rng.uniform(-1.0f + r + 0.01f, 1.0f - r - 0.01f);
        float y = rng.uniform(-1.0f + r + 0.01f, 1.0f - r - 0.01f);
        float z = rng.uniform(r + 0.01f, 1.8f - r);
        bool ok = true;
        for (auto& o : placed) {
            float d = std::sqrt((x - o.x) * (x - o.x) + (y - o.y) * (y - o.y) + (z - o.z) * (z - o.z));
            if (d < r + o.r + 0.02f) ok = false;
        }
        float bx = x + 0.45f, by = y - 0.35f;
        if (std::sqrt(bx * bx + by * by) < r + 0.4f && z < 1.2f + r + 0.02f) ok = false;
        if (!ok) continue;
        placed.push_back({x, y, z, r});
    }
    for (auto& s : placed) {
        Mesh m;
        m.use16 = true;
        add_icosphere(m, 2);
        float alb[3] = {rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f)};
        float rough = rng.uniform(0.05f, 1.0f);
        bool metal = rng.uniform() < 0.3f;
        m.mat = make_material(alb[0], alb[1], alb[2], rough, metal);
        m.mat.category_id = 2;
        g.meshes.push_back(m);
        float t[16];
        trs(t, s.x, s.y, s.z, 0.0f, s.r, s.r, s.r);
        add_instance(g, (uint32_t)g.meshes.size() - 1, t);
    }
    g.camera = pbrt_camera_pose{{0.0f, -3.0f, 1.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 1.0f}, 60.0f, 0.1f, 1000.0f};
}

// the atrium's camera path: a closed Catmull-Rom loop through 8 keys (eye, look-at)
// at eye height in the gallery between the colonnade and the walls
constexpr float kAtriumKeys[8][6] = {{-14, -14, 1.7f, 0, 0, 3}, {-4, -16, 2.0f, 2, 0, 4}, {10, -13, 1.8f, 0, 2, 3},
                                     {15, -2, 2.5f, -2, 2, 5},  {12, 12, 1.7f, 0, 0, 4},  {0, 15, 2.2f, 0, -2, 3},
                                     {-13, 10, 1.9f, 2, -2, 6}, {-16, -4, 1.7f, 2, 0, 3}};

// the eye's ground-plane track, `n` points per key interval
std::vector<std::pair<float, float>> atrium_eye_track(int n) {
    std::vector<std::pair<float, float>> out;
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < n; ++j) {
            const float* k0 = kAtriumKeys[(i + 7) % 8];
            const float* k1 = kAtriumKeys[i];
            const float* k2 = kAtriumKeys[(i + 1) % 8];
            const float* k3 = kAtriumKeys[(i + 2) % 8];
            const float t = (float)j / (float)n;
            out.push_back({catmull_rom(k0[0], k1[0], k2[0], k3[0], t), catmull_rom(k0[1], k1[1], k2[1], k3[1], t)});
        }
    return out;
}

void build_atrium(pbrt_generated_scene& g, uint64_t seed) {
    Pcg32 rng(seed);
    for (int k = 0; k < 8; ++k) g.textures.push_back(make_texture(k, 1024, (uint32_t)(seed * 131 + k)));
    float ident[16];
    identity(ident);
    auto textured = [&](Mesh& m, int diffuse, float tint, float rough) {
        m.mat = make_material(tint, tint, tint, rough, false);
        m.tex[0] = diffuse;
    };
    // floor 40 x 40 with tile texture + normal map
    {
        Mesh m;
        float p0[3] = {-20, -20, 0}, p1[3] = {20, -20, 0}, p3[3] = {-20, 20, 0};
        add_quad(m, p0, p1, p3, 32, 32, 8.0f, 8.0f, true);
        textured(m, 0, 1.0f, 0.35f);
        m.tex[1] = 6;
        m.mat.category_id = 1;
        g.meshes.push_back(m);
        add_instance(g, (uint32_t)g.meshes.size() - 1, ident);
    }
    // four brick walls, height 16
    {
        const float H = 16.0f;
        float w[4][3][3] = {{{-20, 20, 0}, {20, 20, 0}, {-20, 20, H}},
                            {{20, -20, 0}, {-20, -20, 0}, {20, -20, H}},
                            {{-20, -20, 0}, {-20, 20, 0}, {-20, -20, H}},
                            {{20, 20, 0}, {20, -20, 0}, {20, 20, H}}};
        for (auto& q : w) {
            Mesh m;
            add_quad(m, q[0], q[1], q[2], 32, 16, 6.0f, 3.0f, true);
            textured(m, 1, 1.0f, 0.8f);
            m.mat.category_id = 2;
            g.meshes.push_back(m);
            add_instance(g, (uint32_t)g.meshes.size() - 1, ident);
        }
    }
    // gallery slabs at two levels (boxes)
    uint32_t slab_mesh;
    {
        Mesh m;
        add_unit_box(m, 4.0f);
        textured(m, 2, 0.9f, 0.5f);
        m.mat.category_id = 3;
        g.meshes.push_back(m);
        slab_mesh = (uint32_t)g.meshes.size() - 1;
        for (float z : {6.0f, 11.0f}) {
            float t[16];
            trs(t, 0, 17.5f, z, 0, 40, 5, 0.4f);
            add_instance(g, slab_mesh, t);
            trs(t, 0, -17.5f, z, 0, 40, 5, 0.4f);
            add_instance(g, slab_mesh, t);
            trs(t, 17.5f, 0, z, 0, 5, 30, 0.4f);
            add_instance(g, slab_mesh, t);
            trs(t, -17.5f, 0, z, 0, 5, 30, 0.4f);
            add_instance(g, slab_mesh, t);
        }
    }
    // three column meshes (marble + specular map, wood, plain metal), 72 instances
    uint32_t col_mesh[3];
    for (int k = 0; k < 3; ++k) {
        Mesh m;
        add_cylinder(m, 64, 22);
        if (k == 0) {
            textured(m, 2, 1.0f, 0.4f);
            m.tex[2] = 7;
        } else if (k == 1) {
            textured(m, 3, 1.0f, 0.6f);
        } else {
            m.mat = make_material(0.9f, 0.75f, 0.45f, 0.3f, true);
        }
        m.mat.category_id = 4;
        g.meshes.push_back(m);
        col_mesh[k] = (uint32_t)g.meshes.size() - 1;
    }
    std::vector<std::pair<float, float>> col_xy;
    for (int i = 0; i < 6; ++i) {
        float c = -12.5f + 5.0f * i;
        col_xy.push_back({c, 12.5f});
        col_xy.push_back({c, -12.5f});
    }
    for (int i = 1; i < 5; ++i) {
        float c = -12.5f + 5.0f * i;
        col_xy.push_back({12.5f, c});
        col_xy.push_back({-12.5f, c});
    }
    // 20 column positions x 3 levels (+ 12 free-standing inner columns) = 72
    for (int level = 0; level < 3; ++level) {
        float z0 = level * 5.5f, h = level == 2 ? 4.5f : 5.5f;
        for (size_t i = 0; i < col_xy.size(); ++i) {
            float t[16];
            trs(t, col_xy[i].first, col_xy[i].second, z0, 0, 0.45f, 0.45f, h);
            add_instance(g, col_mesh[(i + level) % 3], t);
        }
    }
    for (int i = 0; i < 12; ++i) {
        float a = 6.28318530718f * i / 12;
        float t[16];
        trs(t, 6.0f * std::cos(a), 6.0f * std::sin(a), 0, a, 0.3f, 0.3f, rng.log_uniform(1.0f, 8.0f));
        add_instance(g, col_mesh[i % 3], t);
    }
    // arches between neighbouring perimeter columns (2 levels)
    {
        Mesh m;
        add_arch(m, 32, 8, 0.18f);
        textured(m, 3, 0.9f, 0.5f);
        m.mat.category_id = 5;
        g.meshes.push_back(m);
        uint32_t arch = (uint32_t)g.meshes.size() - 1;
        for (int level = 0; level < 2; ++level) {
            float z = 5.5f * level + 4.0f;
            for (int i = 0; i < 5; ++i) {
                float c = -10.0f + 5.0f * i;
                float t[16];
                trs(t, c, 12.5f, z, 0, 2.5f, 1, 1.3f);
                add_instance(g, arch, t);
                trs(t, c, -12.5f, z, 0, 2.5f, 1, 1.3f);
                add_instance(g, arch, t);
                trs(t, 12.5f, c, z, 1.5707963f, 2.5f, 1, 1.3f);
                add_instance(g, arch, t);
                trs(t, -12.5f, c, z, 1.5707963f, 2.5f, 1, 1.3f);
                add_instance(g, arch, t);
            }
        }
    }
    // 200 boxes with log-uniform scale in [0.1, 10]
    {
        const auto track = atrium_eye_track(64);
        uint32_t box_mesh[4];
        for (int k = 0; k < 4; ++k) {
            Mesh m;
            add_unit_box(m, 1.0f);
            float alb[3] = {rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f)};
            bool metal = rng.uniform() < 0.3f;
            m.mat = make_material(alb[0], alb[1], alb[2], rng.uniform(0.05f, 1.0f), metal);
            if (k == 1) m.tex[0] = 3;
            m.mat.category_id = 6;
            g.meshes.push_back(m);
            box_mesh[k] = (uint32_t)g.meshes.size() - 1;
        }
        for (int i = 0; i < 200; ++i) {
            float s = rng.log_uniform(0.1f, 10.0f);
            float sx = s * rng.uniform(0.5f, 1.0f), sy = s * rng.uniform(0.5f, 1.0f), sz = s * rng.uniform(0.2f, 1.0f);
            sx = std::fmin(sx, 6.0f);
            sy = std::fmin(sy, 6.0f);
            sz = std::fmin(sz, 4.0f);
            // positions whose bounding circle comes within 1.5 of the camera's track are
            // drawn again (a box without such a position in 256 draws is left out), so the
            // camera never enters or grazes a box
            const float clear = 0.5f * std::sqrt(sx * sx + sy * sy) + 1.5f;
            float x = 0.0f, y = 0.0f;
            bool ok = false;
            for (int tries = 0; tries < 256 && !ok; ++tries) {
                x = rng.uniform(-18.0f, 18.0f);
                y = rng.uniform(-18.0f, 18.0f);
                ok = true;
                for (const auto& e : track)
                    if ((e.first - x) * (e.first - x) + (e.second - y) * (e.second - y) < clear * clear) {
                        ok = false;
                        break;
                    }
            }
            const float angle = rng.uniform(0.0f, 6.2831853f);
            if (!ok) continue;
            float t[16];
            trs(t, x, y, 0.5f * sz, angle, sx, sy, sz);
            add_instance(g, box_mesh[i % 4], t);
        }
    }
    // alpha-masked foliage panels (any-hit) and gallery grilles
    {
        Mesh leaves;
        float p0[3] = {-0.5f, 0, 0}, p1[3] = {0.5f, 0, 0}, p3[3] = {-0.5f, 0, 1};
        add_quad(leaves, p0, p1, p3, 8, 8, 1.0f, 1.0f, true);
        textured(leaves, 4, 1.0f, 0.7f);
        leaves.mat.category_id = 7;
        g.meshes.push_back(leaves);
        uint32_t leaf_mesh = (uint32_t)g.meshes.size() - 1;
        for (int i = 0; i < 120; ++i) {
            float a = rng.uniform(0.0f, 6.2831853f), r = rng.uniform(2.0f, 10.0f);
            float t[16];
            trs(t, r * std::cos(a), r * std::sin(a), 0.0f, rng.uniform(0.0f, 3.14159f), rng.uniform(1.0f, 2.5f), 1,
                rng.uniform(1.0f, 3.0f));
            add_instance(g, leaf_mesh, t);
        }
        Mesh grille;
        add_quad(grille, p0, p1, p3, 8, 8, 2.0f, 1.0f, true);
        textured(grille, 5, 0.8f, 0.4f);
        grille.mat = make_material(0.6f, 0.6f, 0.62f, 0.35f, true);
        grille.tex[0] = 5;
        grille.mat.category_id = 8;
        g.meshes.push_back(grille);
        uint32_t grille_mesh = (uint32_t)g.meshes.size() - 1;
        for (int level = 0; level < 2; ++level) {
            float z = level == 0 ? 6.2f : 11.2f;
            for (int i = 0; i < 10; ++i) {
                float c = -13.5f + 3.0f * i;
                float t[16];
                trs(t, c, 15.0f, z, 0, 3, 1, 1.1f);
                add_instance(g, grille_mesh, t);
                trs(t, c, -15.0f, z, 0, 3, 1, 1.1f);
                add_instance(g, grille_mesh, t);
            }
        }
    }
    // three emissive quads (area lights, 2 triangles each)
    {
        const float em[3][3] = {{30, 26, 20}, {20, 24, 30}, {28, 28, 24}};
        const float pos[3][3] = {{-8, 0, 14.5f}, {8, 4, 14.5f}, {0, -8, 9.5f}};
        for (int k = 0; k < 3; ++k) {
            Mesh m;
            float x = pos[k][0], y = pos[k][1], z = pos[k][2];
            float p0[3] = {x - 0.6f, y + 0.6f, z}, p1[3] = {x + 0.6f, y + 0.6f, z}, p3[3] = {x - 0.6f, y - 0.6f, z};
            add_quad(m, p0, p1, p3, 1, 1, 1.0f, 1.0f, true);
            m.mat = make_material(0.8f, 0.8f, 0.8f, 1.0f, false);
            m.mat.emission_texture_id[0] = em[k][0];
            m.mat.emission_texture_id[1] = em[k][1];
            m.mat.emission_texture_id[2] = em[k][2];
            m.mat.category_id = 9;
            g.meshes.push_back(m);
            add_instance(g, (uint32_t)g.meshes.size() - 1, ident);
        }
    }
    // the sun: directional light, attenuation (0.5, 0, 0) like the default sun
    {
        pbrt_light l{};
        l.v0_type[3] = 1.0f;
        float d[3] = {0.35f, 0.25f, -0.9f};
        float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; ++a) l.dir_angle2[a] = d[a] / len;
        for (int a = 0; a < 3; ++a) {
            l.col_ambient[a] = 1.2f;
            l.col_diffuse[a] = 1.2f;
            l.col_specular[a] = 0.0f;
        }
        l.col_ambient[2] = 1.05f;
        l.strengths[0] = 0.5f;
        g.lights.push_back(l);
    }
    for (auto& k : kAtriumKeys)
        g.camera_keys.push_back(pbrt_camera_pose{{k[0], k[1], k[2]}, {k[3], k[4], k[5]}, {0, 0, 1}, 60.0f, 0.1f, 1000.0f});
    g.camera = g.camera_keys[0];
}

// Config 5: a city of 4096 axis-aligned towers (6 subdivided box variants of
// ~1.2k triangles each, log-uniform heights) on a 64 x 64 block grid, a ground
// plane, 65,536 small emissive window quads on random facades (131,072 emissive
// triangles, so light sampling switches to UNIFORM) and a dim sun; a 128-frame
// Catmull-Rom flight over the roofs.
void build_city(pbrt_generated_scene& g, uint64_t seed) {
    Pcg32 rng(seed);
    float ident[16];
    identity(ident);
    const int grid = 64;
    const float spacing = 20.0f, half = 0.5f * grid * spacing;
    {
        Mesh ground;
        float p0[3] = {-half, -half, 0}, p1[3] = {half, -half, 0}, p3[3] = {-half, half, 0};
        add_quad(ground, p0, p1, p3, 64, 64, 64.0f, 64.0f, true);
        ground.mat = make_material(0.18f, 0.18f, 0.19f, 0.8f, false);
        ground.mat.category_id = 1;
        g.meshes.push_back(ground);
        add_instance(g, 0, ident);
    }
    // building variants: unit box [-0.5, 0.5]^2 x [0, 1], sides subdivided 8 x 16, roof 8 x 8
    const int variants = 6;
    const uint32_t first_building = (uint32_t)g.meshes.size();
    for (int v = 0; v < variants; ++v) {
        Mesh m;
        m.use16 = true;
        const float c[4][3] = {{-0.5f, -0.5f, 0}, {0.5f, -0.5f, 0}, {0.5f, 0.5f, 0}, {-0.5f, 0.5f, 0}};
        for (int f = 0; f < 4; ++f) {
            const float* a = c[f];
            const float* b = c[(f + 1) % 4];
            float top[3] = {a[0], a[1], 1.0f};
            add_quad(m, a, b, top, 8, 16, 4.0f, 8.0f, true);
        }
        float r0[3] = {-0.5f, -0.5f, 1.0f}, r1[3] = {0.5f, -0.5f, 1.0f}, r3[3] = {-0.5f, 0.5f, 1.0f};
        add_quad(m, r0, r1, r3, 8, 8, 1.0f, 1.0f, true);
        const float tone = 0.35f + 0.1f * v;
        const bool metal = v == 5;
        m.mat = make_material(tone, tone * (0.95f + 0.02f * v), tone * 0.9f, 0.3f + 0.1f * v, metal);
        m.mat.category_id = (uint32_t)(2 + v);
        g.meshes.push_back(m);
    }
    struct Tower {
        float x, y, sx, sy, sz;
    };
    std::vector<Tower> towers;
    towers.reserve(grid * grid);
    for (int j = 0; j < grid; ++j)
        for (int i = 0; i < grid; ++i) {
            Tower t;
            t.x = -half + (i + 0.5f) * spacing + rng.uniform(-2.0f, 2.0f);
            t.y = -half + (j + 0.5f) * spacing + rng.uniform(-2.0f, 2.0f);
            t.sx = rng.uniform(8.0f, 14.0f);
            t.sy = rng.uniform(8.0f, 14.0f);
            t.sz = rng.log_uniform(10.0f, 120.0f);
            float tr[16];
            trs(tr, t.x, t.y, 0.0f, 0.0f, t.sx, t.sy, t.sz);
            add_instance(g, first_building + (uint32_t)(rng.next() % variants), tr);
            towers.push_back(t);
        }
    // emissive windows: a unit quad in the local y-z plane facing +x
    {
        Mesh w;
        float p0[3] = {0, -0.5f, -0.5f}, p1[3] = {0, 0.5f, -0.5f}, p3[3] = {0, -0.5f, 0.5f};
        add_quad(w, p0, p1, p3, 1, 1, 1.0f, 1.0f, true);
        w.mat = make_material(0.9f, 0.85f, 0.7f, 1.0f, false);
        w.mat.emission_texture_id[0] = 6.0f;
        w.mat.emission_texture_id[1] = 5.0f;
        w.mat.emission_texture_id[2] = 3.5f;
        w.mat.category_id = 9;
        g.meshes.push_back(w);
        const uint32_t window = (uint32_t)g.meshes.size() - 1;
        for (int k = 0; k < 65536; ++k) {
            const Tower& t = towers[rng.next() % towers.size()];
            const int face = (int)(rng.next() % 4);  // +x, +y, -x, -y
            const float along = rng.uniform(-0.45f, 0.45f), z = rng.uniform(2.0f, t.sz - 1.0f);
            float x = t.x, y = t.y, ang = 0.0f;
            const float off = 0.02f;
            if (face == 0) {
                x += 0.5f * t.sx + off;
                y += along * t.sy;
            } else if (face == 1) {
                y += 0.5f * t.sy + off;
                x += along * t.sx;
                ang = 1.5707963f;
            } else if (face == 2) {
                x -= 0.5f * t.sx + off;
                y += along * t.sy;
                ang = 3.1415927f;
            } else {
                y -= 0.5f * t.sy + off;
                x += along * t.sx;
                ang = 4.7123890f;
            }
            float tr[16];
            trs(tr, x, y, z, ang, 1.0f, 0.8f, 1.2f);
            add_instance(g, window, tr);
        }
    }
    {
        pbrt_light l{};
        l.v0_type[3] = 1.0f;
        float d[3] = {0.3f, 0.5f, -0.8f};
        float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; ++a) l.dir_angle2[a] = d[a] / len;
        for (int a = 0; a < 3; ++a) {
            l.col_ambient[a] = 0.3f;
            l.col_diffuse[a] = 0.3f;
        }
        l.strengths[0] = 0.5f;
        g.lights.push_back(l);
    }
    const float keys[8][6] = {{-500, -500, 160, 0, 0, 20},   {-100, -560, 120, 100, 0, 30},
                              {420, -420, 180, 0, 0, 10},     {560, 60, 140, -100, 50, 40},
                              {380, 420, 100, 0, 0, 30},      {0, 560, 150, 0, -100, 20},
                              {-420, 380, 200, 50, 0, 10},    {-580, -60, 130, 0, 0, 40}};
    for (auto& k : keys)
        g.camera_keys.push_back(
            pbrt_camera_pose{{k[0], k[1], k[2]}, {k[3], k[4], k[5]}, {0, 0, 1}, 60.0f, 0.5f, 5000.0f});
    g.path_frames = 128;
    g.camera = g.camera_keys[0];
}

void finalize(pbrt_generated_scene& g) {
    g.mesh_descs.clear();
    for (Mesh& m : g.meshes) {
        pbrt_mesh_desc d{};
        d.vertex_count = (uint32_t)(m.pos.size() / 3);
        d.positions = m.pos.data();
        if (m.zero_normals) std::fill(m.nor.begin(), m.nor.end(), 0.0f);
        d.normals = m.nor.data();
        d.uvs = m.uv.data();
        if (m.use16) {
            d.index_count = (uint32_t)m.idx16.size();
            d.index_stride = 2;
            d.indices = m.idx16.data();
        } else {
            d.index_count = (uint32_t)m.idx32.size();
            d.index_stride = 4;
            d.indices = m.idx32.data();
        }
        d.material = m.mat;
        d.diffuse_texture = m.tex[0];
        d.normal_texture = m.tex[1];
        d.specular_texture = m.tex[2];
        d.emissive_texture = m.tex[3];
        g.mesh_descs.push_back(d);
    }
    g.tex_descs.clear();
    for (Texture& t : g.textures) g.tex_descs.push_back(pbrt_texture_desc{t.w, t.h, t.rgba.data()});
}

}  // namespace

extern "C" int32_t pbrt_generate_scene(uint32_t config, uint64_t seed, pbrt_generated_scene** out_scene) {
    return guard([&] {
        PBRT_REQUIRE(out_scene, "pbrt_generate_scene: NULL output");
        *out_scene = nullptr;
        auto g = std::make_unique<pbrt_generated_scene>();
        g->config = config;
        if (config == 1)
            build_cornell(*g, seed);
        else if (config == 2)
            build_atrium(*g, seed);
        else if (config == 5)
            build_city(*g, seed);
        else
            fail(PBRT_ERR_UNSUPPORTED, "pbrt_generate_scene: unknown config " + std::to_string(config));
        finalize(*g);
        *out_scene = g.release();
    });
}

extern "C" int32_t pbrt_generated_scene_desc(const pbrt_generated_scene* g, pbrt_scene_desc* d) {
    return guard([&] {
        PBRT_REQUIRE(g && d, "pbrt_generated_scene_desc: NULL argument");
        std::memset(d, 0, sizeof(*d));
        d->mesh_count = (uint32_t)g->mesh_descs.size();
        d->meshes = g->mesh_descs.data();
        d->instance_count = (uint32_t)g->instances.size();
        d->instances = g->instances.data();
        d->texture_count = (uint32_t)g->tex_descs.size();
        d->textures = g->tex_descs.empty() ? nullptr : g->tex_descs.data();
        d->light_count = (uint32_t)g->lights.size();
        d->lights = g->lights.empty() ? nullptr : g->lights.data();
        d->bvh_builder = PBRT_BVH_SAH_HOST;
    });
}

extern "C" int32_t pbrt_generated_scene_camera(const pbrt_generated_scene* g, uint32_t frame,
                                               pbrt_camera_pose* out) {
    return guard([&] {
        PBRT_REQUIRE(g && out, "pbrt_generated_scene_camera: NULL argument");
        if (g->camera_keys.empty()) {
            *out = g->camera;
            return;
        }
        // path_frames frames over the closed loop of keys
        const uint32_t n = (uint32_t)g->camera_keys.size();
        float s = (float)(frame % g->path_frames) / (float)g->path_frames * n;
        uint32_t i1 = (uint32_t)s % n;
        float t = s - std::floor(s);
        const pbrt_camera_pose& k0 = g->camera_keys[(i1 + n - 1) % n];
        const pbrt_camera_pose& k1 = g->camera_keys[i1];
        const pbrt_camera_pose& k2 = g->camera_keys[(i1 + 1) % n];
        const pbrt_camera_pose& k3 = g->camera_keys[(i1 + 2) % n];
        *out = k1;
        for (int a = 0; a < 3; ++a) {
            out->eye[a] = catmull_rom(k0.eye[a], k1.eye[a], k2.eye[a], k3.eye[a], t);
            out->center[a] = catmull_rom(k0.center[a], k1.center[a], k2.center[a], k3.center[a], t);
        }
    });
}

extern "C" int32_t pbrt_generated_scene_destroy(pbrt_generated_scene* g) {
    return guard([&] { delete g; });
}
