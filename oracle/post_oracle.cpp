// CPU oracle of the screen-space passes (test infrastructure only).
//
// Restates accumulator.comp, AccumulationBuffer::copy_to_back_images, taa.comp
// with Taa::copy_final_image, formatConverter.comp and the BMFR passes
// (bmfrGeneral.comp, bmfrPre.comp, bmfrFit.comp, bmfrPost.comp) in FP32, plus
// an FP64 least-squares solve of each BMFR block (Householder QR in double).
// Image byte orders and the undefined-behaviour rules are the ones documented
// in the product headers (RG16F motion written in full, B8G8R8A8 byte order,
// raw-byte TAA history copy, RNE half/unorm stores, FP32 bilinear REPEAT).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

namespace orc {

using V3f = V3<float>;

static int wrap(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// texture() with LINEAR / REPEAT on a single image; fetch(x, y, out[4])
template <class F>
static void bilinear(F fetch, int w, int h, float u, float v, float* out, int nc) {
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float x0f = std::floor(x), y0f = std::floor(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = wrap((int)x0f, w), x1 = wrap((int)x0f + 1, w), y0 = wrap((int)y0f, h), y1 = wrap((int)y0f + 1, h);
    float a[4], b[4], c[4], d[4];
    fetch(x0, y0, a);
    fetch(x1, y0, b);
    fetch(x0, y1, c);
    fetch(x1, y1, d);
    float gx = 1.0f - fx, gy = 1.0f - fy;
    for (int k = 0; k < nc; ++k) {
        float top = a[k] * gx + b[k] * fx;
        float bot = c[k] * gx + d[k] * fx;
        out[k] = top * gy + bot * fy;
    }
}

static void inverse4d(const float* m, float* out) {
    // Gauss-Jordan in double on the column-major matrix
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[4 * c + r];
            a[r][4 + c] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[p][k]);
        double piv = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= piv;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
            }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[4 * c + r] = (float)a[r][4 + c];
}

}  // namespace orc

using namespace orc;

extern "C" {

// accumulator.comp (SEPARATE_MATRICES).  consts: inv_proj[16], inv_view[16], prev_view[16], prev_origin[4]
void orc_accumulate(uint32_t w, uint32_t h, uint32_t frame, const float* consts, const float* illum,
                    const float* depth, const float* prev_depth, const uint8_t* prev_spp, const uint16_t* prev_acc,
                    uint16_t* motion, uint8_t* spp, uint16_t* acc) {
    const float* inv_proj = consts;
    const float* inv_view = consts + 16;
    const float* prev_view = consts + 32;
    const float* prev_origin = consts + 48;
    float proj[16], m[16];
    inverse4d(inv_proj, proj);
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += proj[4 * k + row] * prev_view[4 * col + k];
            m[4 * col + row] = s;
        }
    const int iw = (int)w, ih = (int)h;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            size_t pix = (size_t)y * w + x;
            float pixel_spp = 1.0f / 256.0f;
            float d = depth[pix];
            V4<float> pos{inv_view[12], inv_view[13], inv_view[14], 1.0f};
            float cx = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
            float cy = ((float)y + 0.5f) / (float)h * 2.0f - 1.0f;
            V4<float> dir = mat4_mul<float>(inv_proj, V4<float>{cx, cy, 1.0f, 1.0f});
            V3f dn = normalize(V3f(dir.x, dir.y, dir.z));
            dir = mat4_mul<float>(inv_view, V4<float>{dn.x, dn.y, dn.z, 0.0f});
            V4<float> p{pos.x + d * dir.x, pos.y + d * dir.y, pos.z + d * dir.z, pos.w + d * dir.w};
            V4<float> pp = mat4_mul<float>(m, p);
            float pre_depth = length(V3f(p.x, p.y, p.z) - V3f(prev_origin[0], prev_origin[1], prev_origin[2]));
            float pw = pp.w;
            pp = V4<float>{pp.x / pw, pp.y / pw, pp.z / pw, pp.w / pw};
            float ux = (pp.x + 1.0f) * 0.5f * ((float)w / ((float)w - 0.5f));
            float uy = (pp.y + 1.0f) * 0.5f * ((float)h / ((float)h - 0.5f));
            bool rep = false;
            float pc[4] = {0, 0, 0, 0};
            if (frame > 0) {
                float tpd[4];
                bilinear([&](int sx, int sy, float* o) { o[0] = prev_depth[(size_t)sy * w + sx]; }, iw, ih, ux, uy,
                         tpd, 1);
                float dis = tpd[0] / pre_depth - 1.0f;
                if (ux >= 0 && uy >= 0 && ux <= 1 && uy <= 1 && std::fabs(dis) <= 0.01f) {
                    rep = true;
                    bilinear(
                        [&](int sx, int sy, float* o) {
                            for (int k = 0; k < 4; ++k) o[k] = half_to_float(prev_acc[4 * ((size_t)sy * w + sx) + k]);
                        },
                        iw, ih, ux, uy, pc, 4);
                    float ps[4];
                    bilinear([&](int sx, int sy, float* o) { o[0] = (float)prev_spp[(size_t)sy * w + sx] / 255.0f; },
                             iw, ih, ux, uy, ps, 1);
                    pixel_spp += ps[0];
                }
            }
            motion[2 * pix] = float_to_half(rep ? ux : -1.0f);
            motion[2 * pix + 1] = float_to_half(rep ? uy : -1.0f);
            spp[pix] = unorm8(pixel_spp);
            V3f dm(illum[4 * pix], illum[4 * pix + 1], illum[4 * pix + 2]);
            if (rep) {
                float a = std::fmax(1.0f / (pixel_spp * 256.0f), 0.1f);
                dm = vmix(V3f(pc[0], pc[1], pc[2]), dm, a);
            }
            acc[4 * pix] = float_to_half(dm.x);
            acc[4 * pix + 1] = float_to_half(dm.y);
            acc[4 * pix + 2] = float_to_half(dm.z);
            acc[4 * pix + 3] = float_to_half(1.0f);
        }
}

static V3f ycocg(V3f c) {
    return V3f(dot(c, V3f(1.f, 2.f, 1.f)), dot(c, V3f(2.f, 0.f, -2.f)), dot(c, V3f(-1.f, 2.f, -1.f)));
}

static V3f bgra(const uint8_t* p) { return V3f((float)p[2] / 255.0f, (float)p[1] / 255.0f, (float)p[0] / 255.0f); }
static void put_bgra(uint8_t* p, V3f c) {
    p[0] = unorm8(c.z);
    p[1] = unorm8(c.y);
    p[2] = unorm8(c.x);
    p[3] = 255;
}

// taa.comp + Taa::copy_final_image (raw byte copy into the R8G8B8A8 history)
void orc_taa(uint32_t w, uint32_t h, uint32_t frame, const uint16_t* motion, const uint8_t* denoised,
             uint8_t* history, uint8_t* final_image) {
    const int iw = (int)w, ih = (int)h;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            size_t pix = (size_t)y * w + x;
            V3f cur = bgra(denoised + 4 * pix);
            float mx = half_to_float(motion[2 * pix]), my = half_to_float(motion[2 * pix + 1]);
            if (frame == 0 || mx < 0 || my < 0 || mx > 1 || my > 1) {
                put_bgra(final_image + 4 * pix, cur);
                continue;
            }
            const float big = 2147483648.0f;
            V3f mnb(big), mnc(big), mxb(-big), mxc(-big);
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int sx = (int)x + dx, sy = (int)y + dy;
                    if (sx < 0 || sy < 0 || sx >= iw || sy >= ih) continue;
                    V3f s = (dx == 0 && dy == 0) ? cur : bgra(denoised + 4 * ((size_t)sy * w + sx));
                    s = ycocg(s);
                    auto upd = [](V3f& lo, V3f& hi, V3f v) {
                        lo = V3f(std::fmin(lo.x, v.x), std::fmin(lo.y, v.y), std::fmin(lo.z, v.z));
                        hi = V3f(std::fmax(hi.x, v.x), std::fmax(hi.y, v.y), std::fmax(hi.z, v.z));
                    };
                    if (dx == 0 || dy == 0) upd(mnc, mxc, s);
                    upd(mnb, mxb, s);
                }
            float pv[4];
            bilinear(
                [&](int sx, int sy, float* o) {
                    const uint8_t* c = history + 4 * ((size_t)sy * w + sx);
                    for (int k = 0; k < 4; ++k) o[k] = (float)c[k] / 255.0f;
                },
                iw, ih, mx, my, pv, 4);
            V3f prev(pv[0], pv[1], pv[2]);
            V3f py = ycocg(prev);
            V3f lo = (mnb + mnc) * 0.5f, hi = (mxb + mxc) * 0.5f;
            bool inside = py.x >= lo.x && py.y >= lo.y && py.z >= lo.z && py.x <= hi.x && py.y <= hi.y && py.z <= hi.z;
            V3f res = inside ? 0.4f * cur + (1.0f - 0.4f) * prev : cur;
            put_bgra(final_image + 4 * pix, res);
        }
    std::memcpy(history, final_image, (size_t)w * h * 4);
}

// formatConverter.comp (rgba8 store into B8G8R8A8)
void orc_convert(uint32_t w, uint32_t h, const float* src, uint8_t* dst) {
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        dst[4 * i] = unorm8(src[4 * i + 2]);
        dst[4 * i + 1] = unorm8(src[4 * i + 1]);
        dst[4 * i + 2] = unorm8(src[4 * i]);
        dst[4 * i + 3] = unorm8(src[4 * i + 3]);
    }
}

// ---------------------------------------------------------------------------
// BMFR
// ---------------------------------------------------------------------------
static const float kOffsets[16][2] = {{.7f, .85f},  {.95f, .5f},  {.43f, .76f}, {.97f, .03f}, {.37f, .58f},
                                      {.03f, .36f}, {.81f, .46f}, {0.f, .78f},  {.36f, -.08f}, {-.06f, 0.f},
                                      {.95f, .1f},  {.85f, .61f}, {.06f, .1f},  {.43f, .16f}, {0.f, .5f},
                                      {.73f, .38f}};

static int mirror(int x, int s) {
    if (x < 0) return std::abs(x) - 1;
    if (x >= s) return 2 * s - x - 1;
    return x;
}

static float bmfr_random(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    uint32_t max_u = 0xffffffffu;
    return (float)a / (float)max_u;
}

struct BmfrGeo {
    int w, h, bs, wp, hp, bx, by, offx, offy;
    uint32_t frame;
};

static BmfrGeo geo(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame) {
    BmfrGeo g;
    g.w = (int)w;
    g.h = (int)h;
    g.bs = (int)bs;
    g.wp = (g.w / g.bs + 2) * g.bs;
    g.hp = (g.h / g.bs + 2) * g.bs;
    g.bx = g.wp / g.bs;
    g.by = g.hp / g.bs;
    g.offx = (int)((float)g.bs * kOffsets[frame % 16][0]);
    g.offy = (int)((float)g.bs * kOffsets[frame % 16][1]);
    g.frame = frame;
    return g;
}

struct PixIn {
    float depth;
    V3f n, noisy;
    int ix, iy;
    bool inside;
};

static PixIn pix_in(const BmfrGeo& g, int gx, int gy, const float* depth, const float* normal,
                    const uint16_t* noisy) {
    PixIn p;
    int ax = gx - g.offx, ay = gy - g.offy;
    p.ix = mirror(ax, g.w);
    p.iy = mirror(ay, g.h);
    p.inside = ax == p.ix && ay == p.iy;
    size_t pix = (size_t)p.iy * g.w + p.ix;
    p.depth = depth[pix];
    float th = normal[2 * pix], ph = normal[2 * pix + 1];
    p.n = V3f(std::cos(ph) * std::sin(th), std::sin(ph) * std::sin(th), std::cos(th));
    p.noisy = V3f(half_to_float(noisy[4 * pix]), half_to_float(noisy[4 * pix + 1]), half_to_float(noisy[4 * pix + 2]));
    return p;
}

// bmfrPre: fp16 feature buffer [13][hp][wp]
void orc_bmfr_pre(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame, const float* depth, const float* normal,
                  const uint16_t* noisy, uint16_t* features) {
    BmfrGeo g = geo(w, h, bs, frame);
    size_t plane = (size_t)g.wp * g.hp;
    std::vector<PixIn> px((size_t)g.bs * g.bs);
    for (int by = 0; by < g.by; ++by)
        for (int bx = 0; bx < g.bx; ++bx) {
            float mn = 3.4e38f, mx = -3.4e38f;
            for (int ly = 0; ly < g.bs; ++ly)
                for (int lx = 0; lx < g.bs; ++lx) {
                    PixIn& p = px[(size_t)ly * g.bs + lx];
                    p = pix_in(g, bx * g.bs + lx, by * g.bs + ly, depth, normal, noisy);
                    mn = std::fmin(mn, p.depth);
                    mx = std::fmax(mx, p.depth);
                }
            for (int ly = 0; ly < g.bs; ++ly)
                for (int lx = 0; lx < g.bs; ++lx) {
                    const PixIn& p = px[(size_t)ly * g.bs + lx];
                    float z = (p.depth - mn) / (mx - mn + 1e-6f);
                    float fx = (float)lx / (float)(g.bs - 1), fy = (float)ly / (float)(g.bs - 1);
                    float f[13] = {1.0f, p.n.x, p.n.y, p.n.z, fx, fy, z, fx * fx, fy * fy, z * z,
                                   p.noisy.x, p.noisy.y, p.noisy.z};
                    size_t at = (size_t)(by * g.bs + ly) * g.wp + (bx * g.bs + lx);
                    for (int i = 0; i < 13; ++i) features[i * plane + at] = float_to_half(f[i]);
                }
        }
}

// The noised fit matrix of block (bx, by): rows i (bmfrFit.comp:16-24), 13 columns.
static void fit_matrix(const BmfrGeo& g, const uint16_t* features, int bx, int by, std::vector<float>& A) {
    const int R = g.bs * g.bs;
    const uint32_t pb = (uint32_t)R;
    size_t plane = (size_t)g.wp * g.hp;
    A.assign((size_t)R * 13, 0.0f);
    for (int col = 0; col < 13; ++col)
        for (int i = 0; i < R; ++i) {
            int px = i / g.bs + bx * g.bs, py = i % g.bs + by * g.bs;
            float v = half_to_float(features[col * plane + (size_t)py * g.wp + px]);
            if (col < 10) {
                uint32_t key = (uint32_t)i + (uint32_t)col * pb * pb + g.frame * 13u * pb * pb;
                const float noise_amount = 1e-4f;
                v = v + (noise_amount * 2.f) * (bmfr_random(key) - .5f);
            }
            A[(size_t)i * 13 + col] = v;
        }
}

// Tree reduction of one logical subgroup of `sg` lanes (xor butterfly), then the
// first invocation folds the subgroup totals in order (bmfrGeneral.comp:59-71).
static float wg_sum(const std::vector<float>& v, int sg) {
    int n = (int)v.size();
    std::vector<float> totals;
    for (int base = 0; base < n; base += sg) {
        std::vector<float> lane(v.begin() + base, v.begin() + std::min(n, base + sg));
        lane.resize(sg, 0.0f);
        for (int off = sg / 2; off > 0; off >>= 1) {
            std::vector<float> nx(sg);
            for (int l = 0; l < sg; ++l) nx[l] = lane[l] + lane[l ^ off];
            lane.swap(nx);
        }
        totals.push_back(lane[0]);
    }
    float t = totals[0];
    for (size_t i = 1; i < totals.size(); ++i) t += totals[i];
    return t;
}

// bmfrFit (FP32 Householder, logical workgroup of T threads and subgroups of sg lanes)
// weights out: [blocks_y][blocks_x][30]
void orc_bmfr_fit(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame, int32_t subgroup, const uint16_t* features,
                  float* weights) {
    BmfrGeo g = geo(w, h, bs, frame);
    const int R = g.bs * g.bs;
    const int T = g.bs == 8 ? 64 : 256;
    const int S = R / T;
    std::vector<float> A;
    for (int by = 0; by < g.by; ++by)
        for (int bx = 0; bx < g.bx; ++bx) {
            fit_matrix(g, features, bx, by, A);
            // feat[thread][sub][col]
            std::vector<float> feat((size_t)T * S * 13);
            for (int id = 0; id < T; ++id)
                for (int s = 0; s < S; ++s)
                    for (int c = 0; c < 13; ++c) feat[((size_t)id * S + s) * 13 + c] = A[(size_t)(id + s * T) * 13 + c];
            auto F = [&](int id, int s, int c) -> float& { return feat[((size_t)id * S + s) * 13 + c]; };
            float uls = 0.0f;
            std::vector<float> u((size_t)T * S), part(T);
            for (int col = 0; col < 10; ++col) {
                for (int id = 0; id < T; ++id) {
                    float v2 = 0.0f;
                    for (int s = 0; s < S; ++s) {
                        int index = id + s * T;
                        u[(size_t)id * S + s] = F(id, s, col);
                        if (index > col) v2 += u[(size_t)id * S + s] * u[(size_t)id * S + s];
                    }
                    part[id] = v2;
                }
                float vls = wg_sum(part, subgroup);
                for (int id = 0; id < T; ++id) {
                    if (id < col) {
                        u[(size_t)id * S] = 0.0f;
                    } else if (id == col) {
                        float vl = std::sqrt(vls + u[(size_t)id * S] * u[(size_t)id * S]);
                        u[(size_t)id * S] -= vl;
                        float t = vls + u[(size_t)id * S] * u[(size_t)id * S];
                        uls = t;
                        F(id, 0, col) = vl;
                    } else {
                        F(id, 0, col) = 0.0f;
                    }
                }
                for (int f = col + 1; f < 13; ++f) {
                    for (int id = 0; id < T; ++id) {
                        float v = 0.0f;
                        for (int s = 0; s < S; ++s)
                            if (id + s * T >= col) v += F(id, s, f) * u[(size_t)id * S + s];
                        part[id] = v;
                    }
                    float vv = wg_sum(part, subgroup);
                    for (int id = 0; id < T; ++id)
                        for (int s = 0; s < S; ++s)
                            if (id + s * T >= col) F(id, s, f) -= 2 * u[(size_t)id * S + s] * vv / uls;
                }
            }
            V3f ws[10];
            for (int i = 9; i >= 0; --i) {
                V3f wv(F(i, 0, 10), F(i, 0, 11), F(i, 0, 12));
                for (int x = i + 1; x < 10; ++x) wv = wv - ws[x] * F(i, 0, x);
                wv = wv / F(i, 0, i);
                ws[i] = wv;
            }
            float* out = weights + ((size_t)by * g.bx + bx) * 30;
            for (int i = 0; i < 10; ++i) {
                V3f wv = uls == 0.0f ? V3f(0.2f) : ws[i];
                out[3 * i] = wv.x;
                out[3 * i + 1] = wv.y;
                out[3 * i + 2] = wv.z;
            }
        }
}

// FP64 least squares of every block (Householder QR in double on the same noised
// fp32 matrix).  weights out: [blocks_y][blocks_x][30] as double.
void orc_bmfr_lsq64(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame, const uint16_t* features,
                    double* weights) {
    BmfrGeo g = geo(w, h, bs, frame);
    const int R = g.bs * g.bs;
    std::vector<float> A;
    std::vector<double> M((size_t)R * 13);
    for (int by = 0; by < g.by; ++by)
        for (int bx = 0; bx < g.bx; ++bx) {
            fit_matrix(g, features, bx, by, A);
            for (size_t i = 0; i < M.size(); ++i) M[i] = A[i];
            for (int k = 0; k < 10; ++k) {
                double nrm = 0;
                for (int i = k; i < R; ++i) nrm += M[(size_t)i * 13 + k] * M[(size_t)i * 13 + k];
                nrm = std::sqrt(nrm);
                double akk = M[(size_t)k * 13 + k];
                double alpha = akk > 0 ? -nrm : nrm;
                std::vector<double> v(R, 0.0);
                for (int i = k; i < R; ++i) v[i] = M[(size_t)i * 13 + k];
                v[k] -= alpha;
                double vv = 0;
                for (int i = k; i < R; ++i) vv += v[i] * v[i];
                if (vv == 0) continue;
                for (int c = k; c < 13; ++c) {
                    double d = 0;
                    for (int i = k; i < R; ++i) d += v[i] * M[(size_t)i * 13 + c];
                    double f = 2 * d / vv;
                    for (int i = k; i < R; ++i) M[(size_t)i * 13 + c] -= f * v[i];
                }
            }
            double* out = weights + ((size_t)by * g.bx + bx) * 30;
            for (int c = 0; c < 3; ++c) {
                double x[10];
                for (int i = 9; i >= 0; --i) {
                    double s = M[(size_t)i * 13 + 10 + c];
                    for (int k = i + 1; k < 10; ++k) s -= M[(size_t)i * 13 + k] * x[k];
                    x[i] = s / M[(size_t)i * 13 + i];
                }
                for (int i = 0; i < 10; ++i) out[3 * i + c] = x[i];
            }
        }
}

// The noised fit matrix of one block (for textbook-routine checks): out R x 13 floats.
void orc_bmfr_block_matrix(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame, const uint16_t* features,
                           uint32_t bx, uint32_t by, float* out) {
    BmfrGeo g = geo(w, h, bs, frame);
    std::vector<float> A;
    fit_matrix(g, features, (int)bx, (int)by, A);
    std::memcpy(out, A.data(), A.size() * sizeof(float));
}

// bmfrPost.  weights: [blocks_y][blocks_x][30] floats; denoised: 2 layers RGBA16F.
void orc_bmfr_post(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame, const float* depth, const float* normal,
                   const uint8_t* albedo, const uint16_t* motion, const uint8_t* spp, const uint16_t* noisy,
                   const float* weights, uint16_t* denoised, uint8_t* final_image) {
    BmfrGeo g = geo(w, h, bs, frame);
    const size_t plane = (size_t)w * h;
    std::vector<PixIn> px((size_t)g.bs * g.bs);
    // all blocks read layer frame&1 before any write of layer (frame&1)^1: distinct layers
    for (int by = 0; by < g.by; ++by)
        for (int bx = 0; bx < g.bx; ++bx) {
            float mn = 3.4e38f, mx = -3.4e38f;
            for (int ly = 0; ly < g.bs; ++ly)
                for (int lx = 0; lx < g.bs; ++lx) {
                    PixIn& p = px[(size_t)ly * g.bs + lx];
                    p = pix_in(g, bx * g.bs + lx, by * g.bs + ly, depth, normal, noisy);
                    mn = std::fmin(mn, p.depth);
                    mx = std::fmax(mx, p.depth);
                }
            const float* wt = weights + ((size_t)by * g.bx + bx) * 30;
            float wc[30];
            for (int i = 0; i < 30; ++i) wc[i] = (std::isinf(wt[i]) || std::isnan(wt[i])) ? 0.0f : wt[i];
            for (int ly = 0; ly < g.bs; ++ly)
                for (int lx = 0; lx < g.bs; ++lx) {
                    const PixIn& p = px[(size_t)ly * g.bs + lx];
                    if (!p.inside) continue;
                    size_t pix = (size_t)p.iy * w + p.ix;
                    float mux = half_to_float(motion[2 * pix]), muy = half_to_float(motion[2 * pix + 1]);
                    bool accept = mux >= 0;
                    float pspp = ((float)spp[pix] / 255.0f) * 256.0f;
                    float z = (p.depth - mn) / (mx - mn + 1e-6f);
                    float fx = (float)lx / (float)(g.bs - 1), fy = (float)ly / (float)(g.bs - 1);
                    float f[10] = {1.0f, p.n.x, p.n.y, p.n.z, fx, fy, z, fx * fx, fy * fy, z * z};
                    V3f den(0.0f);
                    for (int i = 0; i < 10; ++i) den = den + V3f(wc[3 * i], wc[3 * i + 1], wc[3 * i + 2]) * f[i];
                    den = vclamp(den, 0.0f, 10.0f);
                    V3f prev(0.0f);
                    float alpha = 1.0f;
                    if (frame > 0 && accept) {
                        const uint16_t* layer = denoised + 4 * plane * (frame & 1u);
                        float pv[4];
                        bilinear(
                            [&](int sx, int sy, float* o) {
                                for (int k = 0; k < 4; ++k) o[k] = half_to_float(layer[4 * ((size_t)sy * w + sx) + k]);
                            },
                            (int)w, (int)h, mux, muy, pv, 4);
                        prev = V3f(pv[0], pv[1], pv[2]);
                        alpha = std::fmax(1.0f / pspp, 0.1f);
                    }
                    den = alpha * den + (1.0f - alpha) * prev;
                    uint16_t* out = denoised + 4 * plane * ((frame & 1u) ^ 1u) + 4 * pix;
                    out[0] = float_to_half(den.x);
                    out[1] = float_to_half(den.y);
                    out[2] = float_to_half(den.z);
                    out[3] = float_to_half(1.0f);
                    V3f alb((float)albedo[4 * pix] / 255.0f + 1e-6f, (float)albedo[4 * pix + 1] / 255.0f + 1e-6f,
                            (float)albedo[4 * pix + 2] / 255.0f + 1e-6f);
                    V3f t = alb * den;
                    t = V3f(std::fmax(t.x, 0.0f), std::fmax(t.y, 0.0f), std::fmax(t.z, 0.0f));
                    t = vclamp(vpow(t, 0.454545f), 0.0f, 1.0f);
                    uint8_t* o = final_image + 4 * pix;
                    o[0] = unorm8(t.z);
                    o[1] = unorm8(t.y);
                    o[2] = unorm8(t.x);
                    o[3] = 255;
                }
        }
}


// bfr.comp:202-308 with parallel_reduction_alpha (bfr.comp:100-142): per block, 7
// features, `iterations` Adam steps (40 in the reference) on the block-summed
// gradient (sums in double, in pixel order), reconstruction, temporal blend,
// remodulation.  The stop test reduces |gradient| over all 7 alpha components
// (only a NaN gradient stops early).  weights (optional): [blocks_y][blocks_x][21].
void orc_bfr(uint32_t w, uint32_t h, uint32_t bs, uint32_t frame, int iterations, const float* depth,
             const float* normal, const uint8_t* albedo, const uint16_t* motion, const uint8_t* spp,
             const uint16_t* noisy, uint16_t* denoised, uint8_t* final_image, float* weights) {
    static const int kOff[16][2] = {{-7, -11}, {-14, -8}, {-5, -12}, {-15, -1}, {-5, -9},  {-1, -4},
                                    {-14, -7}, {0, -13},  {-5, -1},  {-1, 0},   {-15, -2}, {-14, -10},
                                    {-1, -1},  {-6, -3},  {0, -8},   {-10, -4}};
    const float ALPHA = .863f, K = .116f, BETA1 = .3f, BETA2 = .7314f, EPS = 1e-8f;
    const float ALPHA_L = 1.1f, BETA1_L = .45f, BETA2_L = .75f;
    const int W = (int)w, H = (int)h, b = (int)bs, n = b * b;
    const int nbx = W / b + 2, nby = H / b + 2;
    const int offx = kOff[frame % 16][0], offy = kOff[frame % 16][1];
    const size_t plane = (size_t)W * H;
    struct Px {
        int ix, iy;
        bool inside;
        float f[7];
        V3f c;
        float spp;
        bool l1;
    };
    std::vector<Px> px(n);
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            float mn = 3.4e38f, mx = -3.4e38f;
            int l1 = 0;
            for (int r = 0; r < n; ++r) {
                Px& p = px[r];
                const int lx = r / b, ly = r % b;
                const int ax = bx * b + lx + offx, ay = by * b + ly + offy;
                p.ix = mirror(ax, W);
                p.iy = mirror(ay, H);
                p.inside = ax == p.ix && ay == p.iy;
                const size_t pix = (size_t)p.iy * W + p.ix;
                p.c = V3f(half_to_float(noisy[4 * pix]), half_to_float(noisy[4 * pix + 1]),
                          half_to_float(noisy[4 * pix + 2]));
                const float th = normal[2 * pix], ph = normal[2 * pix + 1];
                p.f[0] = 1.0f;
                p.f[1] = (float)lx / (float)(b - 1) - 0.5f;
                p.f[2] = (float)ly / (float)(b - 1) - 0.5f;
                p.f[3] = depth[pix];
                p.f[4] = std::cos(ph) * std::sin(th);
                p.f[5] = std::sin(ph) * std::sin(th);
                p.f[6] = std::cos(th);
                p.spp = ((float)spp[pix] / 255.0f) * 256.0f;
                p.l1 = p.spp >= 10.0f;
                l1 += p.l1 ? 1 : 0;
                mn = std::fmin(mn, p.f[3]);
                mx = std::fmax(mx, p.f[3]);
            }
            for (int r = 0; r < n; ++r) {
                float z = px[r].f[3] - mn;
                z /= mx - mn + EPS;
                px[r].f[3] = z * 2.0f - 1.0f;
            }
            const float l1r = (float)l1 * 1.0f / (float)n;
            const float am = l1r * ALPHA_L + (1.0f - l1r) * ALPHA;
            const float b1 = l1r * BETA1_L + (1.0f - l1r) * BETA1;
            const float b2 = l1r * BETA2_L + (1.0f - l1r) * BETA2;
            float alpha[21] = {0}, m[21] = {0}, v[21] = {0};
            for (int it = 0; it < iterations; ++it) {
                double g[21] = {0};
                for (int r = 0; r < n; ++r) {
                    const Px& p = px[r];
                    float pr[3] = {0.0f, 0.0f, 0.0f};
                    for (int j = 0; j < 7; ++j)
                        for (int c = 0; c < 3; ++c) pr[c] += p.f[j] * alpha[3 * j + c];
                    float d[3] = {p.c.x - pr[0], p.c.y - pr[1], p.c.z - pr[2]};
                    if (p.l1)
                        for (int c = 0; c < 3; ++c) d[c] = d[c] > 0.0f ? 1.0f : (d[c] < 0.0f ? -1.0f : 0.0f);
                    for (int j = 0; j < 7; ++j)
                        for (int c = 0; c < 3; ++c) g[3 * j + c] += (double)(p.f[j] * d[c]);
                }
                const float t = (float)(it + 1);
                const float step =
                    am * std::exp(-K * t) * std::sqrt(1.0f - std::pow(BETA2, t)) / (1.0f - std::pow(BETA1, t));
                float rest = 0.0f;
                for (int k = 0; k < 21; ++k) {
                    const float dl = (float)g[k];
                    m[k] = b1 * m[k] + (1.0f - BETA1) * dl;
                    v[k] = b2 * v[k] + (1.0f - BETA2) * (dl * dl);
                    alpha[k] += step * (m[k] / (std::sqrt(v[k]) + EPS));
                    rest += std::fabs(dl);
                }
                if (!(rest >= 0.0f)) break;
            }
            if (weights) std::memcpy(weights + ((size_t)by * nbx + bx) * 21, alpha, sizeof(alpha));
            for (int r = 0; r < n; ++r) {
                const Px& p = px[r];
                if (!p.inside) continue;
                const size_t pix = (size_t)p.iy * W + p.ix;
                V3f den(0.0f);
                for (int j = 0; j < 7; ++j) den = den + V3f(alpha[3 * j], alpha[3 * j + 1], alpha[3 * j + 2]) * p.f[j];
                den = vclamp(den, 0.0f, 10.0f);
                const float mux = half_to_float(motion[2 * pix]), muy = half_to_float(motion[2 * pix + 1]);
                V3f prev(0.0f);
                float blend = 1.0f;
                if (frame > 0 && mux >= 0.0f) {
                    const uint16_t* layer = denoised + 4 * plane * (frame & 1u);
                    float pv[4];
                    bilinear(
                        [&](int sx, int sy, float* o) {
                            for (int k = 0; k < 4; ++k) o[k] = half_to_float(layer[4 * ((size_t)sy * W + sx) + k]);
                        },
                        W, H, mux, muy, pv, 4);
                    prev = V3f(pv[0], pv[1], pv[2]);
                    blend = std::fmax(1.0f / p.spp, 0.1f);
                }
                den = blend * den + (1.0f - blend) * prev;
                uint16_t* out = denoised + 4 * plane * ((frame & 1u) ^ 1u) + 4 * pix;
                out[0] = float_to_half(den.x);
                out[1] = float_to_half(den.y);
                out[2] = float_to_half(den.z);
                out[3] = float_to_half(1.0f);
                V3f alb((float)albedo[4 * pix] / 255.0f + 1e-6f, (float)albedo[4 * pix + 1] / 255.0f + 1e-6f,
                        (float)albedo[4 * pix + 2] / 255.0f + 1e-6f);
                V3f t = alb * den;
                t = V3f(std::fmax(t.x, 0.0f), std::fmax(t.y, 0.0f), std::fmax(t.z, 0.0f));
                t = vclamp(vpow(t, 0.454545f), 0.0f, 1.0f);
                uint8_t* o = final_image + 4 * pix;
                o[0] = unorm8(t.z);
                o[1] = unorm8(t.y);
                o[2] = unorm8(t.x);
                o[3] = 255;
            }
        }
}

// bfrBlender.comp:22-69.  Window taps outside the image read as zero; a NaN
// deviation takes t = 1 (IEEE minNum in min(t, 1)).
static float mix3(float a, float b, float c, float t, float mid, float maxdev) {
    t /= maxdev;
    t = std::fmin(t, 1.0f);
    float fa = std::fmax(1.0f - (t / mid), 0.0f);
    float fb = (t < mid) ? t / mid : 1.0f - (t - mid) / (1.0f - mid);
    float fc = 1.0f - fa - fb;
    return fa * a + fb * b + fc * c;
}

void orc_bfr_blend(uint32_t w, uint32_t h, uint32_t radius, const uint16_t* average, const uint16_t* average_sq,
                   const uint8_t* den8, const uint8_t* den16, const uint8_t* den32, uint8_t* final_image) {
    const int W = (int)w, H = (int)h, R = (int)radius;
    auto mean3 = [](const uint16_t* img, size_t pix) {
        float s = half_to_float(img[4 * pix]) * (1.0f / 3.0f);
        s += half_to_float(img[4 * pix + 1]) * (1.0f / 3.0f);
        s += half_to_float(img[4 * pix + 2]) * (1.0f / 3.0f);
        return s;
    };
    auto mix = [](float x, float y, float a) { return x * (1.0f - a) + y * a; };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sq = 0.0f, av = 0.0f;
            int count = 0;
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    const int sx = x + dx, sy = y + dy;
                    float cur = 0.0f;
                    if (sx >= 0 && sy >= 0 && sx < W && sy < H) cur = mean3(average, (size_t)sy * W + sx);
                    ++count;
                    sq = mix(sq, cur * cur, 1.0f / (float)count);
                    av = mix(av, cur, 1.0f / (float)count);
                }
            const size_t pix = (size_t)y * W + x;
            av = mix(av, mean3(average, pix), 0.5f);
            sq = mix(sq, mean3(average_sq, pix), 0.5f);
            const float sd = std::sqrt(sq - av * av);
            uint8_t* o = final_image + 4 * pix;
            for (int c = 0; c < 3; ++c) {
                // B8G8R8A8: colour c (r, g, b) sits at byte 2 - c
                const int bi = 2 - c;
                float d2 = (float)den8[4 * pix + bi] / 255.0f, d1 = (float)den16[4 * pix + bi] / 255.0f,
                      d0 = (float)den32[4 * pix + bi] / 255.0f;
                o[bi] = unorm8(mix3(d0, d1, d2, sd, 0.5f, 1.0f));
            }
            o[3] = 255;
        }
}

}  // extern "C"
