// S6: perspective warp of A and B onto the union canvas with feather or
// Laplacian multiband blending, for any number of canvases per launch.
// Semantics: src/s6/s6.cu:60-433 - FP64 inverse map and bilinear sampling
// (FP64 -> FP32 / 255), border-distance weights, ownership A if s_a >= s_b,
// pull-push extension to 1x1, [1 4 6 4 1]/16 pyramids with reflect-101,
// band mix by blurred ownership, collapse, floor(clamp(255 x) + 0.5).
// The depth is generic: 3, 5, 6 or 7 reductions.
//
// MI355X layout: FP32 pyramid levels in HBM, one launch per level covering
// every canvas of the batch. Normalisation is fused into the push pass
// (top-down, so the coarse level is final when it is expanded), and band
// mixing is fused with the collapse - at level 0 also with the final
// quantisation - which removes three full-resolution passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "stitch/jobs.hpp"
#include "stitch/s6_jobs.hpp"

namespace stitch {
namespace {

using jobs::S6Job;
using jobs::S6Level;
using jobs::kS6MaxLevels;

constexpr int kBlock = 256;

// Reflect-101 index for the filter footprints here, which overshoot a
// border by at most 2: one fold is exact for n >= 3 (the reference's general
// modulo form, s6.cu:122-128, gives the same index).
__device__ __forceinline__ int reflect101(int x, int n) {
    if (n <= 2) return n == 1 ? 0 : (x & 1);
    return x < 0 ? -x : (x >= n ? 2 * n - 2 - x : x);
}

__device__ __forceinline__ float tap(int x) {
    const int t = x < 0 ? -x : x;
    return t == 0 ? 0.375f : (t == 1 ? 0.25f : 0.0625f);
}

__device__ bool center_coordinate(double& x, int n, double tol) {
    if (!isfinite(x)) return false;
    const double nearest = nearbyint(x);
    const double roundoff = 16 * 2.2204460492503131e-16 * fmax(1.0, fabs(x));
    if (fabs(x - nearest) <= fmax(tol, roundoff)) x = nearest;
    return x >= 0 && x <= n - 1.0;
}

__device__ bool sample(const std::uint8_t* img, int w, int h, int pitch, double x, double y, float* rgb, double& score,
                       double tol) {
    if (!center_coordinate(x, w, tol) || !center_coordinate(y, h, tol)) return false;
    const int x0 = static_cast<int>(floor(x)), y0 = static_cast<int>(floor(y));
    const int x1 = x0 < w - 1 ? x0 + 1 : x0, y1 = y0 < h - 1 ? y0 + 1 : y0;
    const double ax = x - x0, ay = y - y0;
    const std::uint8_t* r0 = img + static_cast<size_t>(y0) * pitch;
    const std::uint8_t* r1 = img + static_cast<size_t>(y1) * pitch;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double v0 = (1 - ax) * r0[3 * x0 + c] + ax * r0[3 * x1 + c];
        const double v1 = (1 - ax) * r1[3 * x0 + c] + ax * r1[3 * x1 + c];
        rgb[c] = static_cast<float>(((1 - ay) * v0 + ay * v1) / 255.0);
    }
    // min((x+.5)/w, (w-.5-x)/w) == min(x+.5, w-.5-x)/w exactly: a correctly
    // rounded division by a positive w is monotonic, so it commutes with min.
    score = fmin(fmin(x + 0.5, w - 0.5 - x) / w, fmin(y + 0.5, h - 0.5 - y) / h);
    return true;
}

__device__ __forceinline__ std::uint8_t quantize(float x) {
    return static_cast<std::uint8_t>(floorf(fminf(255.0f, fmaxf(0.0f, 255.0f * x)) + 0.5f));
}

template <class Begin>
__device__ int find_job_by(const dev::JobSpan<S6Job>& span, long long t, Begin begin) {
    if (!span.table) return 0;
    int lo = 0, hi = span.jobs() - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin(span.table[mid]) <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Row-segment tiles: tile = (row, 256-pixel segment); no per-thread division.
__device__ __forceinline__ unsigned tiles_x(int w) { return (static_cast<unsigned>(w) + kBlock - 1) / kBlock; }

__device__ __forceinline__ bool tile_pixel(long long local, int w, int& x, int& y) {
    const unsigned lt = static_cast<unsigned>(local);
    const unsigned tx = tiles_x(w);
    const unsigned row = lt / tx;
    x = static_cast<int>((lt - row * tx) * kBlock + threadIdx.x);
    y = static_cast<int>(row);
    return x < w;
}

constexpr unsigned kVA = 1u, kVB = 2u, kUA = 4u, kUB = 8u;

__device__ __forceinline__ void s6_clear_stats_job(const S6Job& j) {
    if (threadIdx.x == 0 && j.stats) {
        j.stats->union_pixels = 0;
        j.stats->overlap_pixels = 0;
    }
}

__global__ void s6_clear_stats(dev::JobSpan<S6Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s6_clear_stats_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

__device__ void flush_stats(const S6Job& j, unsigned long long u, unsigned long long o,
                            unsigned long long (&red)[2][kBlock / 64]) {
    u = dev::wave_sum(u);
    o = dev::wave_sum(o);
    const int lane = dev::lane_id(), wave = threadIdx.x / 64;
    if (lane == 0) {
        red[0][wave] = u;
        red[1][wave] = o;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long su = 0, so = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            su += red[0][w];
            so += red[1][w];
        }
        if (su) atomicAdd(reinterpret_cast<unsigned long long*>(&j.stats->union_pixels), su);
        if (so) atomicAdd(reinterpret_cast<unsigned long long*>(&j.stats->overlap_pixels), so);
    }
    __syncthreads();
}

// Canvas pixel -> A (integer shift) and B (FP64 inverse map), level-0 colours
// and flags or the feather output, union mask, union/overlap counters.
// A is sampled at integer positions, where the bilinear weights are exactly
// (1, 0): its colour is float(v / 255.0) of the source byte, read from a
// 256-entry table computed with the same expression.
__global__ void __launch_bounds__(kBlock) s6_warp(dev::JobSpan<S6Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ unsigned long long red[2][kBlock / 64];
    __shared__ float lut[256];
    lut[threadIdx.x] = static_cast<float>(static_cast<double>(threadIdx.x) / 255.0);
    __syncthreads();
    int cur = -1;
    unsigned long long uni = 0, ovl = 0;
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = find_job_by(span, t, [](const S6Job& j) { return j.warp_begin; });
        if (ji != cur) {
            if (cur >= 0 && dev::job_at(span, cur).stats) flush_stats(dev::job_at(span, cur), uni, ovl, red);
            cur = ji;
            uni = ovl = 0;
        }
        const S6Job& j = dev::job_at(span, ji);
        int x, y;
        if (!tile_pixel(t - j.warp_begin, j.width, x, y)) continue;
        const size_t p = static_cast<size_t>(y) * j.width + x;
        const int ax = x + j.origin_x, ay = y + j.origin_y;
        const double xa = static_cast<double>(x) + j.origin_x, ya = static_cast<double>(y) + j.origin_y;
        float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
        double sa = 0, sb = 0;
        const bool ma = ax >= 0 && ax <= j.wa - 1 && ay >= 0 && ay <= j.ha - 1;
        if (ma) {
            const std::uint8_t* q = j.a + static_cast<size_t>(ay) * j.pitch_a + 3 * ax;
            ra[0] = lut[q[0]];
            ra[1] = lut[q[1]];
            ra[2] = lut[q[2]];
            sa = fmin(fmin(xa + 0.5, j.wa - 0.5 - xa) / j.wa, fmin(ya + 0.5, j.ha - 0.5 - ya) / j.ha);
        }
        bool mb = false;
        const double* inv = j.inverse;
        const double den = inv[6] * xa + inv[7] * ya + inv[8];
        const double dsc = fabs(inv[6] * xa) + fabs(inv[7] * ya) + fabs(inv[8]);
        if (isfinite(den) && dsc > 0 && fabs(den) > j.denominator_tolerance * dsc) {
            const double xb = (inv[0] * xa + inv[1] * ya + inv[2]) / den;
            const double yb = (inv[3] * xa + inv[4] * ya + inv[5]) / den;
            mb = sample(j.b, j.wb, j.hb, j.pitch_b, xb, yb, rb, sb, j.snap_tolerance);
        }
        j.mask[static_cast<size_t>(y) * j.mask_pitch + x] = ma || mb ? 1 : 0;
        if (j.top >= 0) {
            const S6Level& L = j.lv[0];
            unsigned f = 0;
            if (ma) f |= kVA;
            if (mb) f |= kVB;
            if (ma && (!mb || sa >= sb)) f |= kUA;
            if (mb && (!ma || sb > sa)) f |= kUB;
            L.flags[p] = static_cast<std::uint8_t>(f);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                L.a[c * L.n + p] = ra[c];
                L.b[c * L.n + p] = rb[c];
            }
        } else {
            std::uint8_t* out = j.out + static_cast<size_t>(y) * j.out_pitch + 3 * x;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float color =
                    ma && mb ? static_cast<float>((sa * ra[c] + sb * rb[c]) / (sa + sb)) : (ma ? ra[c] : rb[c]);
                out[c] = quantize(color);
            }
        }
        uni += ma || mb;
        ovl += ma && mb;
    }
    if (cur >= 0 && dev::job_at(span, cur).stats) flush_stats(dev::job_at(span, cur), uni, ovl, red);
}

// ---------------------------------------------------------------- reduce
// 5x5 [1 4 6 4 1]/16 filter + 2x decimation (s6.cu:236-253) on LDS tiles:
// a 32x8 tile of level-l outputs reads a 67x19 region of level l-1, loaded
// once with coalesced rows and the reflect-101 borders applied at load time;
// each output keeps the reference's dy-major, dx-minor sum per channel.
// Tile rows hold the even columns first, then the odd ones, so the stride-2
// reads of neighbouring lanes land in consecutive banks; a row pitch of 80
// words puts the wave's two output rows 32 banks apart.
constexpr int kRW = 32, kRH = 8, kRIW = 2 * kRW + 3, kRIH = 2 * kRH + 3;
constexpr int kREven = (kRIW + 1) / 2, kRPitch = 80;
static_assert(kRIW <= kRPitch && (2 * kRPitch) % 64 == 32, "reduce tile pitch");

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// kind 0: pull (A, B, VA, VB); kind 1: Gaussians of filled A, B and of UA, UB.
__global__ void __launch_bounds__(kBlock) s6_reduce(dev::JobSpan<S6Job> span, dev::Total work, int l, int kind) {
    const long long total = work.get();
    __shared__ float tile[8][kRIH][kRPitch];
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = kind == 0 ? find_job_by(span, t, [l](const S6Job& j) { return j.rb[l]; })
                                 : find_job_by(span, t, [l](const S6Job& j) { return j.gb[l]; });
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& C = j.lv[l];
        const S6Level& F = j.lv[l - 1];
        const unsigned local = static_cast<unsigned>(t - (kind == 0 ? j.rb[l] : j.gb[l]));
        const unsigned tx = (static_cast<unsigned>(C.w) + kRW - 1) / kRW;
        const int x0 = static_cast<int>(local % tx) * kRW, y0 = static_cast<int>(local / tx) * kRH;
        const int ix0 = 2 * x0 - 2, iy0 = 2 * y0 - 2;
        const float* w1 = kind == 0 ? F.va : F.ua;
        const float* w2 = kind == 0 ? F.vb : F.ub;
        const unsigned m1 = kind == 0 ? kVA : kUA, m2 = kind == 0 ? kVB : kUB;
        for (int i = threadIdx.x; i < kRIW * kRIH; i += kBlock) {
            const int ix = i % kRIW, iy = i / kRIW;
            const int sx = clampi(reflect101(ix0 + ix, F.w), 0, F.w - 1);
            const int sy = clampi(reflect101(iy0 + iy, F.h), 0, F.h - 1);
            const size_t q = static_cast<size_t>(sy) * F.w + sx;
            const int cx = (ix & 1) * kREven + (ix >> 1);
            tile[0][iy][cx] = F.a[q];
            tile[1][iy][cx] = F.a[F.n + q];
            tile[2][iy][cx] = F.a[2 * F.n + q];
            tile[3][iy][cx] = F.b[q];
            tile[4][iy][cx] = F.b[F.n + q];
            tile[5][iy][cx] = F.b[2 * F.n + q];
            if (l == 1) {
                const unsigned f = F.flags[q];
                tile[6][iy][cx] = (f & m1) ? 1.0f : 0.0f;
                tile[7][iy][cx] = (f & m2) ? 1.0f : 0.0f;
            } else {
                tile[6][iy][cx] = w1[q];
                tile[7][iy][cx] = w2[q];
            }
        }
        __syncthreads();
        const int lx = threadIdx.x % kRW, ly = threadIdx.x / kRW;
        const int x = x0 + lx, y = y0 + ly;
        if (x < C.w && y < C.h) {
            float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                for (int dx = 0; dx < 5; ++dx) {
                    const float w = tap(dx - 2) * tap(dy - 2);
#pragma unroll
                    for (int c = 0; c < 8; ++c) acc[c] += w * tile[c][2 * ly + dy][(dx & 1) * kREven + lx + (dx >> 1)];
                }
            const size_t p = static_cast<size_t>(y) * C.w + x;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                C.a[c * C.n + p] = acc[c];
                C.b[c * C.n + p] = acc[3 + c];
            }
            if (kind == 0) {
                C.va[p] = acc[6];
                C.vb[p] = acc[7];
            } else {
                C.ua[p] = acc[6];
                C.ub[p] = acc[7];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- expand
// Upsampling reads a 3x3 coarse neighbourhood per fine pixel; a 64x4 fine
// tile needs a 34x4 coarse region, staged in LDS. Taps whose reflected fine
// position is odd carry weight 0 (adding 0 * v to the running sum leaves it
// unchanged), so all lanes run the same 25-step loop without divergence.
constexpr int kEW = 64, kEH = 4, kCW = kEW / 2 + 2, kCH = kEH / 2 + 2;

struct ExpandTaps {
    int xi[5], yi[5];      // LDS column/row per offset
    float wx[5], wy[5];    // tap weight or 0 when the position is odd
};

__device__ __forceinline__ void expand_setup(int x, int y, int cw, int ch, int cx0, int cy0, ExpandTaps& e) {
#pragma unroll
    for (int d = 0; d < 5; ++d) {
        const int rx = reflect101(x + d - 2, 2 * cw);
        const int ry = reflect101(y + d - 2, 2 * ch);
        e.xi[d] = (rx >> 1) - cx0;
        e.yi[d] = (ry >> 1) - cy0;
        e.wx[d] = (rx & 1) ? 0.0f : tap(d - 2);
        e.wy[d] = (ry & 1) ? 0.0f : tap(d - 2);
    }
}

// 4 * sum over dy, dx of (tap(dx) * tap(dy)) * coarse, for 3 interleaved channels.
__device__ __forceinline__ void expand_lds(const float (&src)[3][kCH][kCW], const ExpandTaps& e, float out[3]) {
    float s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const float w = e.wx[dx] * e.wy[dy];
            s0 += w * src[0][e.yi[dy]][e.xi[dx]];
            s1 += w * src[1][e.yi[dy]][e.xi[dx]];
            s2 += w * src[2][e.yi[dy]][e.xi[dx]];
        }
    out[0] = 4 * s0;
    out[1] = 4 * s1;
    out[2] = 4 * s2;
}

__device__ __forceinline__ void load_coarse(const float* src, int cw, int ch, std::size_t cn, int cx0, int cy0,
                                            float (&dst)[3][kCH][kCW]) {
    for (int i = threadIdx.x; i < kCW * kCH; i += kBlock) {
        const int cx = clampi(cx0 + i % kCW, 0, cw - 1), cy = clampi(cy0 + i / kCW, 0, ch - 1);
        const size_t q = static_cast<size_t>(cy) * cw + cx;
        dst[0][i / kCW][i % kCW] = src[q];
        dst[1][i / kCW][i % kCW] = src[cn + q];
        dst[2][i / kCW][i % kCW] = src[2 * cn + q];
    }
}

// normalize_color + push_missing for level l, top-down (s6.cu:254-270,
// :399-410). At level 0 the validity is exactly 0 or 1, so valid pixels keep
// their colour (x / 1 == x) and only missing ones are written.
__global__ void __launch_bounds__(kBlock) s6_push(dev::JobSpan<S6Job> span, dev::Total work, int l) {
    const long long total = work.get();
    __shared__ float ca[3][kCH][kCW], cb[3][kCH][kCW];
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = find_job_by(span, t, [l](const S6Job& j) { return j.pb[l]; });
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& L = j.lv[l];
        const unsigned local = static_cast<unsigned>(t - j.pb[l]);
        const unsigned tx = (static_cast<unsigned>(L.w) + kEW - 1) / kEW;
        const int x0 = static_cast<int>(local % tx) * kEW, y0 = static_cast<int>(local / tx) * kEH;
        const bool coarse = l < j.top;
        const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1;
        if (coarse) {
            const S6Level& C = j.lv[l + 1];
            load_coarse(C.a, C.w, C.h, C.n, cx0, cy0, ca);
            load_coarse(C.b, C.w, C.h, C.n, cx0, cy0, cb);
            __syncthreads();
        }
        const int x = x0 + static_cast<int>(threadIdx.x) % kEW, y = y0 + static_cast<int>(threadIdx.x) / kEW;
        if (x < L.w && y < L.h) {
            const size_t p = static_cast<size_t>(y) * L.w + x;
            float wa, wb;
            if (l == 0) {
                const unsigned f = L.flags[p];
                wa = (f & kVA) ? 1.0f : 0.0f;
                wb = (f & kVB) ? 1.0f : 0.0f;
            } else {
                wa = L.va[p];
                wb = L.vb[p];
            }
            ExpandTaps e;
            if (coarse && (!(wa > 0) || !(wb > 0))) expand_setup(x, y, j.lv[l + 1].w, j.lv[l + 1].h, cx0, cy0, e);
            float v[3];
            if (wa > 0) {
                if (l > 0)
#pragma unroll
                    for (int c = 0; c < 3; ++c) L.a[c * L.n + p] = L.a[c * L.n + p] / wa;
            } else {
                if (coarse) expand_lds(ca, e, v);
                else v[0] = v[1] = v[2] = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) L.a[c * L.n + p] = v[c];
            }
            if (wb > 0) {
                if (l > 0)
#pragma unroll
                    for (int c = 0; c < 3; ++c) L.b[c * L.n + p] = L.b[c * L.n + p] / wb;
            } else {
                if (coarse) expand_lds(cb, e, v);
                else v[0] = v[1] = v[2] = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) L.b[c * L.n + p] = v[c];
            }
        }
        if (coarse) __syncthreads();
    }
}

// mix_band + reconstruct (s6.cu:271-293) for level l, top-down; at level 0
// also finish (s6.cu:294-306), where single-source pixels keep their warped
// colour and only overlap pixels need the band mix.
__global__ void __launch_bounds__(kBlock) s6_mix(dev::JobSpan<S6Job> span, dev::Total work, int l) {
    const long long total = work.get();
    __shared__ float ca[3][kCH][kCW], cb[3][kCH][kCW], cz[3][kCH][kCW];
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = find_job_by(span, t, [l](const S6Job& j) { return j.mb[l]; });
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& L = j.lv[l];
        const unsigned local = static_cast<unsigned>(t - j.mb[l]);
        const unsigned tx = (static_cast<unsigned>(L.w) + kEW - 1) / kEW;
        const int x0 = static_cast<int>(local % tx) * kEW, y0 = static_cast<int>(local / tx) * kEH;
        const bool coarse = l < j.blend_top;
        const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1;
        if (coarse) {
            const S6Level& C = j.lv[l + 1];
            load_coarse(C.a, C.w, C.h, C.n, cx0, cy0, ca);
            load_coarse(C.b, C.w, C.h, C.n, cx0, cy0, cb);
            load_coarse(C.z, C.w, C.h, C.n, cx0, cy0, cz);
            __syncthreads();
        }
        const int x = x0 + static_cast<int>(threadIdx.x) % kEW, y = y0 + static_cast<int>(threadIdx.x) / kEW;
        if (x < L.w && y < L.h) {
            const size_t p = static_cast<size_t>(y) * L.w + x;
            float ua, ub;
            bool both = true;
            unsigned f = 0;
            if (l == 0) {
                f = L.flags[p];
                ua = (f & kUA) ? 1.0f : 0.0f;
                ub = (f & kUB) ? 1.0f : 0.0f;
                both = (f & kVA) && (f & kVB);
            } else {
                ua = L.ua[p];
                ub = L.ub[p];
            }
            if (l == 0 && !both) {
                std::uint8_t* out = j.out + static_cast<size_t>(y) * j.out_pitch + 3 * x;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    out[c] = quantize((f & kVA) ? L.a[c * L.n + p] : ((f & kVB) ? L.b[c * L.n + p] : 0));
            } else {
                float la[3], lb[3], zc[3] = {0, 0, 0};
                if (coarse) {
                    ExpandTaps e;
                    expand_setup(x, y, j.lv[l + 1].w, j.lv[l + 1].h, cx0, cy0, e);
                    float ea[3], eb[3];
                    expand_lds(ca, e, ea);
                    expand_lds(cb, e, eb);
                    expand_lds(cz, e, zc);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        la[c] = L.a[c * L.n + p] - ea[c];
                        lb[c] = L.b[c * L.n + p] - eb[c];
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        la[c] = L.a[c * L.n + p] - 0.0f;
                        lb[c] = L.b[c * L.n + p] - 0.0f;
                    }
                }
                const float tot = ua + ub;
                float z[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float m = tot > 0 ? (ua * la[c] + ub * lb[c]) / tot : 0;
                    z[c] = coarse ? m + zc[c] : m;
                }
                if (l > 0) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) L.z[c * L.n + p] = z[c];
                } else {
                    std::uint8_t* out = j.out + static_cast<size_t>(y) * j.out_pitch + 3 * x;
#pragma unroll
                    for (int c = 0; c < 3; ++c) out[c] = quantize(z[c]);
                }
            }
        }
        if (coarse) __syncthreads();
    }
}

bool valid_geometry(const s6::Geometry& g) {
    if (g.width <= 0 || g.height <= 0 ||
        static_cast<std::size_t>(g.width) > std::numeric_limits<std::size_t>::max() / static_cast<std::size_t>(g.height) ||
        g.width_a <= 0 || g.height_a <= 0 || g.width_b <= 0 || g.height_b <= 0 || g.origin_x > 0 || g.origin_y > 0)
        return false;
    const auto right = static_cast<long long>(g.origin_x) + g.width - 1;
    const auto bottom = static_cast<long long>(g.origin_y) + g.height - 1;
    if (right < g.width_a - 1LL || bottom < g.height_a - 1LL) return false;
    double magnitude = 0;
    for (double v : g.inverse) {
        if (!std::isfinite(v)) return false;
        magnitude = std::max(magnitude, std::fabs(v));
    }
    return magnitude > 0 && std::isfinite(g.denominator_tolerance) && g.denominator_tolerance > 0 &&
           g.denominator_tolerance < 1 && std::isfinite(g.snap_tolerance) && g.snap_tolerance >= 0 &&
           g.snap_tolerance < 0.5;
}

bool valid_mode(s6::BlendMode m) { return s6::requested_reductions(m) >= 0; }

struct Range {
    std::uintptr_t first, last;
};
bool image_range(const ImageView& v, PixelFormat format, Range& r) {
    const std::size_t channels = format == PixelFormat::RGB8 ? 3 : 1;
    if (!v.data || v.width <= 0 || v.height <= 0 || v.pitch_bytes <= 0 || v.format != format) return false;
    const std::size_t row = static_cast<std::size_t>(v.width) * channels;
    if (row > static_cast<std::size_t>(v.pitch_bytes)) return false;
    const auto max = std::numeric_limits<std::uintptr_t>::max();
    const auto rows = static_cast<std::size_t>(v.height - 1);
    const auto pitch = static_cast<std::size_t>(v.pitch_bytes);
    if (rows > (max - row) / pitch) return false;
    const auto span = rows * pitch + row;
    r.first = reinterpret_cast<std::uintptr_t>(v.data);
    if (r.first > max - span) return false;
    r.last = r.first + span;
    return true;
}
bool overlaps(const Range& a, const Range& b) { return a.first < b.last && b.first < a.last; }

std::size_t align256(std::size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

namespace jobs {

// Pyramid layout (s6.cu:60-88): level 0 is the canvas, each level halves with
// ceil down to 1x1; A, B (RGB FP32), VA, VB on every level; UA, UB, Z on the
// blended levels (Z only above level 0, which is fused into the output).
Status s6_layout(const s6::Geometry& g, s6::BlendMode mode, S6Layout* out) {
    if (!valid_geometry(g) || !valid_mode(mode)) return Status::INVALID_ARGUMENT;
    S6Layout L{};
    if (mode == s6::BlendMode::Feather) {
        *out = L;
        return Status::OK;
    }
    const int requested = s6::requested_reductions(mode);
    int w = g.width, h = g.height;
    std::size_t cur = 0;
    const auto add = [&](std::size_t count, int ch, std::size_t& off) {
        cur = align256(cur);
        off = cur;
        cur += count * sizeof(float) * ch;
    };
    for (int l = 0; l < kS6MaxLevels; ++l) {
        S6LevelOffsets& o = L.lv[l];
        o.w = w;
        o.h = h;
        const std::size_t n = static_cast<std::size_t>(w) * h;
        add(n, 3, o.a);
        add(n, 3, o.b);
        if (l == 0) {
            cur = align256(cur);
            o.flags = cur;
            cur += n;
        } else {
            add(n, 1, o.va);
            add(n, 1, o.vb);
            if (l <= requested) {
                add(n, 1, o.ua);
                add(n, 1, o.ub);
                add(n, 3, o.z);
            }
        }
        L.top = l;
        if (w == 1 && h == 1) break;
        if (l == kS6MaxLevels - 1) return Status::INVALID_ARGUMENT;
        w = w / 2 + w % 2;
        h = h / 2 + h % 2;
    }
    L.blend_top = std::min(requested, L.top);
    L.bytes = cur;
    *out = L;
    return Status::OK;
}

Status make_s6_job(const ImageView& a, const ImageView& b, const s6::Geometry& g, s6::BlendMode mode,
                   const ImageView& output, const ImageView& mask, void* ws, std::size_t available,
                   s6::RenderStatistics* stats, S6Job* job) {
    S6Layout layout{};
    const Status st = s6_layout(g, mode, &layout);
    if (st != Status::OK) return st;
    if (a.width != g.width_a || a.height != g.height_a || b.width != g.width_b || b.height != g.height_b ||
        output.width != g.width || output.height != g.height || mask.width != g.width || mask.height != g.height)
        return Status::INVALID_ARGUMENT;
    Range r[4];
    if (!image_range(a, PixelFormat::RGB8, r[0]) || !image_range(b, PixelFormat::RGB8, r[1]) ||
        !image_range(output, PixelFormat::RGB8, r[2]) || !image_range(mask, PixelFormat::GRAY8, r[3]))
        return Status::INVALID_ARGUMENT;
    for (int i = 0; i < 4; ++i)
        for (int k = i + 1; k < 4; ++k)
            if (!(i == 0 && k == 1) && overlaps(r[i], r[k])) return Status::INVALID_ARGUMENT;
    if (available < layout.bytes || (layout.bytes && !ws)) return Status::INSUFFICIENT_WORKSPACE;
    if (layout.bytes) {
        const auto address = reinterpret_cast<std::uintptr_t>(ws);
        if (address % 256 || address > std::numeric_limits<std::uintptr_t>::max() - available)
            return Status::INVALID_ARGUMENT;
        const Range arena{address, address + available};
        for (const auto& x : r)
            if (overlaps(x, arena)) return Status::INVALID_ARGUMENT;
    }
    if (stats) {
        const auto address = reinterpret_cast<std::uintptr_t>(stats);
        if (address % alignof(s6::RenderStatistics)) return Status::INVALID_ARGUMENT;
        const Range counters{address, address + sizeof(s6::RenderStatistics)};
        for (const auto& x : r)
            if (overlaps(x, counters)) return Status::INVALID_ARGUMENT;
        if (ws && available && overlaps(counters, Range{reinterpret_cast<std::uintptr_t>(ws),
                                                        reinterpret_cast<std::uintptr_t>(ws) + available}))
            return Status::INVALID_ARGUMENT;
    }
    S6Job j{};
    j.a = a.data;
    j.pitch_a = a.pitch_bytes;
    j.wa = a.width;
    j.ha = a.height;
    j.b = b.data;
    j.pitch_b = b.pitch_bytes;
    j.wb = b.width;
    j.hb = b.height;
    j.width = g.width;
    j.height = g.height;
    j.origin_x = g.origin_x;
    j.origin_y = g.origin_y;
    j.denominator_tolerance = g.denominator_tolerance;
    j.snap_tolerance = g.snap_tolerance;
    std::memcpy(j.inverse, g.inverse, sizeof j.inverse);
    j.out = output.data;
    j.out_pitch = output.pitch_bytes;
    j.mask = mask.data;
    j.mask_pitch = mask.pitch_bytes;
    j.stats = stats;
    j.mode = static_cast<int>(mode);
    if (mode == s6::BlendMode::Feather) {
        j.top = -1;
        j.blend_top = -1;
    } else {
        j.top = layout.top;
        j.blend_top = layout.blend_top;
        auto* base = static_cast<std::uint8_t*>(ws);
        const auto f = [&](std::size_t off) { return reinterpret_cast<float*>(base + off); };
        for (int l = 0; l <= layout.top; ++l) {
            const S6LevelOffsets& o = layout.lv[l];
            S6Level& L = j.lv[l];
            L.w = o.w;
            L.h = o.h;
            L.n = static_cast<std::size_t>(o.w) * o.h;
            L.a = f(o.a);
            L.b = f(o.b);
            if (l == 0) {
                L.flags = base + o.flags;
            } else {
                L.va = f(o.va);
                L.vb = f(o.vb);
                if (l <= layout.blend_top) {
                    L.ua = f(o.ua);
                    L.ub = f(o.ub);
                    L.z = f(o.z);
                }
            }
        }
    }
    *job = j;
    return Status::OK;
}

S6Totals s6_prefix(S6Job* js, int n) {
    S6Totals t{};
    t.top = -1;
    t.blend_top = -1;
    for (int i = 0; i < n; ++i) {
        S6Job& j = js[i];
        j.warp_begin = t.warp;
        t.warp += static_cast<long long>((j.width + kBlock - 1) / kBlock) * j.height;
        for (int l = 0; l < kS6MaxLevels; ++l) {
            j.rb[l] = t.pull[l];
            j.gb[l] = t.gauss[l];
            j.pb[l] = t.push[l];
            j.mb[l] = t.mix[l];
            if (j.top < l) continue;
            const long long w = j.lv[l].w, h = j.lv[l].h;
            const long long rtiles = ((w + kRW - 1) / kRW) * ((h + kRH - 1) / kRH);
            const long long etiles = ((w + kEW - 1) / kEW) * ((h + kEH - 1) / kEH);
            if (l >= 1) t.pull[l] += rtiles;
            t.push[l] += etiles;
            if (j.blend_top >= l) {
                if (l >= 1) t.gauss[l] += rtiles;
                t.mix[l] += etiles;
            }
        }
        t.top = std::max(t.top, j.top);
        t.blend_top = std::max(t.blend_top, j.blend_top);
    }
    return t;
}

hipError_t launch_s6(const S6Job* table, int count, const S6Totals& t, const S6Job* single, hipStream_t stream,
                     const dev::Replay<S6Totals>* replay) {
    dev::JobSpan<S6Job> span{table, count, single ? *single : S6Job{}, replay ? replay->count : nullptr};
    const S6Totals* d = replay ? replay->totals : nullptr;
    const bool any_stats = single ? single->stats != nullptr : true;
    if (any_stats) s6_clear_stats<<<count, 64, 0, stream>>>(span);
    if (t.warp > 0) s6_warp<<<persistent_grid(t.warp), kBlock, 0, stream>>>(span, dev::Total(t.warp, d ? &d->warp : nullptr));
    for (int l = 1; l <= t.top; ++l)
        if (t.pull[l] > 0)
            s6_reduce<<<persistent_grid(t.pull[l]), kBlock, 0, stream>>>(
                span, dev::Total(t.pull[l], d ? &d->pull[l] : nullptr), l, 0);
    for (int l = t.top; l >= 0; --l)
        if (t.push[l] > 0)
            s6_push<<<persistent_grid(t.push[l]), kBlock, 0, stream>>>(span, dev::Total(t.push[l], d ? &d->push[l] : nullptr), l);
    for (int l = 1; l <= t.blend_top; ++l)
        if (t.gauss[l] > 0)
            s6_reduce<<<persistent_grid(t.gauss[l]), kBlock, 0, stream>>>(
                span, dev::Total(t.gauss[l], d ? &d->gauss[l] : nullptr), l, 1);
    for (int l = t.blend_top; l >= 0; --l)
        if (t.mix[l] > 0)
            s6_mix<<<persistent_grid(t.mix[l]), kBlock, 0, stream>>>(span, dev::Total(t.mix[l], d ? &d->mix[l] : nullptr), l);
    return hipGetLastError();
}

}  // namespace jobs

namespace s6 {

int effective_reductions(const Geometry& g, BlendMode mode) {
    if (!valid_mode(mode) || !valid_geometry(g)) return -1;
    if (mode == BlendMode::Feather) return 0;
    int r = 0, w = g.width, h = g.height;
    const int requested = requested_reductions(mode);
    while ((w > 1 || h > 1) && r < requested) {
        w = w / 2 + w % 2;
        h = h / 2 + h % 2;
        ++r;
    }
    return r;
}

Status s6_workspace_bytes(const Geometry& g, BlendMode mode, std::size_t* bytes) {
    if (!bytes) return Status::INVALID_ARGUMENT;
    *bytes = 0;
    jobs::S6Layout layout{};
    const Status st = jobs::s6_layout(g, mode, &layout);
    if (st == Status::OK) *bytes = layout.bytes;
    return st;
}

Status render(const ImageView& a, const ImageView& b, const Geometry& g, BlendMode mode, const ImageView& output,
              const ImageView& mask, void* workspace, std::size_t available, hipStream_t stream,
              RenderStatistics* statistics) {
    jobs::S6Job j{};
    const Status st = jobs::make_s6_job(a, b, g, mode, output, mask, workspace, available, statistics, &j);
    if (st != Status::OK) return st;
    const jobs::S6Totals t = jobs::s6_prefix(&j, 1);
    return jobs::launch_s6(nullptr, 1, t, &j, stream) == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

}  // namespace s6
}  // namespace stitch
