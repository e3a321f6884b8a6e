// S6 over a chain of frames onto one canvas (no counterpart in the
// reference, which renders pairs; the blend generalises the pair render of
// src/s6/s6.cu:60-433 to n frames, and equals it for two).
//
// Frame 0 sits at an integer offset, frames 1.. are sampled through their
// inverse maps (FP64 positions and bilinear weights, as the pair render's
// reference semantics). Ownership goes to the valid frame with the largest
// border score (lowest index on ties); feather averages the valid frames by
// score; multiband runs every frame's pull-push extension, Gaussian and
// ownership pyramids and mixes the bands over all frames.
//
// Layout: a frame only matters near its own footprint, so every per-frame
// plane covers a window of its level instead of the whole canvas:
//   level 0   colour, validity and border score over the frame's canvas box;
//   pull      raw sums and weights over the box of level l plus 3 pixels
//             (outside it they are exactly zero, and readers take them so);
//   filled    the pull-push colour over the windows the finer levels read;
//   Gaussian  colour and ownership pyramids (levels 1..blend top) over
//             windows that shrink level by level as the reductions read them.
// The canvas keeps the owner and valid-count maps and one mix plane per
// blended level. Windows are sized so that every value that reaches the
// output is read inside its window; reads are clamped to the window in any
// case, so no access leaves its allocation. One launch per frame and level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "stitch/stages.hpp"

namespace stitch::s6 {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxLevels = 40;
// Window margins in level pixels. A frame's ownership weight at level l is
// zero outside its level-0 box scaled by 2^-l and widened by 2 (each reduce
// widens the support by under one coarse pixel, and the widenings sum to
// less than 2); the mix zone keeps one more. The Gaussian levels are read
// over the zone, through the expansion by the mix one level finer (4) and by
// the level-0 finish (2).
constexpr int kZMargin = 3;
constexpr int kGTop = 4;

struct Box {  // half-open [x0, x1) x [y0, y1) in level pixels
    int x0, y0, x1, y1;
    __host__ __device__ int w() const { return x1 - x0; }
    __host__ __device__ int h() const { return y1 - y0; }
    __host__ __device__ long long n() const { return static_cast<long long>(w()) * h(); }
    __host__ __device__ bool has(int x, int y) const { return x >= x0 && x < x1 && y >= y0 && y < y1; }
};

// Windowed plane of `ch` floats per pixel, pixel-major.
struct Plane {
    float* p;
    Box b;
    int ch;
    __device__ float at(int x, int y, int c) const {
        x = min(max(x, b.x0), b.x1 - 1);
        y = min(max(y, b.y0), b.y1 - 1);
        return p[(static_cast<size_t>(y - b.y0) * b.w() + (x - b.x0)) * ch + c];
    }
    __device__ float* ptr(int x, int y) const {
        return p + (static_cast<size_t>(y - b.y0) * b.w() + (x - b.x0)) * ch;
    }
};

struct Level0 {
    const std::uint8_t* img;
    int pitch, w, h;       // source frame
    double inv[9];         // canvas A-coordinates -> frame (frame 0: unused)
    double den_tol, snap_tol;
    int shift;             // frame 0: sampled at its integer offset
    Box box;               // canvas pixels the frame can cover
    float* color;          // box, 3 per pixel (0 where invalid)
    std::uint8_t* valid;   // box
    double* score;         // box
};

__device__ __forceinline__ int reflect101(int x, int n) {
    if (n <= 2) return n == 1 ? 0 : (x & 1);
    return x < 0 ? -x : (x >= n ? 2 * n - 2 - x : x);
}

__device__ __forceinline__ float tap(int d) {
    const int t = d < 0 ? -d : d;
    return t == 0 ? 0.375f : (t == 1 ? 0.25f : 0.0625f);
}

__device__ bool centered(double& x, int n, double tol) {
    if (!isfinite(x)) return false;
    const double nearest = nearbyint(x);
    const double roundoff = 16 * 2.2204460492503131e-16 * fmax(1.0, fabs(x));
    if (fabs(x - nearest) <= fmax(tol, roundoff)) x = nearest;
    return x >= 0 && x <= n - 1.0;
}

// Bilinear sample of an RGB8 frame in FP64 (weights and blend), /255 to FP32,
// and the border score of the snapped position.
__device__ bool sample(const std::uint8_t* img, int w, int h, int pitch, double x, double y, float* rgb,
                       double* score, double tol) {
    if (!centered(x, w, tol) || !centered(y, h, tol)) return false;
    const int x0 = static_cast<int>(floor(x)), y0 = static_cast<int>(floor(y));
    const int x1 = x0 < w - 1 ? x0 + 1 : x0, y1 = y0 < h - 1 ? y0 + 1 : y0;
    const double ax = x - x0, ay = y - y0;
    const std::uint8_t* r0 = img + static_cast<size_t>(y0) * pitch;
    const std::uint8_t* r1 = img + static_cast<size_t>(y1) * pitch;
    for (int c = 0; c < 3; ++c) {
        const double v0 = (1 - ax) * r0[3 * x0 + c] + ax * r0[3 * x1 + c];
        const double v1 = (1 - ax) * r1[3 * x0 + c] + ax * r1[3 * x1 + c];
        rgb[c] = static_cast<float>(((1 - ay) * v0 + ay * v1) / 255.0);
    }
    *score = fmin(fmin((x + 0.5) / w, (w - 0.5 - x) / w), fmin((y + 0.5) / h, (h - 0.5 - y) / h));
    return true;
}

__device__ __forceinline__ std::uint8_t quantize(float x) {
    return static_cast<std::uint8_t>(floorf(fminf(255.0f, fmaxf(0.0f, 255.0f * x)) + 0.5f));
}

#define GRID_STRIDE(i, n) \
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
         i += static_cast<long long>(gridDim.x) * blockDim.x)

// Level 0 of one frame over its box.
__global__ void __launch_bounds__(kThreads) chain_warp(Level0 f, int origin_x, int origin_y) {
    GRID_STRIDE(i, f.box.n()) {
        const int x = f.box.x0 + static_cast<int>(i % f.box.w()), y = f.box.y0 + static_cast<int>(i / f.box.w());
        const double xa = static_cast<double>(x) + origin_x, ya = static_cast<double>(y) + origin_y;
        float rgb[3] = {0, 0, 0};
        double s = 0;
        bool m = false;
        if (f.shift) {
            m = sample(f.img, f.w, f.h, f.pitch, xa, ya, rgb, &s, f.snap_tol);
        } else {
            const double den = f.inv[6] * xa + f.inv[7] * ya + f.inv[8];
            const double dsc = fabs(f.inv[6] * xa) + fabs(f.inv[7] * ya) + fabs(f.inv[8]);
            if (isfinite(den) && dsc > 0 && fabs(den) > f.den_tol * dsc) {
                const double xb = (f.inv[0] * xa + f.inv[1] * ya + f.inv[2]) / den;
                const double yb = (f.inv[3] * xa + f.inv[4] * ya + f.inv[5]) / den;
                m = sample(f.img, f.w, f.h, f.pitch, xb, yb, rgb, &s, f.snap_tol);
            }
        }
        for (int c = 0; c < 3; ++c) f.color[i * 3 + c] = m ? rgb[c] : 0.0f;
        f.valid[i] = m ? 1 : 0;
        f.score[i] = m ? s : 0.0;
    }
}

struct Frames {
    Level0 f[kMaxChainFrames];
    int n;
};

// Owner and valid count per canvas pixel; the mask, the statistics, and the
// output of every pixel the band mix does not decide (feather: all of them).
__global__ void __launch_bounds__(kThreads) chain_own(const Frames* fr, int W, int H, bool feather, std::int8_t* owner,
                                                     std::uint8_t* count, std::uint8_t* out, int out_pitch,
                                                     std::uint8_t* mask, int mask_pitch,
                                                     unsigned long long* stats) {
    unsigned long long uni = 0, ovl = 0;
    GRID_STRIDE(i, static_cast<long long>(W) * H) {
        const int x = static_cast<int>(i % W), y = static_cast<int>(i / W);
        int own = -1, valid = 0;
        double best = 0, fs = 0, fc[3] = {0, 0, 0};
        float single[3] = {0, 0, 0};
        for (int k = 0; k < fr->n; ++k) {
            const Level0& f = fr->f[k];
            if (!f.box.has(x, y)) continue;
            const size_t j = static_cast<size_t>(y - f.box.y0) * f.box.w() + (x - f.box.x0);
            if (!f.valid[j]) continue;
            const double s = f.score[j];
            ++valid;
            if (own < 0 || s > best) {
                own = k;
                best = s;
            }
            for (int c = 0; c < 3; ++c) {
                const float v = f.color[j * 3 + c];
                fc[c] = valid == 1 ? s * v : fc[c] + s * v;
                single[c] = v;
            }
            fs = valid == 1 ? s : fs + s;
        }
        owner[i] = static_cast<std::int8_t>(own);
        count[i] = static_cast<std::uint8_t>(min(valid, 255));
        mask[static_cast<size_t>(y) * mask_pitch + x] = valid ? 1 : 0;
        uni += valid > 0;
        ovl += valid > 1;
        std::uint8_t* o = out + static_cast<size_t>(y) * out_pitch + 3 * x;
        if (feather || valid < 2)
            for (int c = 0; c < 3; ++c)
                o[c] = quantize(valid > 1 ? static_cast<float>(fc[c] / fs) : single[c]);
    }
    if (stats) {
        // One atomic pair per block (the grid is capped at a few blocks per CU).
        __shared__ unsigned long long part[2][kThreads / 64];
        for (int off = 32; off > 0; off >>= 1) {
            uni += __shfl_xor(uni, off, 64);
            ovl += __shfl_xor(ovl, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            part[0][threadIdx.x / 64] = uni;
            part[1][threadIdx.x / 64] = ovl;
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            unsigned long long v = 0;
            for (int w = 0; w < kThreads / 64; ++w) v += part[threadIdx.x][w];
            if (v) atomicAdd(&stats[threadIdx.x], v);
        }
    }
}

__global__ void chain_clear_stats(unsigned long long* stats) {
    if (threadIdx.x < 2) stats[threadIdx.x] = 0;
}

// Pull into level 1: 5x5 reduce of colour x validity and validity from the
// frame's level 0 (zero outside its box; zero taps are skipped, which leaves
// every sum as the full filter gives it).
__global__ void __launch_bounds__(kThreads) chain_pull1(Level0 f, int W0, int H0, Plane dst) {
    GRID_STRIDE(i, dst.b.n()) {
        const int x = dst.b.x0 + static_cast<int>(i % dst.b.w()), y = dst.b.y0 + static_cast<int>(i / dst.b.w());
        float v[4] = {0, 0, 0, 0};
        for (int dy = -2; dy <= 2; ++dy) {
            const int yy = reflect101(2 * y + dy, H0);
            for (int dx = -2; dx <= 2; ++dx) {
                const int xx = reflect101(2 * x + dx, W0);
                if (!f.box.has(xx, yy)) continue;
                const size_t j = static_cast<size_t>(yy - f.box.y0) * f.box.w() + (xx - f.box.x0);
                const float t = tap(dx) * tap(dy);
                const float ok = f.valid[j] ? 1.0f : 0.0f;
                for (int c = 0; c < 3; ++c) v[c] = fmaf(t, f.color[j * 3 + c], v[c]);
                v[3] = fmaf(t, ok, v[3]);
            }
        }
        float* d = dst.ptr(x, y);
        for (int c = 0; c < 4; ++c) d[c] = v[c];
    }
}

// Pull into level l >= 2 from level l - 1 (zero outside the source window).
__global__ void __launch_bounds__(kThreads) chain_pull(Plane src, int fw, int fh, Plane dst) {
    GRID_STRIDE(i, dst.b.n()) {
        const int x = dst.b.x0 + static_cast<int>(i % dst.b.w()), y = dst.b.y0 + static_cast<int>(i / dst.b.w());
        float v[4] = {0, 0, 0, 0};
        for (int dy = -2; dy <= 2; ++dy) {
            const int yy = reflect101(2 * y + dy, fh);
            for (int dx = -2; dx <= 2; ++dx) {
                const int xx = reflect101(2 * x + dx, fw);
                if (!src.b.has(xx, yy)) continue;
                const float t = tap(dx) * tap(dy);
                for (int c = 0; c < 4; ++c) v[c] = fmaf(t, src.at(xx, yy, c), v[c]);
            }
        }
        float* d = dst.ptr(x, y);
        for (int c = 0; c < 4; ++c) d[c] = v[c];
    }
}

// Expansion of a coarse plane (cw x ch level) at fine pixel (x, y): the taps
// whose reflected fine position is even, times 4.
__device__ __forceinline__ float expand_at(const Plane& src, int cw, int ch, int x, int y, int c) {
    float sum = 0;
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(y + dy, 2 * ch);
        if (yy & 1) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101(x + dx, 2 * cw);
            if (xx & 1) continue;
            sum = fmaf(tap(dx) * tap(dy), src.at(xx / 2, yy / 2, c), sum);
        }
    }
    return 4 * sum;
}

// Pull value (colour sum c or weight 3) of a windowed pull plane: zero outside.
__device__ __forceinline__ float pulled(const Plane& p, int x, int y, int c) {
    return p.b.has(x, y) ? p.at(x, y, c) : 0.0f;
}

// Normalised pull where the level has weight, else the expansion of the
// filled level above (none at the top level: zero).
__global__ void __launch_bounds__(kThreads) chain_push(Plane pull, Plane dst, Plane coarse, int cw, int ch,
                                                      bool top) {
    GRID_STRIDE(i, dst.b.n()) {
        const int x = dst.b.x0 + static_cast<int>(i % dst.b.w()), y = dst.b.y0 + static_cast<int>(i / dst.b.w());
        const float wgt = pulled(pull, x, y, 3);
        float* d = dst.ptr(x, y);
        for (int c = 0; c < 3; ++c) {
            if (wgt > 0) d[c] = pulled(pull, x, y, c) / wgt;
            else d[c] = top ? 0.0f : expand_at(coarse, cw, ch, x, y, c);
        }
    }
}

// Filled level-0 colour of a frame (3 channels): its own colour where valid,
// else the expansion of its filled level 1 (each channel summed in the
// expansion's tap order).
__device__ __forceinline__ void filled0(const Level0& f, const Plane& f1, int w1, int h1, int x, int y, float* v) {
    if (f.box.has(x, y)) {
        const size_t j = static_cast<size_t>(y - f.box.y0) * f.box.w() + (x - f.box.x0);
        if (f.valid[j]) {
            for (int c = 0; c < 3; ++c) v[c] = f.color[j * 3 + c];
            return;
        }
    }
    float sum[3] = {0, 0, 0};
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(y + dy, 2 * h1);
        if (yy & 1) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101(x + dx, 2 * w1);
            if (xx & 1) continue;
            const float t = tap(dx) * tap(dy);
            for (int c = 0; c < 3; ++c) sum[c] = fmaf(t, f1.at(xx / 2, yy / 2, c), sum[c]);
        }
    }
    for (int c = 0; c < 3; ++c) v[c] = 4 * sum[c];
}

// Gaussian level 1 of the filled frame and its ownership weight. A block
// makes a kG1W x kG1H tile of level-1 pixels from the (2 kG1W + 3) x
// (2 kG1H + 3) level-0 pixels the 5x5 reduce reads, staged once in LDS
// (filled colour, and ownership as 0/1), so each filled value is formed once
// per tile instead of once per tap. Input positions are reflected into the
// canvas as they are staged; positions past what the tile's real outputs
// read are clamped (their values are never used).
constexpr int kG1W = 64, kG1H = 8;
constexpr int kG1InW = 2 * kG1W + 3, kG1InH = 2 * kG1H + 3;

__global__ void __launch_bounds__(kThreads) chain_gauss1(Level0 f, int k, const std::int8_t* owner, int W0, int H0,
                                                        Plane f1, int w1, int h1, Plane g, Plane u) {
    __shared__ float s_c[3][kG1InH][kG1InW];
    __shared__ float s_o[kG1InH][kG1InW];
    const int tiles_x = (g.b.w() + kG1W - 1) / kG1W;
    const int ox0 = g.b.x0 + static_cast<int>(blockIdx.x % tiles_x) * kG1W;
    const int oy0 = g.b.y0 + static_cast<int>(blockIdx.x / tiles_x) * kG1H;
    const int ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
    for (int i = threadIdx.x; i < kG1InW * kG1InH; i += blockDim.x) {
        const int lx = i % kG1InW, ly = i / kG1InW;
        const int xx = reflect101(min(ix0 + lx, W0 + 1), W0), yy = reflect101(min(iy0 + ly, H0 + 1), H0);
        float v[3];
        filled0(f, f1, w1, h1, xx, yy, v);
        for (int c = 0; c < 3; ++c) s_c[c][ly][lx] = v[c];
        s_o[ly][lx] = owner[static_cast<size_t>(yy) * W0 + xx] == k ? 1.0f : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kG1W * kG1H; i += blockDim.x) {
        const int lx = i % kG1W, ly = i / kG1W;
        const int x = ox0 + lx, y = oy0 + ly;
        if (x >= g.b.x1 || y >= g.b.y1) continue;
        float v[3] = {0, 0, 0}, o = 0;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const float t = tap(dx) * tap(dy);
                const int sy = 2 * ly + dy + 2, sx = 2 * lx + dx + 2;
                for (int c = 0; c < 3; ++c) v[c] = fmaf(t, s_c[c][sy][sx], v[c]);
                o = fmaf(t, s_o[sy][sx], o);
            }
        float* d = g.ptr(x, y);
        for (int c = 0; c < 3; ++c) d[c] = v[c];
        *u.ptr(x, y) = o;
    }
}

// Gaussian level l >= 2: reduce of the level below (colour and ownership).
__global__ void __launch_bounds__(kThreads) chain_gauss(Plane gs, Plane us, int fw, int fh, Plane g, Plane u) {
    GRID_STRIDE(i, g.b.n()) {
        const int x = g.b.x0 + static_cast<int>(i % g.b.w()), y = g.b.y0 + static_cast<int>(i / g.b.w());
        float v[3] = {0, 0, 0}, o = 0;
        for (int dy = -2; dy <= 2; ++dy) {
            const int yy = reflect101(2 * y + dy, fh);
            for (int dx = -2; dx <= 2; ++dx) {
                const int xx = reflect101(2 * x + dx, fw);
                const float t = tap(dx) * tap(dy);
                for (int c = 0; c < 3; ++c) v[c] = fmaf(t, gs.at(xx, yy, c), v[c]);
                o = fmaf(t, us.at(xx, yy, 0), o);
            }
        }
        float* d = g.ptr(x, y);
        for (int c = 0; c < 3; ++c) d[c] = v[c];
        *u.ptr(x, y) = o;
    }
}

// Per-frame planes of one blended level, for the mix.
struct MixFrame {
    Plane g, u, g_up;  // Gaussian and ownership at this level, Gaussian one level up
    Box zone;          // where this frame may have ownership weight
};
struct MixLevel {
    MixFrame f[kMaxChainFrames];
    int n;
};

// Band mix and collapse of level l (1..blend top): Z = sum U L / sum U plus
// the expansion of the collapsed level above.
__global__ void __launch_bounds__(kThreads) chain_mix(const MixLevel* m, int lw, int lh, int cw, int ch, bool coarse,
                                                     Plane z, Plane z_up) {
    GRID_STRIDE(i, static_cast<long long>(lw) * lh) {
        const int x = static_cast<int>(i % lw), y = static_cast<int>(i / lw);
        float total = 0, acc[3] = {0, 0, 0};
        bool any = false;
        for (int k = m->n - 1; k >= 0; --k) {
            const MixFrame& f = m->f[k];
            if (!f.zone.has(x, y)) continue;
            const float uk = f.u.at(x, y, 0);
            if (uk == 0) continue;
            for (int c = 0; c < 3; ++c) {
                const float lk = f.g.at(x, y, c) - (coarse ? expand_at(f.g_up, cw, ch, x, y, c) : 0.0f);
                acc[c] = any ? fmaf(uk, lk, acc[c]) : uk * lk;
            }
            any = true;
        }
        // The denominator in frame order.
        for (int k = 0; k < m->n; ++k) {
            const MixFrame& f = m->f[k];
            if (f.zone.has(x, y)) total += f.u.at(x, y, 0);
        }
        float* d = z.ptr(x, y);
        for (int c = 0; c < 3; ++c) {
            const float mix = total > 0 ? acc[c] / total : 0.0f;
            d[c] = coarse ? mix + expand_at(z_up, cw, ch, x, y, c) : mix;
        }
    }
}

// Overlap pixels of level 0: the owner's band plus the collapsed level 1.
__global__ void __launch_bounds__(kThreads) chain_finish(const Frames* fr, const MixLevel* m1, const std::int8_t* owner,
                                                        const std::uint8_t* count, int W, int H, int w1, int h1,
                                                        bool coarse, Plane z1, std::uint8_t* out, int out_pitch) {
    GRID_STRIDE(i, static_cast<long long>(W) * H) {
        if (count[i] < 2) continue;
        const int x = static_cast<int>(i % W), y = static_cast<int>(i / W);
        const int k = owner[i];
        const Level0& f = fr->f[k];
        const size_t j = static_cast<size_t>(y - f.box.y0) * f.box.w() + (x - f.box.x0);
        std::uint8_t* o = out + static_cast<size_t>(y) * out_pitch + 3 * x;
        for (int c = 0; c < 3; ++c) {
            const float own = f.color[j * 3 + c];
            // One owner with weight 1: the mix is its Laplacian.
            const float band = coarse ? own - expand_at(m1->f[k].g, w1, h1, x, y, c) : own;
            o[c] = quantize(coarse ? band + expand_at(z1, w1, h1, x, y, c) : band);
        }
    }
}

unsigned grid_for(long long n, long long cap = 1 << 16) {
    const long long g = (n + kThreads - 1) / kThreads;
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>(g, cap)));
}

// ------------------------------------------------------------------ layout
struct FramePlan {
    Box box;                  // level-0 canvas box
    Box pull[kMaxLevels];     // levels 1..top
    Box fill[kMaxLevels];     // levels 1..top
    Box gauss[kMaxLevels];    // levels 1..blend top
    Box zone[kMaxLevels];     // levels 1..blend top
    std::size_t off_color, off_valid, off_score;
    std::size_t off_pull[kMaxLevels], off_fill[kMaxLevels], off_gauss[kMaxLevels], off_own[kMaxLevels];
};

struct Plan {
    int n, top, blend_top;
    int lw[kMaxLevels], lh[kMaxLevels];
    std::vector<FramePlan> f;
    std::size_t off_owner, off_count, off_z[kMaxLevels], off_frames, off_mix[kMaxLevels], bytes;
};

Box clip(Box b, int w, int h) {
    b.x0 = std::clamp(b.x0, 0, w);
    b.x1 = std::clamp(b.x1, 0, w);
    b.y0 = std::clamp(b.y0, 0, h);
    b.y1 = std::clamp(b.y1, 0, h);
    if (b.x1 <= b.x0 || b.y1 <= b.y0) b = {0, 0, 1, 1};  // never empty: a 1-pixel window of zeros
    return b;
}

// Level-l box of a level-0 box, widened by m level pixels.
Box scaled(const Box& b0, int l, int m, int w, int h) {
    const long long d = 1LL << l;
    Box b{static_cast<int>(b0.x0 / d) - m, static_cast<int>(b0.y0 / d) - m,
          static_cast<int>((b0.x1 + d - 1) / d) + m, static_cast<int>((b0.y1 + d - 1) / d) + m};
    return clip(b, w, h);
}

// Canvas box a frame can cover. A canvas point samples the frame when its
// image lies in the frame rectangle widened by the snap tolerance; where the
// forward map (the adjugate of the canvas -> frame map, long double) keeps
// one denominator sign over that rectangle, its image is the convex hull of
// the four mapped corners, widened here by 2 pixels for rounding. Otherwise
// (or when a corner cannot be mapped) the box is the whole canvas.
Box frame_box(const Geometry& g, bool shift) {
    if (shift) return clip({-g.origin_x, -g.origin_y, -g.origin_x + g.width_a, -g.origin_y + g.height_a}, g.width,
                           g.height);
    using R = long double;
    const double* m = g.inverse;
    R a[9];
    a[0] = R(m[4]) * m[8] - R(m[5]) * m[7];
    a[1] = R(m[2]) * m[7] - R(m[1]) * m[8];
    a[2] = R(m[1]) * m[5] - R(m[2]) * m[4];
    a[3] = R(m[5]) * m[6] - R(m[3]) * m[8];
    a[4] = R(m[0]) * m[8] - R(m[2]) * m[6];
    a[5] = R(m[2]) * m[3] - R(m[0]) * m[5];
    a[6] = R(m[3]) * m[7] - R(m[4]) * m[6];
    a[7] = R(m[1]) * m[6] - R(m[0]) * m[7];
    a[8] = R(m[0]) * m[4] - R(m[1]) * m[3];
    R x0 = 1e300L, y0 = 1e300L, x1 = -1e300L, y1 = -1e300L;
    const R t = R(g.snap_tolerance) + 1e-3L;
    int sign = 0;
    for (int c = 0; c < 4; ++c) {
        const R u = (c & 1) ? g.width_b - 1 + t : -t, v = (c & 2) ? g.height_b - 1 + t : -t;
        const R w = a[6] * u + a[7] * v + a[8];
        const int s = w > 0 ? 1 : (w < 0 ? -1 : 0);
        if (s == 0 || (sign != 0 && s != sign)) return {0, 0, g.width, g.height};
        sign = s;
        const R x = (a[0] * u + a[1] * v + a[2]) / w - g.origin_x, y = (a[3] * u + a[4] * v + a[5]) / w - g.origin_y;
        if (!std::isfinite(static_cast<double>(x)) || !std::isfinite(static_cast<double>(y)))
            return {0, 0, g.width, g.height};
        x0 = std::min(x0, x);
        y0 = std::min(y0, y);
        x1 = std::max(x1, x);
        y1 = std::max(y1, y);
    }
    const auto lo = [](R v) { return static_cast<int>(std::max<R>(-1e9L, std::floor(v) - 2)); };
    const auto hi = [](R v) { return static_cast<int>(std::min<R>(1e9L, std::ceil(v) + 3)); };
    return clip({lo(x0), lo(y0), hi(x1), hi(y1)}, g.width, g.height);
}

bool consistent(int n, const Geometry* g) {
    if (n < 1 || n > kMaxChainFrames || !g) return false;
    for (int k = 0; k < n; ++k) {
        const Geometry& a = g[k];
        if (a.width <= 0 || a.height <= 0 || a.width_b <= 0 || a.height_b <= 0 || a.width != g[0].width ||
            a.height != g[0].height || a.origin_x != g[0].origin_x || a.origin_y != g[0].origin_y ||
            a.width_a != g[0].width_a || a.height_a != g[0].height_a || a.origin_x > 0 || a.origin_y > 0 ||
            static_cast<long long>(a.origin_x) + a.width < a.width_a ||
            static_cast<long long>(a.origin_y) + a.height < a.height_a || !std::isfinite(a.denominator_tolerance) ||
            !(a.denominator_tolerance > 0) || a.denominator_tolerance >= 1 || !std::isfinite(a.snap_tolerance) ||
            a.snap_tolerance < 0 || a.snap_tolerance >= 0.5)
            return false;
        bool nonzero = false;
        for (double v : a.inverse) {
            if (!std::isfinite(v)) return false;
            nonzero |= v != 0;
        }
        if (!nonzero) return false;
    }
    return g[0].width_b == g[0].width_a && g[0].height_b == g[0].height_a;
}

Status make_plan(int n, const Geometry* g, BlendMode mode, Plan* plan) {
    const int requested = requested_reductions(mode);
    if (requested < 0 || !consistent(n, g)) return Status::INVALID_ARGUMENT;
    if (static_cast<std::size_t>(g[0].width) > std::numeric_limits<std::size_t>::max() / g[0].height)
        return Status::INVALID_ARGUMENT;
    Plan& P = *plan;
    P.n = n;
    P.lw[0] = g[0].width;
    P.lh[0] = g[0].height;
    P.top = 0;
    while (P.lw[P.top] > 1 || P.lh[P.top] > 1) {
        if (P.top + 1 >= kMaxLevels) return Status::INVALID_ARGUMENT;
        P.lw[P.top + 1] = P.lw[P.top] / 2 + P.lw[P.top] % 2;
        P.lh[P.top + 1] = P.lh[P.top] / 2 + P.lh[P.top] % 2;
        ++P.top;
    }
    P.blend_top = std::min(requested, P.top);
    const bool bands = mode != BlendMode::Feather;
    std::size_t cur = 0;
    const auto take = [&](std::size_t bytes) {
        cur = (cur + 255) / 256 * 256;
        const std::size_t at = cur;
        cur += std::max<std::size_t>(bytes, 1);
        return at;
    };
    // Margins: Gaussian windows double (plus 4) level by level downwards,
    // because each reduce reads twice its width one level below; filled
    // windows cover what the finer filled level and the level-1 Gaussian
    // read (halving upwards, plus 3).
    int mg[kMaxLevels] = {}, mf[kMaxLevels] = {};
    if (P.blend_top >= 1) {
        mg[P.blend_top] = kGTop;
        for (int l = P.blend_top - 1; l >= 1; --l) mg[l] = std::min(2 * mg[l + 1] + 4, 1 << 24);
    }
    mf[1] = (P.blend_top >= 1 ? mg[1] : kGTop) + 4;
    for (int l = 2; l <= P.top; ++l) mf[l] = (mf[l - 1] + 1) / 2 + 3;
    const std::size_t canvas = static_cast<std::size_t>(P.lw[0]) * P.lh[0];
    P.f.assign(n, FramePlan{});
    for (int k = 0; k < n; ++k) {
        FramePlan& F = P.f[k];
        F.box = frame_box(g[k], k == 0);
        F.off_color = take(static_cast<std::size_t>(F.box.n()) * 12);
        F.off_valid = take(static_cast<std::size_t>(F.box.n()));
        F.off_score = take(static_cast<std::size_t>(F.box.n()) * 8);
        if (!bands) continue;
        for (int l = 1; l <= P.top; ++l) {
            F.pull[l] = scaled(F.box, l, 3, P.lw[l], P.lh[l]);
            F.fill[l] = scaled(F.box, l, mf[l], P.lw[l], P.lh[l]);
            F.off_pull[l] = take(static_cast<std::size_t>(F.pull[l].n()) * 16);
            F.off_fill[l] = take(static_cast<std::size_t>(F.fill[l].n()) * 12);
            if (l <= P.blend_top) {
                F.gauss[l] = scaled(F.box, l, mg[l], P.lw[l], P.lh[l]);
                F.zone[l] = scaled(F.box, l, kZMargin, P.lw[l], P.lh[l]);
                F.off_gauss[l] = take(static_cast<std::size_t>(F.gauss[l].n()) * 12);
                F.off_own[l] = take(static_cast<std::size_t>(F.gauss[l].n()) * 4);
            }
        }
    }
    P.off_owner = take(canvas);
    P.off_count = take(canvas);
    P.off_frames = take(sizeof(Frames));
    for (int l = 1; l <= P.blend_top; ++l) {
        P.off_z[l] = take(static_cast<std::size_t>(P.lw[l]) * P.lh[l] * 12);
        P.off_mix[l] = take(sizeof(MixLevel));
    }
    P.bytes = cur;
    return Status::OK;
}

bool image_ok(const ImageView& v, PixelFormat f, int w, int h) {
    return v.data && v.format == f && v.width == w && v.height == h &&
           v.pitch_bytes >= v.width * (f == PixelFormat::RGB8 ? 3 : 1);
}

// Byte range [lo, hi) from an image's first byte to its last pixel byte.
struct Span {
    std::uintptr_t lo, hi;
    bool meets(const Span& o) const { return lo < o.hi && o.lo < hi; }
};
Span span_of(const ImageView& v, int bytes_per_pixel) {
    const auto lo = reinterpret_cast<std::uintptr_t>(v.data);
    return {lo, lo + static_cast<std::uintptr_t>(v.height - 1) * v.pitch_bytes +
                    static_cast<std::uintptr_t>(v.width) * bytes_per_pixel};
}

}  // namespace

Status chain_workspace_bytes(int n, const Geometry* g, BlendMode mode, std::size_t* bytes) {
    if (!bytes) return Status::INVALID_ARGUMENT;
    *bytes = 0;
    Plan plan{};
    const Status s = make_plan(n, g, mode, &plan);
    if (s == Status::OK) *bytes = plan.bytes;
    return s;
}

Status render_chain(int n, const ImageView* frames, const Geometry* g, BlendMode mode, const ImageView& output,
                    const ImageView& mask, void* workspace, std::size_t workspace_bytes, hipStream_t stream,
                    RenderStatistics* statistics) {
    Plan P{};
    const Status s = make_plan(n, g, mode, &P);
    if (s != Status::OK) return s;
    if (!frames || !image_ok(output, PixelFormat::RGB8, g[0].width, g[0].height) ||
        !image_ok(mask, PixelFormat::GRAY8, g[0].width, g[0].height))
        return Status::INVALID_ARGUMENT;
    for (int k = 0; k < n; ++k)
        if (!image_ok(frames[k], PixelFormat::RGB8, g[k].width_b, g[k].height_b)) return Status::INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < P.bytes) return Status::INSUFFICIENT_WORKSPACE;
    if (reinterpret_cast<std::uintptr_t>(workspace) % 256) return Status::INVALID_ARGUMENT;
    // The outputs may overlap neither each other, nor a frame, nor the
    // workspace; frames may share memory (they are only read).
    const Span pano = span_of(output, 3), union_mask = span_of(mask, 1);
    const auto ws_lo = reinterpret_cast<std::uintptr_t>(workspace);
    const Span arena{ws_lo, ws_lo + P.bytes};
    const auto st_lo = reinterpret_cast<std::uintptr_t>(statistics);
    const Span counters{st_lo, st_lo + (statistics ? sizeof(RenderStatistics) : 0)};
    if (reinterpret_cast<std::uintptr_t>(statistics) % alignof(RenderStatistics)) return Status::INVALID_ARGUMENT;
    if (pano.meets(union_mask) || pano.meets(arena) || union_mask.meets(arena) || counters.meets(pano) ||
        counters.meets(union_mask) || counters.meets(arena))
        return Status::INVALID_ARGUMENT;
    for (int k = 0; k < n; ++k) {
        const Span in = span_of(frames[k], 3);
        if (in.meets(pano) || in.meets(union_mask) || in.meets(arena) || in.meets(counters))
            return Status::INVALID_ARGUMENT;
    }
    auto* base = static_cast<std::uint8_t*>(workspace);
    const auto fptr = [&](std::size_t off) { return reinterpret_cast<float*>(base + off); };
    const int W = P.lw[0], H = P.lh[0];
    const bool bands = mode != BlendMode::Feather;
    // Host copies of the per-launch tables, uploaded once.
    Frames fr{};
    fr.n = n;
    for (int k = 0; k < n; ++k) {
        Level0& f = fr.f[k];
        const FramePlan& F = P.f[k];
        f.img = frames[k].data;
        f.pitch = frames[k].pitch_bytes;
        f.w = frames[k].width;
        f.h = frames[k].height;
        std::copy(std::begin(g[k].inverse), std::end(g[k].inverse), f.inv);
        f.den_tol = g[k].denominator_tolerance;
        f.snap_tol = g[k].snap_tolerance;
        f.shift = k == 0;
        f.box = F.box;
        f.color = fptr(F.off_color);
        f.valid = base + F.off_valid;
        f.score = reinterpret_cast<double*>(base + F.off_score);
    }
    std::vector<MixLevel> mix(P.blend_top + 1);
    for (int l = 1; l <= P.blend_top; ++l) {
        MixLevel& m = mix[l];
        m.n = n;
        for (int k = 0; k < n; ++k) {
            const FramePlan& F = P.f[k];
            m.f[k].g = {fptr(F.off_gauss[l]), F.gauss[l], 3};
            m.f[k].u = {fptr(F.off_own[l]), F.gauss[l], 1};
            m.f[k].g_up = l < P.blend_top ? Plane{fptr(F.off_gauss[l + 1]), F.gauss[l + 1], 3} : m.f[k].g;
            m.f[k].zone = F.zone[l];
        }
    }
    auto* d_frames = reinterpret_cast<Frames*>(base + P.off_frames);
    if (hipMemcpyAsync(d_frames, &fr, sizeof fr, hipMemcpyHostToDevice, stream) != hipSuccess) return Status::CUDA_ERROR;
    for (int l = 1; l <= P.blend_top; ++l)
        if (hipMemcpyAsync(base + P.off_mix[l], &mix[l], sizeof(MixLevel), hipMemcpyHostToDevice, stream) != hipSuccess)
            return Status::CUDA_ERROR;
    auto* owner = reinterpret_cast<std::int8_t*>(base + P.off_owner);
    auto* count = base + P.off_count;
    auto* stats = reinterpret_cast<unsigned long long*>(statistics);
    if (stats) chain_clear_stats<<<1, 64, 0, stream>>>(stats);
    for (int k = 0; k < n; ++k)
        chain_warp<<<grid_for(fr.f[k].box.n()), kThreads, 0, stream>>>(fr.f[k], g[0].origin_x, g[0].origin_y);
    chain_own<<<grid_for(static_cast<long long>(W) * H, 256 * 8), kThreads, 0, stream>>>(
        d_frames, W, H, !bands, owner, count, output.data, output.pitch_bytes, mask.data, mask.pitch_bytes, stats);
    if (bands) {
        for (int k = 0; k < n; ++k) {
            const FramePlan& F = P.f[k];
            const auto pull = [&](int l) { return Plane{fptr(F.off_pull[l]), F.pull[l], 4}; };
            const auto fill = [&](int l) { return Plane{fptr(F.off_fill[l]), F.fill[l], 3}; };
            chain_pull1<<<grid_for(F.pull[1].n()), kThreads, 0, stream>>>(fr.f[k], W, H, pull(1));
            for (int l = 2; l <= P.top; ++l)
                chain_pull<<<grid_for(F.pull[l].n()), kThreads, 0, stream>>>(pull(l - 1), P.lw[l - 1], P.lh[l - 1],
                                                                             pull(l));
            for (int l = P.top; l >= 1; --l) {
                const bool top = l == P.top;
                const Plane up = top ? fill(l) : fill(l + 1);
                chain_push<<<grid_for(F.fill[l].n()), kThreads, 0, stream>>>(
                    pull(l), fill(l), up, top ? 1 : P.lw[l + 1], top ? 1 : P.lh[l + 1], top);
            }
            if (P.blend_top >= 1) {
                const unsigned g1_tiles = static_cast<unsigned>(((F.gauss[1].w() + kG1W - 1) / kG1W) *
                                                                ((F.gauss[1].h() + kG1H - 1) / kG1H));
                chain_gauss1<<<g1_tiles, kThreads, 0, stream>>>(
                    fr.f[k], k, owner, W, H, fill(1), P.lw[1], P.lh[1], mix[1].f[k].g, mix[1].f[k].u);
                for (int l = 2; l <= P.blend_top; ++l)
                    chain_gauss<<<grid_for(F.gauss[l].n()), kThreads, 0, stream>>>(
                        mix[l - 1].f[k].g, mix[l - 1].f[k].u, P.lw[l - 1], P.lh[l - 1], mix[l].f[k].g, mix[l].f[k].u);
            }
        }
        Plane z_up{};
        for (int l = P.blend_top; l >= 1; --l) {
            const Plane z{fptr(P.off_z[l]), {0, 0, P.lw[l], P.lh[l]}, 3};
            const bool coarse = l < P.blend_top;
            chain_mix<<<grid_for(static_cast<long long>(P.lw[l]) * P.lh[l]), kThreads, 0, stream>>>(
                reinterpret_cast<const MixLevel*>(base + P.off_mix[l]), P.lw[l], P.lh[l], coarse ? P.lw[l + 1] : 1,
                coarse ? P.lh[l + 1] : 1, coarse, z, coarse ? z_up : z);
            z_up = z;
        }
        const Plane z1 = P.blend_top >= 1 ? z_up : Plane{fptr(P.off_count), {0, 0, 1, 1}, 1};
        chain_finish<<<grid_for(static_cast<long long>(W) * H), kThreads, 0, stream>>>(
            d_frames, P.blend_top >= 1 ? reinterpret_cast<const MixLevel*>(base + P.off_mix[1]) : nullptr, owner, count,
            W, H, P.lw[1], P.lh[1], P.blend_top >= 1, z1, output.data, output.pitch_bytes);
    }
    return hipGetLastError() == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

}  // namespace stitch::s6
