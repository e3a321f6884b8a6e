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
// case, so no access leaves its allocation. One launch per level covers
// every frame (grid y = frame), from tables uploaded in one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
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

struct Frames {
    Level0 f[kMaxChainFrames];
    int n;
};

// Per-frame arguments of one launch; a launch covers every frame, frame k
// taking the blocks with blockIdx.y == k, which stride over its window.
struct PullJob {  // pull into level l >= 2
    Plane src, dst;
    int fw, fh;   // level l - 1
};
struct PushJob {  // fill level l
    Plane pull, dst, coarse;
    int cw, ch;   // level l + 1
};
struct Gauss1Job {
    Plane f1, g, u;
};
struct GaussJob {  // Gaussian into level l >= 2
    Plane gs, us, g, u;
    int fw, fh;
};

#define FRAME_STRIDE(i, n) \
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
         i += static_cast<long long>(gridDim.x) * blockDim.x)

// Level 0 of every frame over its box.
__global__ void __launch_bounds__(kThreads) chain_warp(const Frames* fr, int origin_x, int origin_y) {
    // A copy: stores to the frame's planes could otherwise alias the table.
    const Level0 f = fr->f[blockIdx.y];
    FRAME_STRIDE(i, f.box.n()) {
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

// What the per-pixel kernels read of each frame, staged in LDS once per
// block (reads through the global table would be repeated after every store
// the compiler cannot prove apart from it).
struct FrameRef {
    Box box;
    const float* color;
    const std::uint8_t* valid;
    const double* score;
};

__device__ __forceinline__ int stage_frames(const Frames* fr, FrameRef* refs) {
    const int n = fr->n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const Level0& f = fr->f[k];
        refs[k] = {f.box, f.color, f.valid, f.score};
    }
    __syncthreads();
    return n;
}

// Owner and valid count per canvas pixel; the mask, the statistics, and the
// output of every pixel the band mix does not decide (feather: all of them).
__global__ void __launch_bounds__(kThreads) chain_own(const Frames* fr, int W, int H, bool feather, std::int8_t* owner,
                                                     std::uint8_t* count, std::uint8_t* out, int out_pitch,
                                                     std::uint8_t* mask, int mask_pitch,
                                                     unsigned long long* stats) {
    __shared__ FrameRef refs[kMaxChainFrames];
    const int n = stage_frames(fr, refs);
    unsigned long long uni = 0, ovl = 0;
    GRID_STRIDE(i, static_cast<long long>(W) * H) {
        const int x = static_cast<int>(i % W), y = static_cast<int>(i / W);
        int own = -1, valid = 0;
        double best = 0, fs = 0, fc[3] = {0, 0, 0};
        float single[3] = {0, 0, 0};
        for (int k = 0; k < n; ++k) {
            const FrameRef& f = refs[k];
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
                // The score-weighted sums only feather uses.
                if (feather) fc[c] = valid == 1 ? s * v : fc[c] + s * v;
                single[c] = v;
            }
            if (feather) fs = valid == 1 ? s : fs + s;
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
        // One atomic pair per block (the grid is capped at 64 blocks per CU).
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
// frame's level 0. A block makes a kG1W x kG1H tile of level-1 pixels from
// the (2 kG1W + 3) x (2 kG1H + 3) level-0 pixels it reads, staged in LDS with
// zeros outside the frame's box (a zero tap adds exactly nothing to these
// non-negative sums, so the result is the full filter's).
constexpr int kG1W = 32, kG1H = 8;  // 256 outputs: one per thread
constexpr int kG1InW = 2 * kG1W + 3, kG1InH = 2 * kG1H + 3;

__global__ void __launch_bounds__(kThreads) chain_pull1(const Frames* fr, int W0, int H0, const Plane* dsts) {
    __shared__ float s_v[4][kG1InH][kG1InW];
    const Level0 f = fr->f[blockIdx.y];
    const Plane dst = dsts[blockIdx.y];
    const int tiles_x = (dst.b.w() + kG1W - 1) / kG1W;
    const int tiles = tiles_x * ((dst.b.h() + kG1H - 1) / kG1H);
    for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int ox0 = dst.b.x0 + (t % tiles_x) * kG1W;
        const int oy0 = dst.b.y0 + (t / tiles_x) * kG1H;
        const int ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
        for (int i = threadIdx.x; i < kG1InW * kG1InH; i += blockDim.x) {
            const int lx = i % kG1InW, ly = i / kG1InW;
            const int xx = reflect101(min(ix0 + lx, W0 + 1), W0), yy = reflect101(min(iy0 + ly, H0 + 1), H0);
            float v[4] = {0, 0, 0, 0};
            if (f.box.has(xx, yy)) {
                const size_t j = static_cast<size_t>(yy - f.box.y0) * f.box.w() + (xx - f.box.x0);
                for (int c = 0; c < 3; ++c) v[c] = f.color[j * 3 + c];
                v[3] = f.valid[j] ? 1.0f : 0.0f;
            }
            for (int c = 0; c < 4; ++c) s_v[c][ly][lx] = v[c];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kG1W * kG1H; i += blockDim.x) {
            const int lx = i % kG1W, ly = i / kG1W;
            const int x = ox0 + lx, y = oy0 + ly;
            if (x >= dst.b.x1 || y >= dst.b.y1) continue;
            float v[4] = {0, 0, 0, 0};
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx) {
                    const float tp = tap(dx) * tap(dy);
                    const int sy = 2 * ly + dy + 2, sx = 2 * lx + dx + 2;
                    for (int c = 0; c < 4; ++c) v[c] = fmaf(tp, s_v[c][sy][sx], v[c]);
                }
            float* d = dst.ptr(x, y);
            for (int c = 0; c < 4; ++c) d[c] = v[c];
        }
        __syncthreads();
    }
}

// Pull into level l >= 2 from level l - 1 (zero outside the source window).
__global__ void __launch_bounds__(kThreads) chain_pull(const PullJob* jobs) {
    const PullJob J = jobs[blockIdx.y];
    FRAME_STRIDE(i, J.dst.b.n()) {
        const int x = J.dst.b.x0 + static_cast<int>(i % J.dst.b.w()), y = J.dst.b.y0 + static_cast<int>(i / J.dst.b.w());
        float v[4] = {0, 0, 0, 0};
        for (int dy = -2; dy <= 2; ++dy) {
            const int yy = reflect101(2 * y + dy, J.fh);
            for (int dx = -2; dx <= 2; ++dx) {
                const int xx = reflect101(2 * x + dx, J.fw);
                if (!J.src.b.has(xx, yy)) continue;
                const float t = tap(dx) * tap(dy);
                const float* sp = J.src.ptr(xx, yy);
                for (int c = 0; c < 4; ++c) v[c] = fmaf(t, sp[c], v[c]);
            }
        }
        float* d = J.dst.ptr(x, y);
        for (int c = 0; c < 4; ++c) d[c] = v[c];
    }
}

// Expansion of the three channels of coarse planes (cw x ch level) at fine
// pixel (x, y): the taps whose reflected fine position is even, times 4,
// each channel summed in the taps' order. Tap positions are found once for
// all channels and, in expand3x2, for both planes (each read clamped to its
// own window).
__device__ __forceinline__ void expand3(const Plane& a, int cw, int ch, int x, int y, float* ea) {
    float sa[3] = {0, 0, 0};
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(y + dy, 2 * ch);
        if (yy & 1) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101(x + dx, 2 * cw);
            if (xx & 1) continue;
            const float t = tap(dx) * tap(dy);
            for (int c = 0; c < 3; ++c) sa[c] = fmaf(t, a.at(xx / 2, yy / 2, c), sa[c]);
        }
    }
    for (int c = 0; c < 3; ++c) ea[c] = 4 * sa[c];
}

__device__ __forceinline__ void expand3x2(const Plane& a, const Plane& b, int cw, int ch, int x, int y, float* ea,
                                          float* eb) {
    float sa[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(y + dy, 2 * ch);
        if (yy & 1) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101(x + dx, 2 * cw);
            if (xx & 1) continue;
            const float t = tap(dx) * tap(dy);
            for (int c = 0; c < 3; ++c) {
                sa[c] = fmaf(t, a.at(xx / 2, yy / 2, c), sa[c]);
                sb[c] = fmaf(t, b.at(xx / 2, yy / 2, c), sb[c]);
            }
        }
    }
    for (int c = 0; c < 3; ++c) {
        ea[c] = 4 * sa[c];
        eb[c] = 4 * sb[c];
    }
}

// Pull value (colour sum c or weight 3) of a windowed pull plane: zero outside.
__device__ __forceinline__ float pulled(const Plane& p, int x, int y, int c) {
    return p.b.has(x, y) ? p.at(x, y, c) : 0.0f;
}

// Normalised pull where the level has weight, else the expansion of the
// filled level above (none at the top level: zero).
__global__ void __launch_bounds__(kThreads) chain_push(const PushJob* jobs, bool top) {
    const PushJob J = jobs[blockIdx.y];
    FRAME_STRIDE(i, J.dst.b.n()) {
        const int x = J.dst.b.x0 + static_cast<int>(i % J.dst.b.w()), y = J.dst.b.y0 + static_cast<int>(i / J.dst.b.w());
        const float wgt = pulled(J.pull, x, y, 3);
        float* d = J.dst.ptr(x, y);
        if (wgt > 0) {
            for (int c = 0; c < 3; ++c) d[c] = pulled(J.pull, x, y, c) / wgt;
        } else if (top) {
            for (int c = 0; c < 3; ++c) d[c] = 0.0f;
        } else {
            float e[3];
            expand3(J.coarse, J.cw, J.ch, x, y, e);
            for (int c = 0; c < 3; ++c) d[c] = e[c];
        }
    }
}

// Gaussian level 1 of the filled frame and its ownership weight. The filled
// level-0 colour is the frame's own where valid, else the expansion of its
// filled level 1. A block makes a kG1W x kG1H tile of level-1 pixels from
// the (2 kG1W + 3) x (2 kG1H + 3) level-0 pixels the 5x5 reduce reads,
// staged once in LDS (filled colour, and ownership as 0/1), so each filled
// value is formed once per tile instead of once per tap. Input positions are reflected into the
// canvas as they are staged; positions past what the tile's real outputs
// read are clamped (their values are never used).
// The filled level 1 that the expansions read is staged too: a
// kF1W x kF1H region starting one coarse pixel before the first input's
// half, which holds every tap of every staged input (reflected inputs lie
// inside the staged input range); a tap outside it, which the layout never
// produces, would be read from memory.
constexpr int kF1W = (kG1InW + 4) / 2 + 3, kF1H = (kG1InH + 4) / 2 + 3;

__global__ void __launch_bounds__(kThreads) chain_gauss1(const Frames* fr, const Gauss1Job* jobs,
                                                        const std::int8_t* owner, int W0, int H0, int w1, int h1) {
    __shared__ float s_c[3][kG1InH][kG1InW];
    __shared__ std::uint8_t s_o[kG1InH][kG1InW];
    __shared__ float s_f[3][kF1H][kF1W];
    const int k = static_cast<int>(blockIdx.y);
    const Level0& f = fr->f[k];
    const Gauss1Job& J = jobs[k];
    const int tiles_x = (J.g.b.w() + kG1W - 1) / kG1W;
    const int tiles = tiles_x * ((J.g.b.h() + kG1H - 1) / kG1H);
    for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int ox0 = J.g.b.x0 + (t % tiles_x) * kG1W;
        const int oy0 = J.g.b.y0 + (t / tiles_x) * kG1H;
        const int ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
        // Staged filled level 1: [fx0, fx0 + kF1W) x [fy0, fy0 + kF1H).
        const int fx0 = max(ix0, 0) / 2 - 2, fy0 = max(iy0, 0) / 2 - 2;
        for (int i = threadIdx.x; i < kF1W * kF1H; i += blockDim.x) {
            const int lx = i % kF1W, ly = i / kF1W;
            for (int c = 0; c < 3; ++c) s_f[c][ly][lx] = J.f1.at(fx0 + lx, fy0 + ly, c);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kG1InW * kG1InH; i += blockDim.x) {
            const int lx = i % kG1InW, ly = i / kG1InW;
            const int xx = reflect101(min(ix0 + lx, W0 + 1), W0), yy = reflect101(min(iy0 + ly, H0 + 1), H0);
            float v[3];
            bool own_colour = false;
            if (f.box.has(xx, yy)) {
                const size_t j = static_cast<size_t>(yy - f.box.y0) * f.box.w() + (xx - f.box.x0);
                if (f.valid[j]) {
                    for (int c = 0; c < 3; ++c) v[c] = f.color[j * 3 + c];
                    own_colour = true;
                }
            }
            if (!own_colour) {
                // The expansion of the filled level 1 (expand3's taps and order).
                float sum[3] = {0, 0, 0};
                for (int dy = -2; dy <= 2; ++dy) {
                    const int ry = reflect101(yy + dy, 2 * h1);
                    if (ry & 1) continue;
                    for (int dx = -2; dx <= 2; ++dx) {
                        const int rx = reflect101(xx + dx, 2 * w1);
                        if (rx & 1) continue;
                        const float tp = tap(dx) * tap(dy);
                        const int cx = rx / 2 - fx0, cy = ry / 2 - fy0;
                        const bool staged = cx >= 0 && cx < kF1W && cy >= 0 && cy < kF1H;
                        for (int c = 0; c < 3; ++c)
                            sum[c] = fmaf(tp, staged ? s_f[c][cy][cx] : J.f1.at(rx / 2, ry / 2, c), sum[c]);
                    }
                }
                for (int c = 0; c < 3; ++c) v[c] = 4 * sum[c];
            }
            for (int c = 0; c < 3; ++c) s_c[c][ly][lx] = v[c];
            s_o[ly][lx] = owner[static_cast<size_t>(yy) * W0 + xx] == k ? 1 : 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kG1W * kG1H; i += blockDim.x) {
            const int lx = i % kG1W, ly = i / kG1W;
            const int x = ox0 + lx, y = oy0 + ly;
            if (x >= J.g.b.x1 || y >= J.g.b.y1) continue;
            float v[3] = {0, 0, 0}, o = 0;
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx) {
                    const float tp = tap(dx) * tap(dy);
                    const int sy = 2 * ly + dy + 2, sx = 2 * lx + dx + 2;
                    for (int c = 0; c < 3; ++c) v[c] = fmaf(tp, s_c[c][sy][sx], v[c]);
                    o = fmaf(tp, s_o[sy][sx] ? 1.0f : 0.0f, o);
                }
            float* d = J.g.ptr(x, y);
            for (int c = 0; c < 3; ++c) d[c] = v[c];
            *J.u.ptr(x, y) = o;
        }
        __syncthreads();
    }
}

// Gaussian level l >= 2: reduce of the level below (colour and ownership).
__global__ void __launch_bounds__(kThreads) chain_gauss(const GaussJob* jobs) {
    const GaussJob J = jobs[blockIdx.y];
    FRAME_STRIDE(i, J.g.b.n()) {
        const int x = J.g.b.x0 + static_cast<int>(i % J.g.b.w()), y = J.g.b.y0 + static_cast<int>(i / J.g.b.w());
        float v[3] = {0, 0, 0}, o = 0;
        for (int dy = -2; dy <= 2; ++dy) {
            const int yy = reflect101(2 * y + dy, J.fh);
            for (int dx = -2; dx <= 2; ++dx) {
                const int xx = reflect101(2 * x + dx, J.fw);
                const float t = tap(dx) * tap(dy);
                for (int c = 0; c < 3; ++c) v[c] = fmaf(t, J.gs.at(xx, yy, c), v[c]);
                o = fmaf(t, J.us.at(xx, yy, 0), o);
            }
        }
        float* d = J.g.ptr(x, y);
        for (int c = 0; c < 3; ++c) d[c] = v[c];
        *J.u.ptr(x, y) = o;
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
    // The frames' planes staged in LDS (see FrameRef).
    __shared__ MixFrame frames[kMaxChainFrames];
    const int n = m->n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) frames[k] = m->f[k];
    __syncthreads();
    GRID_STRIDE(i, static_cast<long long>(lw) * lh) {
        const int x = static_cast<int>(i % lw), y = static_cast<int>(i / lw);
        float total = 0, acc[3] = {0, 0, 0};
        bool any = false;
        for (int k = n - 1; k >= 0; --k) {
            const MixFrame& f = frames[k];
            if (!f.zone.has(x, y)) continue;
            const float uk = f.u.at(x, y, 0);
            if (uk == 0) continue;
            float e[3] = {0, 0, 0};
            if (coarse) expand3(f.g_up, cw, ch, x, y, e);
            for (int c = 0; c < 3; ++c) {
                const float lk = f.g.at(x, y, c) - (coarse ? e[c] : 0.0f);
                acc[c] = any ? fmaf(uk, lk, acc[c]) : uk * lk;
            }
            any = true;
        }
        // The denominator in frame order.
        for (int k = 0; k < n; ++k) {
            const MixFrame& f = frames[k];
            if (f.zone.has(x, y)) total += f.u.at(x, y, 0);
        }
        float ez[3] = {0, 0, 0};
        if (coarse) expand3(z_up, cw, ch, x, y, ez);
        float* d = z.ptr(x, y);
        for (int c = 0; c < 3; ++c) {
            const float mix = total > 0 ? acc[c] / total : 0.0f;
            d[c] = coarse ? mix + ez[c] : mix;
        }
    }
}

// Overlap pixels of level 0: the owner's band plus the collapsed level 1.
__global__ void __launch_bounds__(kThreads) chain_finish(const Frames* fr, const MixLevel* m1, const std::int8_t* owner,
                                                        const std::uint8_t* count, int W, int H, int w1, int h1,
                                                        bool coarse, Plane z1, std::uint8_t* out, int out_pitch) {
    __shared__ FrameRef refs[kMaxChainFrames];
    stage_frames(fr, refs);
    GRID_STRIDE(i, static_cast<long long>(W) * H) {
        if (count[i] < 2) continue;
        const int x = static_cast<int>(i % W), y = static_cast<int>(i / W);
        const int k = owner[i];
        const FrameRef& f = refs[k];
        const size_t j = static_cast<size_t>(y - f.box.y0) * f.box.w() + (x - f.box.x0);
        std::uint8_t* o = out + static_cast<size_t>(y) * out_pitch + 3 * x;
        float eg[3] = {0, 0, 0}, ez[3] = {0, 0, 0};
        if (coarse) expand3x2(m1->f[k].g, z1, w1, h1, x, y, eg, ez);
        for (int c = 0; c < 3; ++c) {
            const float own = f.color[j * 3 + c];
            // One owner with weight 1: the mix is its Laplacian.
            const float band = coarse ? own - eg[c] : own;
            o[c] = quantize(coarse ? band + ez[c] : band);
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

// Launch tables, uploaded in one copy: offsets within the table region.
struct TableLayout {
    std::size_t frames, mix[kMaxLevels], dst1, pull, push, gauss1, gauss, bytes;
};

struct Plan {
    int n, top, blend_top;
    int lw[kMaxLevels], lh[kMaxLevels];
    std::vector<FramePlan> f;
    std::size_t off_owner, off_count, off_z[kMaxLevels], off_tables, bytes;
    TableLayout t;
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
    for (int l = 1; l <= P.blend_top; ++l) P.off_z[l] = take(static_cast<std::size_t>(P.lw[l]) * P.lh[l] * 12);
    // Tables: the frames, one mix table per blended level, and per level one
    // job per frame for the pulls, pushes and Gaussians.
    std::size_t t = 0;
    const auto part = [&](std::size_t bytes) {
        t = (t + 255) / 256 * 256;
        const std::size_t at = t;
        t += std::max<std::size_t>(bytes, 1);
        return at;
    };
    const std::size_t levels = static_cast<std::size_t>(P.top) + 1;
    P.t.frames = part(sizeof(Frames));
    for (int l = 1; l <= P.blend_top; ++l) P.t.mix[l] = part(sizeof(MixLevel));
    P.t.dst1 = part(n * sizeof(Plane));
    P.t.pull = part(levels * n * sizeof(PullJob));
    P.t.push = part(levels * n * sizeof(PushJob));
    P.t.gauss1 = part(n * sizeof(Gauss1Job));
    P.t.gauss = part(levels * n * sizeof(GaussJob));
    P.t.bytes = t;
    P.off_tables = take(t);
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
    // Host image of the launch tables, uploaded in one copy.
    std::vector<std::uint8_t> host(P.t.bytes, 0);
    auto* tables = base + P.off_tables;
    const auto at = [&](std::size_t off) { return host.data() + off; };
    const auto dev = [&](std::size_t off) { return tables + off; };
    Frames& fr = *new (at(P.t.frames)) Frames{};
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
    const auto pull = [&](int k, int l) { return Plane{fptr(P.f[k].off_pull[l]), P.f[k].pull[l], 4}; };
    const auto fill = [&](int k, int l) { return Plane{fptr(P.f[k].off_fill[l]), P.f[k].fill[l], 3}; };
    const auto gauss = [&](int k, int l) { return Plane{fptr(P.f[k].off_gauss[l]), P.f[k].gauss[l], 3}; };
    const auto own = [&](int k, int l) { return Plane{fptr(P.f[k].off_own[l]), P.f[k].gauss[l], 1}; };
    auto* dst1 = reinterpret_cast<Plane*>(at(P.t.dst1));
    auto* pulls = reinterpret_cast<PullJob*>(at(P.t.pull));
    auto* pushes = reinterpret_cast<PushJob*>(at(P.t.push));
    auto* g1s = reinterpret_cast<Gauss1Job*>(at(P.t.gauss1));
    auto* gs = reinterpret_cast<GaussJob*>(at(P.t.gauss));
    // Largest per-frame extent of each launch, for its grid.
    long long n_warp = 0, n_pull[kMaxLevels] = {}, n_push[kMaxLevels] = {}, n_gauss[kMaxLevels] = {}, g1_tiles = 0,
              p1_tiles = 0;
    for (int k = 0; k < n; ++k) n_warp = std::max(n_warp, P.f[k].box.n());
    if (bands) {
        for (int l = 1; l <= P.blend_top; ++l) {
            MixLevel& m = *new (at(P.t.mix[l])) MixLevel{};
            m.n = n;
            for (int k = 0; k < n; ++k) {
                m.f[k].g = gauss(k, l);
                m.f[k].u = own(k, l);
                m.f[k].g_up = l < P.blend_top ? gauss(k, l + 1) : m.f[k].g;
                m.f[k].zone = P.f[k].zone[l];
            }
        }
        for (int k = 0; k < n; ++k) {
            const FramePlan& F = P.f[k];
            dst1[k] = pull(k, 1);
            p1_tiles = std::max<long long>(p1_tiles, static_cast<long long>((F.pull[1].w() + kG1W - 1) / kG1W) *
                                                         ((F.pull[1].h() + kG1H - 1) / kG1H));
            for (int l = 2; l <= P.top; ++l) {
                pulls[l * n + k] = {pull(k, l - 1), pull(k, l), P.lw[l - 1], P.lh[l - 1]};
                n_pull[l] = std::max(n_pull[l], F.pull[l].n());
            }
            for (int l = 1; l <= P.top; ++l) {
                const bool top = l == P.top;
                pushes[l * n + k] = {pull(k, l), fill(k, l), top ? fill(k, l) : fill(k, l + 1), top ? 1 : P.lw[l + 1],
                                     top ? 1 : P.lh[l + 1]};
                n_push[l] = std::max(n_push[l], F.fill[l].n());
            }
            if (P.blend_top >= 1) {
                g1s[k] = {fill(k, 1), gauss(k, 1), own(k, 1)};
                g1_tiles = std::max<long long>(g1_tiles, static_cast<long long>((F.gauss[1].w() + kG1W - 1) / kG1W) *
                                                             ((F.gauss[1].h() + kG1H - 1) / kG1H));
                for (int l = 2; l <= P.blend_top; ++l) {
                    gs[l * n + k] = {gauss(k, l - 1), own(k, l - 1), gauss(k, l), own(k, l), P.lw[l - 1], P.lh[l - 1]};
                    n_gauss[l] = std::max(n_gauss[l], F.gauss[l].n());
                }
            }
        }
    }
    if (hipMemcpyAsync(tables, host.data(), host.size(), hipMemcpyHostToDevice, stream) != hipSuccess)
        return Status::CUDA_ERROR;
    const auto* d_frames = reinterpret_cast<const Frames*>(dev(P.t.frames));
    const auto d_mix = [&](int l) { return reinterpret_cast<const MixLevel*>(dev(P.t.mix[l])); };
    // A launch over every frame: up to 2048 blocks per frame, striding over its window.
    const auto grid = [&](long long extent) { return dim3(grid_for(extent, 2048), static_cast<unsigned>(n)); };
    auto* owner = reinterpret_cast<std::int8_t*>(base + P.off_owner);
    auto* count = base + P.off_count;
    auto* stats = reinterpret_cast<unsigned long long*>(statistics);
    if (stats) chain_clear_stats<<<1, 64, 0, stream>>>(stats);
    // The warp is latency-bound per pixel: one pixel per thread where the
    // grid allows (65536 blocks over all frames).
    chain_warp<<<dim3(grid_for(n_warp, std::max(2048, 65536 / n)), static_cast<unsigned>(n)), kThreads, 0, stream>>>(
        d_frames, g[0].origin_x, g[0].origin_y);
    chain_own<<<grid_for(static_cast<long long>(W) * H, 256 * 64), kThreads, 0, stream>>>(
        d_frames, W, H, !bands, owner, count, output.data, output.pitch_bytes, mask.data, mask.pitch_bytes, stats);
    if (bands) {
        chain_pull1<<<dim3(static_cast<unsigned>(std::min<long long>(p1_tiles, 2048)), static_cast<unsigned>(n)),
                      kThreads, 0, stream>>>(d_frames, W, H, reinterpret_cast<const Plane*>(dev(P.t.dst1)));
        for (int l = 2; l <= P.top; ++l)
            chain_pull<<<grid(n_pull[l]), kThreads, 0, stream>>>(
                reinterpret_cast<const PullJob*>(dev(P.t.pull)) + static_cast<std::size_t>(l) * n);
        for (int l = P.top; l >= 1; --l)
            chain_push<<<grid(n_push[l]), kThreads, 0, stream>>>(
                reinterpret_cast<const PushJob*>(dev(P.t.push)) + static_cast<std::size_t>(l) * n, l == P.top);
        if (P.blend_top >= 1) {
            chain_gauss1<<<dim3(static_cast<unsigned>(std::min<long long>(g1_tiles, 2048)), static_cast<unsigned>(n)),
                           kThreads, 0, stream>>>(d_frames, reinterpret_cast<const Gauss1Job*>(dev(P.t.gauss1)), owner,
                                                  W, H, P.lw[1], P.lh[1]);
            for (int l = 2; l <= P.blend_top; ++l)
                chain_gauss<<<grid(n_gauss[l]), kThreads, 0, stream>>>(
                    reinterpret_cast<const GaussJob*>(dev(P.t.gauss)) + static_cast<std::size_t>(l) * n);
        }
        Plane z_up{};
        for (int l = P.blend_top; l >= 1; --l) {
            const Plane z{fptr(P.off_z[l]), {0, 0, P.lw[l], P.lh[l]}, 3};
            const bool coarse = l < P.blend_top;
            chain_mix<<<grid_for(static_cast<long long>(P.lw[l]) * P.lh[l]), kThreads, 0, stream>>>(
                d_mix(l), P.lw[l], P.lh[l], coarse ? P.lw[l + 1] : 1, coarse ? P.lh[l + 1] : 1, coarse, z,
                coarse ? z_up : z);
            z_up = z;
        }
        const Plane z1 = P.blend_top >= 1 ? z_up : Plane{fptr(P.off_count), {0, 0, 1, 1}, 1};
        chain_finish<<<grid_for(static_cast<long long>(W) * H), kThreads, 0, stream>>>(
            d_frames, P.blend_top >= 1 ? d_mix(1) : nullptr, owner, count, W, H, P.lw[1], P.lh[1], P.blend_top >= 1,
            z1, output.data, output.pitch_bytes);
    }
    return hipGetLastError() == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

}  // namespace stitch::s6
