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

__device__ __forceinline__ int reflect101(long long x, long long n) {
    if (n == 1) return 0;
    const long long period = 2 * n - 2;
    x %= period;
    if (x < 0) x += period;
    return static_cast<int>(x < n ? x : period - x);
}

__device__ __forceinline__ float tap(int x) {
    const int t = x < 0 ? -x : x;
    return t == 0 ? 0.375f : (t == 1 ? 0.25f : 0.0625f);
}

// Upsampled value of a coarse level at fine pixel (x, y): zero-insertion on a
// 2cw x 2ch grid, 5x5 filter, times 4 (s6.cu:133-146).
__device__ __forceinline__ void expanded3(const float* src, int cw, int ch, int x, int y, float out[3]) {
    float s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(static_cast<long long>(y) + dy, 2LL * ch);
        if (yy & 1) continue;
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101(static_cast<long long>(x) + dx, 2LL * cw);
            if (xx & 1) continue;
            const float w = tap(dx) * tap(dy);
            const float* p = src + (static_cast<size_t>(yy / 2) * cw + xx / 2) * 3;
            s0 += w * p[0];
            s1 += w * p[1];
            s2 += w * p[2];
        }
    }
    out[0] = 4 * s0;
    out[1] = 4 * s1;
    out[2] = 4 * s2;
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
    score = fmin(fmin((x + 0.5) / w, (w - 0.5 - x) / w), fmin((y + 0.5) / h, (h - 0.5 - y) / h));
    return true;
}

__device__ __forceinline__ std::uint8_t quantize(float x) {
    return static_cast<std::uint8_t>(floorf(fminf(255.0f, fmaxf(0.0f, 255.0f * x)) + 0.5f));
}

template <class Begin>
__device__ int find_job_by(const dev::JobSpan<S6Job>& span, long long t, Begin begin) {
    if (!span.table) return 0;
    int lo = 0, hi = span.count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin(span.table[mid]) <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void s6_clear_stats(dev::JobSpan<S6Job> span) {
    const S6Job& j = dev::job_at(span, blockIdx.x);
    if (threadIdx.x == 0 && j.stats) {
        j.stats->union_pixels = 0;
        j.stats->overlap_pixels = 0;
    }
}

// Canvas pixel -> A (integer shift) and B (FP64 inverse), level-0 buffers or
// the feather output, union mask, optional union/overlap counters.
__global__ void __launch_bounds__(kBlock) s6_warp(dev::JobSpan<S6Job> span, long long total) {
    __shared__ unsigned long long red[2][kBlock / 64];
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = find_job_by(span, t, [](const S6Job& j) { return j.warp_begin; });
        const S6Job& j = dev::job_at(span, ji);
        const long long p = (t - j.warp_begin) * kBlock + threadIdx.x;
        const long long n = static_cast<long long>(j.width) * j.height;
        bool ma = false, mb = false;
        if (p < n) {
            const int x = static_cast<int>(p % j.width), y = static_cast<int>(p / j.width);
            const double xa = static_cast<double>(x) + j.origin_x, ya = static_cast<double>(y) + j.origin_y;
            float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
            double sa = 0, sb = 0;
            ma = sample(j.a, j.wa, j.ha, j.pitch_a, xa, ya, ra, sa, j.snap_tolerance);
            const double* inv = j.inverse;
            const double den = inv[6] * xa + inv[7] * ya + inv[8];
            const double dsc = fabs(inv[6] * xa) + fabs(inv[7] * ya) + fabs(inv[8]);
            if (isfinite(den) && dsc > 0 && fabs(den) > j.denominator_tolerance * dsc) {
                const double xb = (inv[0] * xa + inv[1] * ya + inv[2]) / den;
                const double yb = (inv[3] * xa + inv[4] * ya + inv[5]) / den;
                mb = sample(j.b, j.wb, j.hb, j.pitch_b, xb, yb, rb, sb, j.snap_tolerance);
            }
            j.mask[static_cast<size_t>(y) * j.mask_pitch + x] = ma || mb ? 1 : 0;
            if (j.top >= 0 && j.mode != 0) {
                const S6Level& L = j.lv[0];
                L.va[p] = ma ? 1.0f : 0.0f;
                L.vb[p] = mb ? 1.0f : 0.0f;
                L.ua[p] = ma && (!mb || sa >= sb) ? 1.0f : 0.0f;
                L.ub[p] = mb && (!ma || sb > sa) ? 1.0f : 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    L.a[3 * p + c] = ra[c];
                    L.b[3 * p + c] = rb[c];
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
        }
        if (j.stats) {
            const unsigned long long bu = __ballot(ma || mb), bo = __ballot(ma && mb);
            const int lane = dev::lane_id(), wave = threadIdx.x / 64;
            if (lane == 0) {
                red[0][wave] = __popcll(bu);
                red[1][wave] = __popcll(bo);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                unsigned long long u = 0, o = 0;
                for (int w = 0; w < kBlock / 64; ++w) {
                    u += red[0][w];
                    o += red[1][w];
                }
                if (u) atomicAdd(reinterpret_cast<unsigned long long*>(&j.stats->union_pixels), u);
                if (o) atomicAdd(reinterpret_cast<unsigned long long*>(&j.stats->overlap_pixels), o);
            }
            __syncthreads();
        }
    }
}

// 5x5 filter + 2x decimation of `channels` interleaved channels (s6.cu:236-253).
__device__ __forceinline__ float reduce_at(const float* src, int w, int h, int x, int y, int channels, int c) {
    float v = 0;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(2LL * y + dy, h);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101(2LL * x + dx, w);
            v += tap(dx) * tap(dy) * src[(static_cast<size_t>(yy) * w + xx) * channels + c];
        }
    }
    return v;
}

// kind 0: pull (A, B, VA, VB); kind 1: Gaussians of the filled images and of
// the ownership (A, B, UA, UB). Destination level `l`.
__global__ void __launch_bounds__(kBlock) s6_reduce(dev::JobSpan<S6Job> span, long long total, int l, int kind) {
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = kind == 0 ? find_job_by(span, t, [l](const S6Job& j) { return j.pb[l]; })
                                 : find_job_by(span, t, [l](const S6Job& j) { return j.mb[l]; });
        const S6Job& j = dev::job_at(span, ji);
        const long long p = (t - (kind == 0 ? j.pb[l] : j.mb[l])) * kBlock + threadIdx.x;
        const S6Level& C = j.lv[l];
        const S6Level& F = j.lv[l - 1];
        if (p >= static_cast<long long>(C.w) * C.h) continue;
        const int x = static_cast<int>(p % C.w), y = static_cast<int>(p / C.w);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            C.a[3 * p + c] = reduce_at(F.a, F.w, F.h, x, y, 3, c);
            C.b[3 * p + c] = reduce_at(F.b, F.w, F.h, x, y, 3, c);
        }
        if (kind == 0) {
            C.va[p] = reduce_at(F.va, F.w, F.h, x, y, 1, 0);
            C.vb[p] = reduce_at(F.vb, F.w, F.h, x, y, 1, 0);
        } else {
            C.ua[p] = reduce_at(F.ua, F.w, F.h, x, y, 1, 0);
            C.ub[p] = reduce_at(F.ub, F.w, F.h, x, y, 1, 0);
        }
    }
}

// normalize_color + push_missing for level l (top-down): valid pixels keep
// numerator / validity, missing ones take the expanded, already final, coarse
// level (s6.cu:254-270, :399-410).
__global__ void __launch_bounds__(kBlock) s6_push(dev::JobSpan<S6Job> span, long long total, int l) {
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = find_job_by(span, t, [l](const S6Job& j) { return j.pb[l]; });
        const S6Job& j = dev::job_at(span, ji);
        const long long p = (t - j.pb[l]) * kBlock + threadIdx.x;
        const S6Level& L = j.lv[l];
        if (p >= static_cast<long long>(L.w) * L.h) continue;
        const int x = static_cast<int>(p % L.w), y = static_cast<int>(p / L.w);
        const bool coarse = l < j.top;
        const float wa = L.va[p], wb = L.vb[p];
        float e[3];
        if (wa > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) L.a[3 * p + c] = L.a[3 * p + c] / wa;
        } else if (coarse) {
            expanded3(j.lv[l + 1].a, j.lv[l + 1].w, j.lv[l + 1].h, x, y, e);
#pragma unroll
            for (int c = 0; c < 3; ++c) L.a[3 * p + c] = e[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) L.a[3 * p + c] = 0;
        }
        if (wb > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) L.b[3 * p + c] = L.b[3 * p + c] / wb;
        } else if (coarse) {
            expanded3(j.lv[l + 1].b, j.lv[l + 1].w, j.lv[l + 1].h, x, y, e);
#pragma unroll
            for (int c = 0; c < 3; ++c) L.b[3 * p + c] = e[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) L.b[3 * p + c] = 0;
        }
    }
}

// mix_band + reconstruct (s6.cu:271-293) for level l, top-down; at level 0
// also finish (s6.cu:294-306): single-source pixels keep their warped colour.
__global__ void __launch_bounds__(kBlock) s6_mix(dev::JobSpan<S6Job> span, long long total, int l) {
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        const int ji = find_job_by(span, t, [l](const S6Job& j) { return j.mb[l]; });
        const S6Job& j = dev::job_at(span, ji);
        const long long p = (t - j.mb[l]) * kBlock + threadIdx.x;
        const S6Level& L = j.lv[l];
        if (p >= static_cast<long long>(L.w) * L.h) continue;
        const int x = static_cast<int>(p % L.w), y = static_cast<int>(p / L.w);
        const bool coarse = l < j.blend_top;
        float la[3], lb[3], zc[3] = {0, 0, 0};
        if (coarse) {
            const S6Level& C = j.lv[l + 1];
            float ea[3], eb[3];
            expanded3(C.a, C.w, C.h, x, y, ea);
            expanded3(C.b, C.w, C.h, x, y, eb);
            expanded3(C.z, C.w, C.h, x, y, zc);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                la[c] = L.a[3 * p + c] - ea[c];
                lb[c] = L.b[3 * p + c] - eb[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                la[c] = L.a[3 * p + c] - 0.0f;
                lb[c] = L.b[3 * p + c] - 0.0f;
            }
        }
        const float ua = L.ua[p], ub = L.ub[p];
        const float tot = ua + ub;
        float z[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float m = tot > 0 ? (ua * la[c] + ub * lb[c]) / tot : 0;
            z[c] = coarse ? m + zc[c] : m;
        }
        if (l > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) L.z[3 * p + c] = z[c];
        } else {
            const float va = L.va[p], vb = L.vb[p];
            std::uint8_t* out = j.out + static_cast<size_t>(y) * j.out_pitch + 3 * x;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float color = va > 0 && vb > 0 ? z[c] : (va > 0 ? L.a[3 * p + c] : (vb > 0 ? L.b[3 * p + c] : 0));
                out[c] = quantize(color);
            }
        }
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
        add(n, 1, o.va);
        add(n, 1, o.vb);
        if (l <= requested) {
            add(n, 1, o.ua);
            add(n, 1, o.ub);
            if (l > 0) add(n, 3, o.z);
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
            L.a = f(o.a);
            L.b = f(o.b);
            L.va = f(o.va);
            L.vb = f(o.vb);
            if (l <= layout.blend_top) {
                L.ua = f(o.ua);
                L.ub = f(o.ub);
                L.z = l > 0 ? f(o.z) : nullptr;
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
        t.warp += (static_cast<long long>(j.width) * j.height + kBlock - 1) / kBlock;
        for (int l = 0; l < kS6MaxLevels; ++l) {
            j.pb[l] = t.pull[l];
            j.mb[l] = t.mix[l];
            if (j.top >= l) {
                const long long tiles = (static_cast<long long>(j.lv[l].w) * j.lv[l].h + kBlock - 1) / kBlock;
                t.pull[l] += tiles;
                if (j.blend_top >= l) t.mix[l] += tiles;
            }
        }
        t.top = std::max(t.top, j.top);
        t.blend_top = std::max(t.blend_top, j.blend_top);
    }
    return t;
}

hipError_t launch_s6(const S6Job* table, int count, const S6Totals& t, const S6Job* single, hipStream_t stream) {
    dev::JobSpan<S6Job> span{table, count, single ? *single : S6Job{}};
    bool any_stats = single ? single->stats != nullptr : true;
    if (any_stats) s6_clear_stats<<<count, 64, 0, stream>>>(span);
    if (t.warp > 0) s6_warp<<<persistent_grid(t.warp), kBlock, 0, stream>>>(span, t.warp);
    for (int l = 1; l <= t.top; ++l)
        if (t.pull[l] > 0) s6_reduce<<<persistent_grid(t.pull[l]), kBlock, 0, stream>>>(span, t.pull[l], l, 0);
    for (int l = t.top; l >= 0; --l)
        if (t.pull[l] > 0) s6_push<<<persistent_grid(t.pull[l]), kBlock, 0, stream>>>(span, t.pull[l], l);
    for (int l = 1; l <= t.blend_top; ++l)
        if (t.mix[l] > 0) s6_reduce<<<persistent_grid(t.mix[l]), kBlock, 0, stream>>>(span, t.mix[l], l, 1);
    for (int l = t.blend_top; l >= 0; --l)
        if (t.mix[l] > 0) s6_mix<<<persistent_grid(t.mix[l]), kBlock, 0, stream>>>(span, t.mix[l], l);
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
