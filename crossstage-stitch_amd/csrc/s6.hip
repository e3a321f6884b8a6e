// S6: perspective warp of A and B onto the union canvas with feather or
// Laplacian multiband blending, for any number of canvases per launch.
// Semantics: src/s6/s6.cu:60-433 - FP64 inverse map and bilinear sampling
// (FP64 -> FP32 / 255), border-distance weights, ownership A if s_a >= s_b,
// pull-push extension to 1x1, [1 4 6 4 1]/16 pyramids with reflect-101,
// band mix by blurred ownership, collapse, floor(clamp(255 x) + 0.5).
// The depth is generic: 3, 5, 6 or 7 reductions. Filter taps, expansions and
// the band mix accumulate with fused multiply-add.
//
// MI355X layout: the full-resolution level is the expensive one, so it is
// never materialised as A colours (A is read from its source bytes through
// a 256-entry table) and never filled in memory: the Gaussian into level 1
// is the pull's raw sums plus the filter of the fill, which is a stencil of
// the pushed level 1 (s6_gauss1). Level 0 keeps one flag byte per pixel and
// B's warped colour; the final band mix writes the panorama and the union
// mask. Levels >= 1 are FP32 planes. One launch per level covers every
// canvas of a batch;
// normalisation is fused into the push pass and band mixing into the
// collapse. Expansions visit only the taps whose reflected fine position has
// the pixel's parity (reflect-101 keeps parity), 9 instead of 25.
#include <hip/hip_runtime.h>

#include <type_traits>

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

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ bool center_coordinate(double& x, int n, double tol) {
    if (!isfinite(x)) return false;
    const double nearest = nearbyint(x);
    const double roundoff = 16 * 2.2204460492503131e-16 * fmax(1.0, fabs(x));
    if (fabs(x - nearest) <= fmax(tol, roundoff)) x = nearest;
    return x >= 0 && x <= n - 1.0;
}

// center_coordinate when the snap threshold max(tol, 16 eps max(1, |x|)) is
// tol for every |x| <= n (tol >= 16 eps max(1, n), checked per job): then
// the two agree for |x| <= n, and for larger |x| (or an infinite one) both
// call the sample invalid - snapping moves x by at most tol < 1/2, and
// 16 eps |x| + tol < 1. A NaN fails every comparison, so needs no test.
__device__ __forceinline__ bool center_coordinate_tol(double& x, int n, double tol) {
    const double nearest = rint(x);
    if (fabs(x - nearest) <= tol) x = nearest;
    return x >= 0 && x <= n - 1.0;
}

// Bytes p[0 .. n) of an RGB row, n = 6 (two neighbouring pixels) or 3 (one),
// through aligned dword loads: lo holds p[0..3], hi p[4..7] (bytes past n
// undefined). A dword that holds no needed byte is not read: its load reads
// the first dword again instead, so the loads stay unconditional (no branch
// between them) and every address read holds a byte of the image - an
// aligned dword holding any byte of a buffer lies in that buffer's page.
// kSkip: dwords holding no needed byte are skipped by a branch instead (fewer
// memory accesses where the loads are not batched with others).
template <bool kSkip = false>
__device__ __forceinline__ void load_rgb(const std::uint8_t* p, bool two, unsigned& lo, unsigned& hi) {
    // Pointer arithmetic (not an integer round trip) keeps the global
    // address space known, so these are global loads, not flat ones.
    const unsigned sh = static_cast<unsigned>(reinterpret_cast<std::uintptr_t>(p) & 3u);
    const unsigned* a = reinterpret_cast<const unsigned*>(p - sh);
    const unsigned last = sh + (two ? 5u : 2u);  // offset of the last needed byte from a
    const unsigned w0 = a[0];
    unsigned w1, w2;
    if (kSkip) {
        w1 = w2 = 0;
        if (last >= 4) w1 = a[1];
        if (last >= 8) w2 = a[2];
    } else {
        w1 = a[last >= 4 ? 1 : 0];
        w2 = two ? a[last >= 8 ? 2 : 0] : w1;
    }
    lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
    hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
}

__device__ __forceinline__ unsigned byte_of(unsigned w, int k) { return (w >> (8 * k)) & 0xffu; }

// Bilinear B sample; the border score only when asked. The position, its
// snapping and the validity are FP64 as in the reference; the blend of the
// four bytes runs in FP32 with the FP64 fractions rounded to FP32, then
// scales by 1/255: the colour (FP32 either way) moves by a few ulps, far
// below what the 8-bit output can show.
// x and y come back snapped, as the border score needs them.
__device__ bool sample(const std::uint8_t* img, int w, int h, int pitch, double& x, double& y, float* rgb, double* score,
                       double tol, float* raw = nullptr, bool tol_snap = false) {
    if (tol_snap ? !center_coordinate_tol(x, w, tol) || !center_coordinate_tol(y, h, tol)
                 : !center_coordinate(x, w, tol) || !center_coordinate(y, h, tol))
        return false;
    const int x0 = static_cast<int>(floor(x)), y0 = static_cast<int>(floor(y));
    const int x1 = x0 < w - 1 ? x0 + 1 : x0, y1 = y0 < h - 1 ? y0 + 1 : y0;
    const float ax = static_cast<float>(x - x0), ay = static_cast<float>(y - y0);
    // Row bytes of pixels x0 and x1 (x1 = x0 at the right border).
    const bool two = x1 != x0;
    unsigned lo0, hi0, lo1, hi1;
    // x0, y0 >= 0 here and every image is below 4 GiB: 32-bit offsets.
    const unsigned cx = 3u * static_cast<unsigned>(x0);
    load_rgb<true>(img + (static_cast<unsigned>(y0) * static_cast<unsigned>(pitch) + cx), two, lo0, hi0);
    load_rgb<true>(img + (static_cast<unsigned>(y1) * static_cast<unsigned>(pitch) + cx), two, lo1, hi1);
    const unsigned q00[3] = {byte_of(lo0, 0), byte_of(lo0, 1), byte_of(lo0, 2)};
    const unsigned q10[3] = {byte_of(lo1, 0), byte_of(lo1, 1), byte_of(lo1, 2)};
    const unsigned q01[3] = {two ? byte_of(lo0, 3) : q00[0], two ? byte_of(hi0, 0) : q00[1], two ? byte_of(hi0, 1) : q00[2]};
    const unsigned q11[3] = {two ? byte_of(lo1, 3) : q10[0], two ? byte_of(hi1, 0) : q10[1], two ? byte_of(hi1, 1) : q10[2]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float v0 = fmaf(ax, static_cast<float>(q01[c]), (1 - ax) * static_cast<float>(q00[c]));
        const float v1 = fmaf(ax, static_cast<float>(q11[c]), (1 - ax) * static_cast<float>(q10[c]));
        const float v = fmaf(ay, v1, (1 - ay) * v0);
        if (raw) raw[c] = v;
        rgb[c] = v * (1.0f / 255.0f);
    }
    // min((x+.5)/w, (w-.5-x)/w) == min(x+.5, w-.5-x)/w exactly: a correctly
    // rounded division by a positive w is monotonic, so it commutes with min.
    if (score) *score = fmin(fmin(x + 0.5, w - 0.5 - x) / w, fmin(y + 0.5, h - 0.5 - y) / h);
    return true;
}

__device__ __forceinline__ std::uint8_t quantize(float x) {
    return static_cast<std::uint8_t>(floorf(fminf(255.0f, fmaxf(0.0f, 255.0f * x)) + 0.5f));
}

// B's warped colour at level 0, stored for the pull into level 1 and the
// level-0 band mix: the bilinear values v in [0, 255] of R, G and B as 8.13
// fixed point in 21-bit fields of one 8-byte word (8 B per pixel instead of
// three FP32 planes, 12 B). The stored colour is within 2^-14 of a grey level
// of the computed one, about the size of the pyramid's own FP32 rounding; the
// warp's own single-source output uses the exact value.
// The fields are rounded half to even, as rintf rounds: v * 8192 is exact,
// and adding 2^23 rounds it to an integer held in the low mantissa bits
// (v * 8192 < 2^21). The word is assembled from 32-bit halves.
__device__ __forceinline__ unsigned long long pack_b(const float v[3]) {
    unsigned f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = __builtin_bit_cast(unsigned, fmaf(v[c], 8192.0f, 8388608.0f)) - 0x4b000000u;
    const unsigned lo = f[0] | (f[1] << 21), hi = (f[1] >> 11) | (f[2] << 10);
    return static_cast<unsigned long long>(hi) << 32 | lo;
}
__device__ __forceinline__ void unpack_b(unsigned long long w, float rgb[3]) {
    const unsigned lo = static_cast<unsigned>(w), hi = static_cast<unsigned>(w >> 32);
    const unsigned f[3] = {lo & 0x1fffffu, __builtin_amdgcn_alignbit(hi, lo, 21) & 0x1fffffu, hi >> 10};
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = static_cast<float>(f[c]) * (1.0f / 8192.0f) * (1.0f / 255.0f);
}

// Incremental form for tiles visited in increasing order (see dev::next_job).
template <class Begin>
__device__ __forceinline__ int next_job_by(const dev::JobSpan<S6Job>& span, long long t, Begin begin, int ji,
                                           int jobs) {
    if (!span.table) return 0;
    return dev::next_job_at(t, [&](int i) { return dev::ld_const(&begin(span.table[i])); }, ji, jobs);
}

constexpr unsigned kVA = 1u, kVB = 2u, kUA = 4u, kUB = 8u;

// float(v / 255.0) for byte k of w: A's level-0 colour. q = v * RN(1/255)
// is within an ulp of v / 255; one correction with the residual v - 255 q
// (exact through a fused multiply-add) rounds it correctly, which an
// exhaustive check over the 256 bytes confirms. Four VALU instructions and
// no table, so no LDS bank conflicts between lanes looking up different bytes.
__device__ __forceinline__ float unit_byte(unsigned w, int k) {
    const float v = static_cast<float>((w >> (8 * k)) & 0xffu);
    constexpr float kInv = 1.0f / 255.0f;
    const float q = v * kInv;
    return fmaf(fmaf(-q, 255.0f, v), kInv, q);
}

__device__ __forceinline__ const std::uint8_t* a_pixel(const S6Job& j, int x, int y) {
    // A's pixel lies inside A here; every image is below 4 GiB.
    return j.a + (static_cast<unsigned>(y + j.origin_y) * static_cast<unsigned>(j.pitch_a) +
                  3u * static_cast<unsigned>(x + j.origin_x));
}

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

// ------------------------------------------------------------------ warp
// Canvas pixel -> A (integer shift) and B (FP64 inverse map, only inside the
// box that can hold B). Multiband: level-0 flags and B's colour where valid.
// Feather: the output and the union mask. Union/overlap counters for both.
// The border scores are only needed where both images are valid.
// Terms of a canvas row shared by its pixels: the y products of the inverse
// map's three rows and A's vertical border score. Each is the same rounded
// value the per-pixel expression would compute (products are rounded on
// their own, and the sums keep the reference's left-to-right order).
struct WarpRow {
    double ya, ty0, ty1, ty2, sya;
};

__device__ __forceinline__ WarpRow warp_row(const S6Job& j, int y) {
    const double ya = static_cast<double>(y + j.origin_y);
    return {ya, j.inverse[1] * ya, j.inverse[4] * ya, j.inverse[7] * ya, fmin(ya + 0.5, j.ha - 0.5 - ya) / j.ha};
}

// Whether A owns an overlap pixel: its border score sa is at least B's, sb
// (s6.cu:206-207). The scores are min(d / n) over the axes (s6.cu:170), d
// the distance to the nearer border plus a half and n the image size, each
// a correctly rounded quotient. Products by the rounded reciprocals 1 / n
// are each within 3 * 2^-53 of their quotient (relative: two roundings of
// the product and one of the quotient), and so is the minimum; when the two
// approximate scores differ by more than 2^-48 of the larger, the exact
// ones differ with the same sign. Only nearer calls (ties among them) take
// the quotients themselves.
__device__ __forceinline__ bool a_owns(const S6Job& j, const WarpRow& r, double xa, double xb, double yb) {
    const double qa = fmin(fmin(xa + 0.5, j.wa - 0.5 - xa) * j.inv_wa, r.sya);
    const double qb = fmin(fmin(xb + 0.5, j.wb - 0.5 - xb) * j.inv_wb, fmin(yb + 0.5, j.hb - 0.5 - yb) * j.inv_hb);
    const double d = qa - qb;
    if (fabs(d) > 0x1p-48 * fmax(qa, qb)) return d > 0;
    const double sa = fmin(fmin(xa + 0.5, j.wa - 0.5 - xa) / j.wa, r.sya);
    const double sb = fmin(fmin(xb + 0.5, j.wb - 0.5 - xb) / j.wb, fmin(yb + 0.5, j.hb - 0.5 - yb) / j.hb);
    return sa >= sb;
}

__device__ __forceinline__ void warp_pixel(const S6Job& j, const WarpRow& r, int x, int y, unsigned& uni,
                                           unsigned& ovl) {
    // Every plane and image is below 4 GiB: 32-bit offsets from their bases.
    const unsigned p = static_cast<unsigned>(y) * static_cast<unsigned>(j.width) + static_cast<unsigned>(x);
    const unsigned pm = static_cast<unsigned>(y) * static_cast<unsigned>(j.mask_pitch) + static_cast<unsigned>(x);
    const unsigned po = static_cast<unsigned>(y) * static_cast<unsigned>(j.out_pitch) + 3u * static_cast<unsigned>(x);
    const int ax = x + j.origin_x, ay = y + j.origin_y;
    const double xa = static_cast<double>(ax), ya = r.ya;
    float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
    const bool ma = ax >= 0 && ax <= j.wa - 1 && ay >= 0 && ay <= j.ha - 1;
    bool mb = false;
    double xb = 0, yb = 0;
    if (x >= j.bx0 && x <= j.bx1 && y >= j.by0 && y <= j.by1) {
        const double* inv = j.inverse;
        const double den = inv[6] * xa + r.ty2 + inv[8];
        const double dsc = fabs(inv[6] * xa) + fabs(r.ty2) + fabs(inv[8]);
        if (isfinite(den) && dsc > 0 && fabs(den) > j.denominator_tolerance * dsc) {
            xb = (inv[0] * xa + r.ty0 + inv[2]) / den;
            yb = (inv[3] * xa + r.ty1 + inv[5]) / den;
            mb = sample(j.b, j.wb, j.hb, j.pitch_b, xb, yb, rb, nullptr, j.snap_tolerance, vb, j.tol_snap);
        }
    }
    if (j.top >= 0) {
        const S6Level& L = j.lv[0];
        const bool owns = !(ma && mb) || a_owns(j, r, xa, xb, yb);
        unsigned f = 0;
        if (ma) f |= kVA;
        if (mb) f |= kVB;
        if (ma && (!mb || owns)) f |= kUA;
        if (mb && (!ma || !owns)) f |= kUB;
        L.flags[p] = static_cast<std::uint8_t>(f);
        if (mb) reinterpret_cast<unsigned long long*>(L.b)[p] = pack_b(vb);
        // Single-source pixels keep their warped colour (s6.cu:294-306):
        // written here; the band mix writes the overlap.
        j.mask[pm] = ma || mb ? 1 : 0;
        if (!(ma && mb)) {
            std::uint8_t* out = j.out + po;
            if (ma) {
                // A's level-0 colour is float(v / 255), which quantises back
                // to v for every byte v: the output is A's bytes.
                unsigned lo, hi;
                load_rgb<true>(a_pixel(j, x, y), false, lo, hi);
#pragma unroll
                for (int c = 0; c < 3; ++c) out[c] = static_cast<std::uint8_t>(byte_of(lo, c));
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) out[c] = quantize(rb[c]);
            }
        }
    } else {
        double sa = 0, sb = 0;
        if (ma && mb) {
            sa = fmin(fmin(xa + 0.5, j.wa - 0.5 - xa) / j.wa, r.sya);
            sb = fmin(fmin(xb + 0.5, j.wb - 0.5 - xb) / j.wb, fmin(yb + 0.5, j.hb - 0.5 - yb) / j.hb);
        }
        if (ma) {
            unsigned lo, hi;
            load_rgb<true>(a_pixel(j, x, y), false, lo, hi);
            ra[0] = unit_byte(lo, 0);
            ra[1] = unit_byte(lo, 1);
            ra[2] = unit_byte(lo, 2);
        }
        std::uint8_t* out = j.out + po;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float color =
                ma && mb ? static_cast<float>((sa * ra[c] + sb * rb[c]) / (sa + sb)) : (ma ? ra[c] : rb[c]);
            out[c] = quantize(color);
        }
        j.mask[pm] = ma || mb ? 1 : 0;
    }
    uni += ma || mb;
    ovl += ma && mb;
}

// Tiles are (row, 1024-pixel segment); each thread takes four pixels 256 apart.
constexpr int kWarpSeg = 4 * kBlock;

__global__ void __launch_bounds__(kBlock) s6_warp(dev::JobSpan<S6Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ unsigned long long red[2][kBlock / 64];
    int cur = -1, ji = 0;
    const int njobs = span.jobs();
    // A thread counts at most four pixels per tile of one canvas: 32 bits.
    unsigned uni = 0, ovl = 0;
    for (dev::TileWalk w = dev::tile_walk(total); w.t < w.end; w.t += w.step) {
        const long long t = w.t;
        ji = next_job_by(span, t, [](const S6Job& j) -> const long long& { return j.warp_begin; }, ji, njobs);
        if (ji != cur) {
            if (cur >= 0 && dev::job_at(span, cur).stats) flush_stats(dev::job_at(span, cur), uni, ovl, red);
            cur = ji;
            uni = ovl = 0;
        }
        const S6Job& j = dev::job_at(span, ji);
        const unsigned local = static_cast<unsigned>(t - j.warp_begin);
        const unsigned segs = (static_cast<unsigned>(j.width) + kWarpSeg - 1) / kWarpSeg;
        const unsigned row = local / segs;
        const int x0 = static_cast<int>((local - row * segs) * kWarpSeg + threadIdx.x);
        const WarpRow wr = warp_row(j, static_cast<int>(row));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = x0 + k * kBlock;
            if (x < j.width) warp_pixel(j, wr, x, static_cast<int>(row), uni, ovl);
        }
    }
    if (cur >= 0 && dev::job_at(span, cur).stats) flush_stats(dev::job_at(span, cur), uni, ovl, red);
}

// ------------------------------------------------------------- expansion
// 4 * sum of tap(dx) * tap(dy) * coarse over the taps whose reflected fine
// position is even (s6.cu's expanded(), in its dy-major, dx-minor order).
// Reflect-101 keeps parity, so those are the offsets with the parity of x
// (and y): three for an even coordinate, two for an odd one, whose third
// slot carries weight 0 - adding 0 * v leaves the sum unchanged, so every
// lane runs the same 9 steps.
struct Expand {
    int cx[3], cy[3];    // LDS column / row per slot
    float wx[3], wy[3];
};

__device__ __forceinline__ void expand_axis(int x, int cn, int c0, int cmax, int (&ci)[3], float (&w)[3]) {
    const bool odd = x & 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int d = odd ? (k == 0 ? -1 : 1) : 2 * k - 2;
        const int r = reflect101(x + d, 2 * cn);
        ci[k] = clampi((r >> 1) - c0, 0, cmax);
        w[k] = (odd && k == 2) ? 0.0f : tap(d);
    }
}

__device__ __forceinline__ void expand_setup(int x, int y, int cw, int ch, int cx0, int cy0, int cols, int rows,
                                             Expand& e) {
    expand_axis(x, cw, cx0, cols - 1, e.cx, e.wx);
    expand_axis(y, ch, cy0, rows - 1, e.cy, e.wy);
}

// RGB of one coarse pixel per LDS slot (row pitch P slots), so a tap is one
// 16-byte LDS read for the three channels; each channel keeps its own sum
// in the reference's order.
template <int P>
__device__ __forceinline__ void expand_rgb(const float4* src, const Expand& e, float out[3]) {
    float s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float w = e.wx[k] * e.wy[j];
            const float4 v = src[e.cy[j] * P + e.cx[k]];
            s0 = fmaf(w, v.x, s0);
            s1 = fmaf(w, v.y, s1);
            s2 = fmaf(w, v.z, s2);
        }
    out[0] = 4 * s0;
    out[1] = 4 * s1;
    out[2] = 4 * s2;
}

// Stages rows [cy0, cy0 + rows) x cols [cx0, cx0 + cols) of the R, G, B
// planes (n floats apart) of a coarse level, clamped to its extent.
template <int P>
__device__ __forceinline__ void stage_rgb(const float* src, std::size_t n, int cw, int ch, int cx0, int cy0, int cols,
                                          int rows, float4* dst) {
    const float* r = src;
    const float* g = src + n;
    const float* b = src + 2 * n;
    for (int i = threadIdx.x; i < cols * rows; i += blockDim.x) {
        const int cx = clampi(cx0 + i % cols, 0, cw - 1), cy = clampi(cy0 + i / cols, 0, ch - 1);
        const unsigned q = static_cast<unsigned>(cy) * static_cast<unsigned>(cw) + static_cast<unsigned>(cx);
        dst[(i / cols) * P + i % cols] = make_float4(r[q], g[q], b[q], 0.0f);
    }
}

// The same staging split in two: load() issues this thread's loads into
// registers, store() writes them to LDS, so a kernel can issue every load of
// a tile before its first barrier (one memory round trip per tile).
template <int P, int Cols, int Rows>
struct StageRGB {
    static constexpr int kN = Cols * Rows, kIt = (kN + kBlock - 1) / kBlock;
    float v[kIt][3];
    __device__ __forceinline__ void load(const float* src, std::size_t n, int cw, int ch, int cx0, int cy0) {
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int i = static_cast<int>(threadIdx.x) + it * kBlock;
            if (i >= kN) continue;
            const int cx = clampi(cx0 + i % Cols, 0, cw - 1), cy = clampi(cy0 + i / Cols, 0, ch - 1);
            const std::size_t q = static_cast<std::size_t>(static_cast<unsigned>(cy) * static_cast<unsigned>(cw) +
                                                           static_cast<unsigned>(cx));
            v[it][0] = src[q];
            v[it][1] = src[n + q];
            v[it][2] = src[2 * n + q];
        }
    }
    __device__ __forceinline__ void store(float4* dst) const {
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int i = static_cast<int>(threadIdx.x) + it * kBlock;
            if (i < kN) dst[(i / Cols) * P + i % Cols] = make_float4(v[it][0], v[it][1], v[it][2], 0.0f);
        }
    }
};

// ---------------------------------------------------------------- reduce
// 5x5 [1 4 6 4 1]/16 filter + 2x decimation (s6.cu:236-253) on LDS tiles:
// a 32x8 tile of level-l outputs reads a 67x19 region of level l-1, loaded
// once with the reflect-101 borders applied at load time; each output keeps
// the reference's dy-major, dx-minor sum per channel. Tile rows hold the
// even columns first, then the odd ones, so the stride-2 reads of
// neighbouring lanes land in consecutive banks; a row pitch of 80 words puts
// the wave's two output rows 32 banks apart.
constexpr int kRW = 32, kRH = 8, kRIW = 2 * kRW + 3, kRIH = 2 * kRH + 3;
constexpr int kREven = (kRIW + 1) / 2, kRPitch = 80;
static_assert(kRIW <= kRPitch && (2 * kRPitch) % 64 == 32, "reduce tile pitch");

// Integer [1 4 6 4 1] taps.
__device__ __forceinline__ int itap(int d) { return d == 0 ? 6 : (d == 1 || d == -1 ? 4 : 1); }

// Loads and stores at a 32-bit byte offset from a launch-uniform base: the
// address is then one scalar base plus one vector offset (no 64-bit address
// arithmetic per access). Every plane and image here is below 4 GiB.
template <class T>
__device__ __forceinline__ T ld_at(const T* base, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ void st_at(T* base, unsigned off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}

// Whether pixel (x, y) of level l lies in image i's support (A 0, B 1):
// outside it the image's raw pull sums and weights are exactly 0 and never
// stored, so readers take 0 there instead of loading.
__device__ __forceinline__ bool in_support(const S6Job& j, int l, int i, int x, int y) {
    const int(&s)[4] = j.sup[l][i];
    return x >= s[0] && x <= s[2] && y >= s[1] && y <= s[3];
}

// Raw buffer resource over [p, p + bytes): loads at a byte offset at or past
// `bytes` return 0 without touching memory (the range check of buffer
// instructions), so a load that only some lanes need takes no branch.
constexpr unsigned kNoLoad = 0xfffffffcu;  // an offset past every buffer here
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, static_cast<int>(bytes), 0x00020000);
}
__device__ __forceinline__ float ld_buf(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
// Store through a buffer resource: at an offset at or past its size the
// store is dropped (no branch around stores that only some lanes make).
__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// Level-1 validity and ownership weights are multiples of 1/256 in [0, 1]
// (sums of tap products of 0/1 flags), stored exactly as 256 w in 16 bits.
__device__ __forceinline__ float w16(unsigned k) { return static_cast<float>(k) * (1.0f / 256.0f); }
__device__ __forceinline__ float ld_w16(const float* plane, unsigned i) {
    return w16(reinterpret_cast<const std::uint16_t*>(plane)[i]);
}
__device__ __forceinline__ float ld_buf_w16(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return w16(__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
}
__device__ __forceinline__ void st_buf_w16(__amdgpu_buffer_rsrc_t r, unsigned off, int k) {
    __builtin_amdgcn_raw_buffer_store_b16(static_cast<unsigned short>(k), r, off, 0, 0);
}
__device__ __forceinline__ unsigned half_off(unsigned off4) { return off4 == kNoLoad ? kNoLoad : off4 / 2; }

// A's level-1 raw pull sums: sum over the 5x5 footprint of the integer tap
// products (1 .. 36, total 256) times A's bytes where A is valid, an integer
// N <= 65280 that the pull forms exactly (every partial sum is an integer
// below 2^24) and stores in 16 bits; the sum is N / (256 * 255).
constexpr float kInvA1 = 1.0f / 65280.0f;
__device__ __forceinline__ float a1(unsigned n) { return static_cast<float>(n) * kInvA1; }

// Modes: 0 pull at l >= 2, 1 Gaussian at l >= 2, 2 pull at l = 1 (from the
// virtual level 0). The Gaussian at l = 1 is s6_gauss1. From level 0 the
// tile holds the six colour channels; the four level-1 weights (validity and
// ownership of A and B) are integer sums of the flag bits, exact (every
// partial sum of tap products is a multiple of 1/256), and the ownership
// weights are the Gaussian's own at level 1.
template <int kMode, bool kW16 = false>
__global__ void __launch_bounds__(kBlock, 3) s6_reduce(dev::JobSpan<S6Job> span, dev::Total work, int l) {
    static_assert(kMode >= 0 && kMode <= 2, "reduce mode");
    // kW16: the finer level is level 1, whose weights are 16-bit.
    const long long total = work.get();
    constexpr bool kFromZero = kMode == 2;
    constexpr int kCh = kFromZero ? 6 : 8;
    // Level 0 stages A's bytes as two integer planes (R | B << 16 and G) and
    // B's colour as three FP32 planes. Level 1 under a pull stages A's sums
    // (16-bit integers N) and the two validity weights (16-bit 256 w) as
    // 16-bit planes, B's sums as FP32; the Gaussians above stage FP32.
    constexpr bool kInt16 = kMode == 0 && kW16;
    constexpr int kPl = kFromZero ? 5 : (kInt16 ? 3 : kCh);
    __shared__ float tile[kPl][kRIH][kRPitch];
    __shared__ std::uint16_t t16[kInt16 ? 5 : 1][kRIH][kRPitch];
    __shared__ std::uint8_t fl[kFromZero ? kRIH : 1][kRIW + 1];
    __shared__ unsigned wave_flags[2][kBlock / 64];  // per wave: OR and AND of the staged flags
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk w = dev::tile_walk(total); w.t < w.end; w.t += w.step) {
        const long long t = w.t;
        ji = kMode != 1 ? next_job_by(span, t, [l](const S6Job& j) -> const long long& { return j.rb[l]; }, ji, njobs)
                        : next_job_by(span, t, [l](const S6Job& j) -> const long long& { return j.gb[l]; }, ji, njobs);
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& C = j.lv[l];
        const S6Level& F = j.lv[l - 1];
        const unsigned local = static_cast<unsigned>(t - (kMode != 1 ? j.rb[l] : j.gb[l]));
        // Pulls cover the region whose pulled values are read, the Gaussian
        // its dependency region.
        const int(&box)[4] = kMode == 1 ? j.gbox[l] : j.ubox[l];
        const int bx = box[0], by = box[1], bw = box[2] - bx + 1;
        const unsigned tx = (static_cast<unsigned>(bw) + kRW - 1) / kRW;
        const int x0 = bx + static_cast<int>(local % tx) * kRW, y0 = by + static_cast<int>(local / tx) * kRH;
        const int ix0 = 2 * x0 - 2, iy0 = 2 * y0 - 2;
        // Source planes: A, B colour and the two weights (level 0: B only).
        const float* planes[kCh];
        {
            // Pull reads the finer level's raw sums (a / b), the Gaussian its
            // Gaussian (ga / gb, which alias a / b above level 1).
            const float* sa = kMode == 1 ? F.ga : F.a;
            const float* sb = kMode == 1 ? F.gb : F.b;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                planes[c] = kFromZero ? nullptr : sa + c * F.n;
                planes[3 + c] = sb + c * F.n;
            }
            if (!kFromZero) {
                planes[kCh - 2] = kMode == 0 ? F.va : F.ua;
                planes[kCh - 1] = kMode == 0 ? F.vb : F.ub;
            }
        }
        // A pull above level 1 reads an image's raw sums only inside its
        // support at level l - 1 (0 elsewhere, where they are not stored),
        // through buffer resources whose range check returns the 0.
        // Elsewhere the load reads the nearest pixel of the support (written
        // by this render, so the address is in use anyway) and the value is
        // dropped: plain loads with no branch.
        int sa0 = 0, sa1 = 0, sa2 = -1, sa3 = -1, sb0 = 0, sb1 = 0, sb2 = -1, sb3 = -1;
        // Supports at level l, where the pull stores.
        int oa0 = 0, oa1 = 0, oa2 = -1, oa3 = -1, ob0 = 0, ob1 = 0, ob2 = -1, ob3 = -1;
        if (kMode != 1) {
            const int(&ka)[4] = j.sup[l][0];
            const int(&kb)[4] = j.sup[l][1];
            oa0 = ka[0], oa1 = ka[1], oa2 = ka[2], oa3 = ka[3];
            ob0 = kb[0], ob1 = kb[1], ob2 = kb[2], ob3 = kb[3];
        }
        if (kMode == 0) {
            const int(&ka)[4] = j.sup[l - 1][0];
            const int(&kb)[4] = j.sup[l - 1][1];
            sa0 = ka[0], sa1 = ka[1], sa2 = ka[2], sa3 = ka[3];
            sb0 = kb[0], sb1 = kb[1], sb2 = kb[2], sb3 = kb[3];
        }
        // Level 1 below a pull into level 2: A's sums are 16-bit planes.
        const std::uint16_t* a16[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) a16[c] = reinterpret_cast<const std::uint16_t*>(F.a) + c * F.n;
        // Pass 1: every load of this thread's tile elements, issued before any
        // is used (positions are clamped into the level, so all are in
        // bounds; level-0 values of invalid pixels are discarded below).
        constexpr int kElems = kRIW * kRIH, kPer = (kElems + kBlock - 1) / kBlock;
        int sxs[kPer], sys[kPer];
        unsigned fls[kPer], abytes[kPer];
        unsigned long long bpk[kFromZero ? kPer : 1];  // level 0: B's packed colour
        bool ins[kPer][2];  // pulls above level 1: element inside A's / B's support
        float raw[kPer][kCh];
        unsigned iraw[kInt16 ? kPer : 1][5];  // level 1 under a pull: A's N (3), then 256 va, 256 vb
        // A tile whose region lies inside the level needs no reflection or
        // clamping (a block-uniform branch); elements past the region take
        // its last row, which is inside as well.
        // Fields every element reads, held in scalar registers (one wait).
        const int fw = dev::hold(F.w), fh = dev::hold(F.h);
        int org_x = 0, org_y = 0, wa = 0, ha = 0, bx0 = 0, bx1 = -1, by0 = 0, by1 = -1, pitch_a = 0;
        const std::uint8_t* abase = nullptr;
        const std::uint8_t* flags0 = nullptr;
        const __amdgpu_buffer_rsrc_t rbpk =
            buffer_of(kFromZero ? dev::hold(F.b) : nullptr,
                      kFromZero ? 8u * static_cast<unsigned>(fw) * static_cast<unsigned>(fh) : 0u);
        if (kFromZero) {
            org_x = dev::hold(j.origin_x), org_y = dev::hold(j.origin_y);
            wa = dev::hold(j.wa), ha = dev::hold(j.ha);
            bx0 = dev::hold(j.bx0), bx1 = dev::hold(j.bx1), by0 = dev::hold(j.by0), by1 = dev::hold(j.by1);
            pitch_a = dev::hold(j.pitch_a);
            abase = dev::hold(j.a);
            flags0 = dev::hold(F.flags);
        }
        const bool interior = ix0 >= 0 && iy0 >= 0 && ix0 + kRIW <= fw && iy0 + kRIH <= fh;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int i = static_cast<int>(threadIdx.x) + e * kBlock;
            const int px = ix0 + i % kRIW, py = iy0 + min(i / kRIW, kRIH - 1);
            if (interior) {
                sxs[e] = px;
                sys[e] = py;
            } else {
                sxs[e] = clampi(reflect101(px, fw), 0, fw - 1);
                sys[e] = clampi(reflect101(py, fh), 0, fh - 1);
            }
            const unsigned q = static_cast<unsigned>(sys[e]) * static_cast<unsigned>(fw) + static_cast<unsigned>(sxs[e]);
            unsigned qa = q, qb = q;
            bool& ina = ins[e][0];
            bool& inb = ins[e][1];
            ina = inb = true;
            if (kMode == 0) {
                // An empty support (never, for a renderable pair) reads pixel 0.
                const int ax = max(sa0, min(sxs[e], sa2)), ay = max(sa1, min(sys[e], sa3));
                const int bx2 = max(sb0, min(sxs[e], sb2)), by2 = max(sb1, min(sys[e], sb3));
                ina = ax == sxs[e] && ay == sys[e];
                inb = bx2 == sxs[e] && by2 == sys[e];
                qa = sa2 < sa0 || sa3 < sa1 ? 0u : static_cast<unsigned>(ay) * static_cast<unsigned>(fw) + ax;
                qb = sb2 < sb0 || sb3 < sb1 ? 0u : static_cast<unsigned>(by2) * static_cast<unsigned>(fw) + bx2;
            }
            if (kFromZero) {
                // A's bytes only on A's rectangle and B's colour only inside
                // B's box: both are known from the position alone, so the
                // loads still go out together; elsewhere the image is invalid.
                fls[e] = flags0[q];
                const int ax = sxs[e] + org_x, ay = sys[e] + org_y;
                const bool in_a = ax >= 0 && ax < wa && ay >= 0 && ay < ha;
                const bool in_b = sxs[e] >= bx0 && sxs[e] <= bx1 && sys[e] >= by0 && sys[e] <= by1;
                // Outside A's rectangle the load reads A's first pixel (in
                // bounds) and the bytes are dropped; outside B's box the
                // buffer load of B's colour returns 0.
                unsigned lo, hi;
                load_rgb(abase + (in_a ? static_cast<unsigned>(ay) * static_cast<unsigned>(pitch_a) + 3u * static_cast<unsigned>(ax) : 0u),
                         false, lo, hi);
                lo = in_a ? lo : 0u;
                abytes[e] = lo;
                bpk[e] = __builtin_bit_cast(unsigned long long,
                                            __builtin_amdgcn_raw_buffer_load_b64(rbpk, in_b ? 8u * q : kNoLoad, 0, 0));
            } else if (kInt16) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    iraw[e][c] = ld_at(a16[c], 2u * qa);
                    raw[e][3 + c] = ld_at(planes[3 + c], 4u * qb);
                }
                iraw[e][3] = ld_at(reinterpret_cast<const std::uint16_t*>(planes[kCh - 2]), 2u * qa);
                iraw[e][4] = ld_at(reinterpret_cast<const std::uint16_t*>(planes[kCh - 1]), 2u * qb);
            } else {
#pragma unroll
                for (int c = 0; c < kCh; ++c) {
                    const unsigned qq = kMode == 0 ? (c < 3 || c == kCh - 2 ? qa : qb) : q;
                    if (kW16 && c >= kCh - 2) raw[e][c] = ld_w16(planes[c], qq);
                    else if (kW16 && kMode == 0 && c < 3)  // A's level-1 sums, 16-bit N
                        raw[e][c] = a1(ld_at(a16[c], 2u * qq));
                    else raw[e][c] = ld_at(planes[c], 4u * qq);
                }
            }
        }
        // Pass 2: level-0 selection, then the LDS tile.
        unsigned f_or = 0, f_and = 0xfu;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int i = static_cast<int>(threadIdx.x) + e * kBlock;
            if (i >= kElems) continue;
            if (kFromZero) {
                f_or |= fls[e];
                f_and &= fls[e];
            }
            const int ix = i % kRIW, iy = i / kRIW;
            const int cx = (ix & 1) * kREven + (ix >> 1);
            if (kFromZero) {
                const unsigned f = fls[e];
                const bool has_a = f & kVA, has_b = f & kVB;
                float bc[3];
                unpack_b(bpk[e], bc);
                const unsigned rgb = has_a ? abytes[e] : 0u;
                tile[0][iy][cx] = __builtin_bit_cast(float, rgb & 0x00ff00ffu);  // R | B << 16
                tile[1][iy][cx] = __builtin_bit_cast(float, (rgb >> 8) & 0xffu);  // G
#pragma unroll
                for (int c = 0; c < 3; ++c) tile[2 + c][iy][cx] = has_b ? bc[c] : 0.0f;
                fl[iy][ix] = static_cast<std::uint8_t>(f);
            } else if (kInt16) {
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    t16[c][iy][cx] = static_cast<std::uint16_t>(ins[e][c == 4 ? 1 : 0] ? iraw[e][c] : 0u);
#pragma unroll
                for (int c = 0; c < 3; ++c) tile[c][iy][cx] = ins[e][1] ? raw[e][3 + c] : 0.0f;
            } else {
#pragma unroll
                for (int c = 0; c < kCh; ++c)
                    tile[c][iy][cx] = ins[e][c < 3 || c == kCh - 2 ? 0 : 1] ? raw[e][c] : 0.0f;
            }
        }
        if (kFromZero) {
            // Which flag bits the staged tile holds somewhere (OR) and
            // everywhere (AND), per wave; combined after the barrier.
            unsigned w_or = 0, w_and = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (__ballot((f_or >> k) & 1u)) w_or |= 1u << k;
                if (!__ballot(!((f_and >> k) & 1u))) w_and |= 1u << k;
            }
            if (dev::lane_id() == 0) {
                wave_flags[0][threadIdx.x / 64] = w_or;
                wave_flags[1][threadIdx.x / 64] = w_and;
            }
        }
        __syncthreads();
        // Level 0: a tile whose staged flags are all equal has integer
        // weights 256 * bit (or 0) at every output, and one with no valid
        // pixel has colour sums of exact zeros (every tap adds 0 * 0).
        unsigned t_or = 0xfu, t_and = 0;
        if (kFromZero) {
            t_or = wave_flags[0][0] | wave_flags[0][1] | wave_flags[0][2] | wave_flags[0][3];
            t_and = wave_flags[1][0] & wave_flags[1][1] & wave_flags[1][2] & wave_flags[1][3];
        }
        const bool uniform = kFromZero && t_or == t_and;
        const int lx = threadIdx.x % kRW, ly = threadIdx.x / kRW;
        const int x = x0 + lx, y = y0 + ly;
        if (x < C.w && y < C.h) {
            float acc[kCh];
#pragma unroll
            for (int c = 0; c < kCh; ++c) acc[c] = 0.0f;
            // FP32 channels [c0, c1); at level 0 those are B's, in planes c - 1.
            const auto filter = [&](auto c0, auto c1) {
#pragma unroll
                for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 5; ++dx) {
                        const float w = tap(dx - 2) * tap(dy - 2);
#pragma unroll
                        for (int c = decltype(c0)::value; c < decltype(c1)::value; ++c)
                            acc[c] = fmaf(w,
                                          tile[kFromZero ? c - 1 : (kInt16 ? c - 3 : c)][2 * ly + dy]
                                              [(dx & 1) * kREven + lx + (dx >> 1)],
                                          acc[c]);
                    }
            };
            // A's bytes at level 0 take the integer taps (1 .. 36): exact sums
            // N <= 255 * 256, two channels to a word (R | B << 16, no carry
            // between the 16-bit halves) and G alone.
            unsigned na[3] = {0, 0, 0};
            const auto filter_a = [&]() {
                unsigned rb = 0, g = 0;
#pragma unroll
                for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 5; ++dx) {
                        const unsigned w = static_cast<unsigned>(itap(dx - 2) * itap(dy - 2));
                        const int col = (dx & 1) * kREven + lx + (dx >> 1);
                        const unsigned u = __builtin_bit_cast(unsigned, tile[0][2 * ly + dy][col]);
                        const unsigned v = __builtin_bit_cast(unsigned, tile[1][2 * ly + dy][col]);
                        __builtin_assume(u < (1u << 24));
                        __builtin_assume(v < (1u << 24));
                        rb += u * w;
                        g += v * w;
                    }
                na[0] = rb & 0xffffu;
                na[1] = g;
                na[2] = rb >> 16;
            };
            // Level 1 under a pull: A's N and the weights 256 w with the
            // integer taps, exact (N sums to at most 255 * 256^2 < 2^24).
            unsigned ni[5] = {0, 0, 0, 0, 0};
            const auto filter_i16 = [&]() {
#pragma unroll
                for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 5; ++dx) {
                        const unsigned w = static_cast<unsigned>(itap(dx - 2) * itap(dy - 2));
#pragma unroll
                        for (int c = 0; c < 5; ++c)
                            ni[c] += w * static_cast<unsigned>(t16[c][2 * ly + dy][(dx & 1) * kREven + lx + (dx >> 1)]);
                    }
            };
            using I0 = std::integral_constant<int, 0>;
            using I3 = std::integral_constant<int, 3>;
            using I6 = std::integral_constant<int, 6>;
            using IC = std::integral_constant<int, kCh>;
            if (kInt16) {
                filter_i16();
                filter(I3{}, I6{});
                // A's sums: N2 / (256 * 256 * 255); weights: W2 / 65536 (exact).
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] = static_cast<float>(ni[c]) * (kInvA1 * (1.0f / 256.0f));
                acc[6] = static_cast<float>(ni[3]) * (1.0f / 65536.0f);
                acc[7] = static_cast<float>(ni[4]) * (1.0f / 65536.0f);
            } else if (!kFromZero) {
                filter(I0{}, IC{});
            } else {
                // A colour sums where some staged pixel has A, B's where some
                // has B; the others are sums of zeros.
                if (t_or & kVA) filter_a();
                if (t_or & kVB) filter(I3{}, IC{});
            }
            const size_t p = static_cast<size_t>(y) * C.w + x;
            float* da = kMode == 1 ? C.ga : C.a;
            float* db = kMode == 1 ? C.gb : C.b;
            // Pulls store an image's sums only inside its support at level l:
            // elsewhere the buffer store's offset is out of range and the
            // store is dropped.
            const unsigned pb4 = static_cast<unsigned>(4 * p), nb4 = static_cast<unsigned>(4 * C.n);
            const unsigned oa = kMode == 1 || (x >= oa0 && x <= oa2 && y >= oa1 && y <= oa3) ? pb4 : kNoLoad;
            const unsigned ob = kMode == 1 || (x >= ob0 && x <= ob2 && y >= ob1 && y <= ob3) ? pb4 : kNoLoad;
            const bool own = kFromZero && x >= j.gbox[1][0] && x <= j.gbox[1][2] && y >= j.gbox[1][1] &&
                             y <= j.gbox[1][3];
            const __amdgpu_buffer_rsrc_t rb = buffer_of(db, 3 * nb4);
            if (kFromZero) {
                // A's sums at level 1: the exact integers N, 16 bits each.
                const unsigned nb2 = static_cast<unsigned>(2 * C.n);
                const __amdgpu_buffer_rsrc_t ra = buffer_of(da, 3 * nb2);
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    st_buf_w16(ra, oa == kNoLoad ? kNoLoad : oa / 2 + c * nb2, static_cast<int>(na[c]));
            } else {
                const __amdgpu_buffer_rsrc_t ra = buffer_of(da, 3 * nb4);
#pragma unroll
                for (int c = 0; c < 3; ++c) st_buf(ra, oa == kNoLoad ? kNoLoad : oa + c * nb4, acc[c]);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) st_buf(rb, ob == kNoLoad ? kNoLoad : ob + c * nb4, acc[3 + c]);
            if (kFromZero) {
                int va = 0, vb = 0, ua = 0, ub = 0;
                if (uniform) {
                    va = (t_or & kVA) ? 256 : 0;
                    vb = (t_or & kVB) ? 256 : 0;
                    ua = (t_or & kUA) ? 256 : 0;
                    ub = (t_or & kUB) ? 256 : 0;
                } else
#pragma unroll 1
                for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 5; ++dx) {
                        const unsigned f = fl[2 * ly + dy][2 * lx + dx];
                        const int w = itap(dx - 2) * itap(dy - 2);
                        va += (f & kVA) ? w : 0;
                        vb += (f & kVB) ? w : 0;
                        ua += (f & kUA) ? w : 0;
                        ub += (f & kUB) ? w : 0;
                    }
                const unsigned nb2 = static_cast<unsigned>(2 * C.n);
                st_buf_w16(buffer_of(C.va, nb2), half_off(oa), va);
                st_buf_w16(buffer_of(C.vb, nb2), half_off(ob), vb);
                // Ownership weights are read only inside the level-1 Gaussian
                // region (by the band mix and the Gaussian into level 2).
                if (own) {
                    reinterpret_cast<std::uint16_t*>(C.ua)[p] = static_cast<std::uint16_t>(ua);
                    reinterpret_cast<std::uint16_t*>(C.ub)[p] = static_cast<std::uint16_t>(ub);
                }
            } else {
                st_buf(buffer_of(kMode == 0 ? C.va : C.ua, nb4), oa, acc[6]);
                st_buf(buffer_of(kMode == 0 ? C.vb : C.ub, nb4), ob, acc[7]);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------ Gaussian into level 1
// The Gaussian level 1 is the filter of the filled level 0, F = m V + (1 - m) E
// with m the validity, V the warped colour and E the expansion of the pushed
// level 1 (s6.cu:236-270). It is linear in F, so
//   G = sum w w m V  +  sum w w (1 - m) E.
// The first sum is exactly the pull into level 1, whose raw sums sit in
// a / b; the second vanishes where the 5x5 footprint is all valid (the
// pulled weight is 1), is a fixed separable stencil of the pushed level 1
// where it is all missing (weight 0), and for B is summed tap by tap where
// its border crosses the footprint (A's validity is a rectangle, so its term
// has a closed form everywhere). So no level-0 colour is read here, and
// level-0 flags only along B's border. The sum is the reference's up to
// float rounding order. The ownership weights come from the pull.
constexpr int kGW = 32, kGH = 8;                         // level-1 outputs per tile
constexpr int kGPW = kGW + 4, kGPH = kGH + 4;            // staged pushed level 1 (2 on each side)
static_assert(kGW == kRW && kGH == kRH && kGW * kGH == kBlock, "level-1 Gaussian tiles are the reduce tiles");


// Coefficients of the coarse indices i-2..i+2 in sum_dx w(dx) E(X(dx)) and in
// the same sum restricted to fine positions X in [lo, hi] (A's columns or
// rows): E is the 1D expansion (factor 2 per axis) at the reflected fine
// position X. Every coefficient is a multiple of 1/128, so the float sums
// are exact. Away from the borders nothing reflects and the full sum is
// [1 28 70 28 1] / 128.
__device__ __forceinline__ void stencil_axis(int i, int n0, int n1, int lo, int hi, float (&c)[5], float (&d)[5]) {
#pragma unroll
    for (int k = 0; k < 5; ++k) d[k] = 0.0f;
    if (i >= 2 && i <= n1 - 3) {
        c[0] = c[4] = 1.0f / 128;
        c[1] = c[3] = 28.0f / 128;
        c[2] = 70.0f / 128;
        // Fine position 2i + dx reaches slots (dx + e) / 2 + 2 for the taps e
        // of its parity: even dx -> e in {-2, 0, 2}, odd dx -> e in {-1, 1}.
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int X = 2 * i + dx;
            if (X < lo || X > hi) continue;
            const float w = tap(dx);
            if (dx & 1) {
                d[(dx - 1) / 2 + 2] = fmaf(w, 0.5f, d[(dx - 1) / 2 + 2]);
                d[(dx + 1) / 2 + 2] = fmaf(w, 0.5f, d[(dx + 1) / 2 + 2]);
            } else {
                d[(dx - 2) / 2 + 2] = fmaf(w, 0.125f, d[(dx - 2) / 2 + 2]);
                d[dx / 2 + 2] = fmaf(w, 0.75f, d[dx / 2 + 2]);
                d[(dx + 2) / 2 + 2] = fmaf(w, 0.125f, d[(dx + 2) / 2 + 2]);
            }
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) c[k] = 0.0f;
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
        const int X = clampi(reflect101(2 * i + dx, n0), 0, n0 - 1);
        const bool odd = X & 1, in = X >= lo && X <= hi;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (odd && k == 2) continue;
            const int e = odd ? (k == 0 ? -1 : 1) : 2 * k - 2;
            const int slot = clampi((reflect101(X + e, 2 * n1) >> 1) - i + 2, 0, 4);
            const float v = tap(dx) * 2.0f * tap(e);
            c[slot] += v;
            if (in) d[slot] += v;
        }
    }
}

// sum over the staged 5x5 coarse neighbourhood of (cy cx^T - dy dx^T) P.
__device__ __forceinline__ void stencil_apply(const float4* pp, int lx, int ly, const float (&cx)[5], const float (&cy)[5],
                                              const float (&dx)[5], const float (&dy)[5], bool with_d, float out[3]) {
    float s0 = 0, s1 = 0, s2 = 0, t0 = 0, t1 = 0, t2 = 0;
#pragma unroll 1
    for (int b = 0; b < 5; ++b) {
        float r0 = 0, r1 = 0, r2 = 0, q0 = 0, q1 = 0, q2 = 0;
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            const float4 v = pp[(ly + b) * kGPW + lx + a];
            r0 = fmaf(cx[a], v.x, r0);
            r1 = fmaf(cx[a], v.y, r1);
            r2 = fmaf(cx[a], v.z, r2);
            if (with_d) {
                q0 = fmaf(dx[a], v.x, q0);
                q1 = fmaf(dx[a], v.y, q1);
                q2 = fmaf(dx[a], v.z, q2);
            }
        }
        s0 = fmaf(cy[b], r0, s0);
        s1 = fmaf(cy[b], r1, s1);
        s2 = fmaf(cy[b], r2, s2);
        if (with_d) {
            t0 = fmaf(dy[b], q0, t0);
            t1 = fmaf(dy[b], q1, t1);
            t2 = fmaf(dy[b], q2, t2);
        }
    }
    out[0] = s0 - t0;
    out[1] = s1 - t1;
    out[2] = s2 - t2;
}

// B's border crosses the footprint: the taps where B is missing, or the full
// stencil less the taps where it is valid, whichever visits fewer taps (flags
// read directly: such pixels only line B's border).
__device__ __forceinline__ void gauss1_b_border(const float4* pb, int lx, int ly, int x, int y, int x0, int y0,
                                             const S6Level& F, const S6Level& C, const float (&cx)[5],
                                             const float (&cy)[5], float fb[3]) {
    unsigned valid = 0;
    int nb = 0;
    for (int ty = 0; ty < 5; ++ty)
        for (int tx2 = 0; tx2 < 5; ++tx2) {
            const int X = clampi(reflect101(2 * x + tx2 - 2, F.w), 0, F.w - 1);
            const int Y = clampi(reflect101(2 * y + ty - 2, F.h), 0, F.h - 1);
            const bool v = F.flags[static_cast<unsigned>(Y) * static_cast<unsigned>(F.w) + X] & kVB;
            valid |= (v ? 1u : 0u) << (ty * 5 + tx2);
            nb += v;
        }
    const bool complement = nb < 13;
    float full[3] = {0, 0, 0};
    if (complement) {
        const float none[5] = {0, 0, 0, 0, 0};
        stencil_apply(pb, lx, ly, cx, cy, none, none, false, full);
    }
    float s0 = 0, s1 = 0, s2 = 0;
    for (int ty = 0; ty < 5; ++ty)
        for (int tx2 = 0; tx2 < 5; ++tx2) {
            if ((((valid >> (ty * 5 + tx2)) & 1u) != 0) != complement) continue;
            const int X = clampi(reflect101(2 * x + tx2 - 2, F.w), 0, F.w - 1);
            const int Y = clampi(reflect101(2 * y + ty - 2, F.h), 0, F.h - 1);
            Expand e;
            expand_setup(X, Y, C.w, C.h, x0 - 2, y0 - 2, kGPW, kGPH, e);
            float E[3];
            expand_rgb<kGPW>(pb, e, E);
            const float w = tap(tx2 - 2) * tap(ty - 2);
            s0 = fmaf(w, E[0], s0);
            s1 = fmaf(w, E[1], s1);
            s2 = fmaf(w, E[2], s2);
        }
    fb[0] = complement ? full[0] - s0 : s0;
    fb[1] = complement ? full[1] - s1 : s1;
    fb[2] = complement ? full[2] - s2 : s2;
}

// The level-1 region a Gaussian tile stages (the tile and two pixels around)
// and the pushed level 2 behind it, from which missing level-1 pixels are
// filled: push_missing for level 1 (s6.cu:399-410) happens here, on the
// staged pixels, instead of in a pass of its own.
constexpr int kGSN = kGPW * kGPH, kGSIt = (kGSN + kBlock - 1) / kBlock;
constexpr int kG2W = kGPW / 2 + 4, kG2H = kGPH / 2 + 4;

__global__ void __launch_bounds__(kBlock, 3) s6_gauss1(dev::JobSpan<S6Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ float4 pp[2][kGPH][kGPW];  // pushed level 1, A and B
    __shared__ float4 p2[2][kG2H][kG2W];  // pushed level 2, A and B
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk w = dev::tile_walk(total); w.t < w.end; w.t += w.step) {
        const long long t = w.t;
        ji = next_job_by(span, t, [](const S6Job& j) -> const long long& { return j.gb[1]; }, ji, njobs);
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& C = j.lv[1];
        const S6Level& F = j.lv[0];
        const unsigned local = static_cast<unsigned>(t - j.gb[1]);
        const int bx = j.gbox[1][0], by = j.gbox[1][1];
        const unsigned tx = (static_cast<unsigned>(j.gbox[1][2] - bx + 1) + kGW - 1) / kGW;
        const int x0 = bx + static_cast<int>(local % tx) * kGW, y0 = by + static_cast<int>(local / tx) * kGH;
        const int lx = threadIdx.x % kGW, ly = threadIdx.x / kGW;
        const int x = x0 + lx, y = y0 + ly;
        const bool inside = x < C.w && y < C.h;
        const size_t p = inside ? static_cast<size_t>(y) * C.w + x : 0;
        const bool coarse = j.top >= 2;
        // This pixel's raw pull sums and weights first. The raw sums (a / b)
        // are never written here: the result goes to ga / gb, so
        // neighbouring tiles read unchanged halos.
        // Raw level-1 planes through buffer resources: 0 outside an image's
        // support, where they are not stored.
        // The weights are 16-bit (256 w) at level 1.
        const unsigned nb = static_cast<unsigned>(4 * C.n), nb2 = static_cast<unsigned>(2 * C.n);
        // A's raw sums are 16-bit integers N at level 1 (N / 65280).
        const std::uint16_t* a16 = reinterpret_cast<const std::uint16_t*>(C.a);
        const __amdgpu_buffer_rsrc_t pa[4] = {buffer_of(a16, nb2), buffer_of(a16 + C.n, nb2),
                                              buffer_of(a16 + 2 * C.n, nb2), buffer_of(C.va, nb2)};
        const __amdgpu_buffer_rsrc_t pb[4] = {buffer_of(C.b, nb), buffer_of(C.b + C.n, nb), buffer_of(C.b + 2 * C.n, nb),
                                              buffer_of(C.vb, nb2)};
        const unsigned oa = inside && in_support(j, 1, 0, x, y) ? static_cast<unsigned>(4 * p) : kNoLoad;
        const unsigned ob = inside && in_support(j, 1, 1, x, y) ? static_cast<unsigned>(4 * p) : kNoLoad;
        float ra[3], rb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ra[c] = a1(__builtin_amdgcn_raw_buffer_load_b16(pa[c], half_off(oa), 0, 0));
            rb[c] = ld_buf(pb[c], ob);
        }
        const float wa = ld_buf_w16(pa[3], half_off(oa)), wb = ld_buf_w16(pb[3], half_off(ob));
        // An image's pushed level 1 is needed only if some output of the
        // tile misses it somewhere in its footprint (weight below 1); only
        // then are its staged level-1 pixels and level 2 read. That can only
        // happen where the tile's level-0 footprint leaves A's rectangle, or
        // B's inner box: known from the geometry, so those stagings are
        // issued with this pixel's own loads (one memory round trip). A vote
        // over the loaded weights confirms it; an image the geometry ruled
        // out but some weight needs is staged afterwards.
        const auto stage = [&](bool need_a, bool need_b) {
            if (!need_a && !need_b) return;
            float sa[kGSIt][4], sb[kGSIt][4];
#pragma unroll
            for (int it = 0; it < kGSIt; ++it) {
                const int i = static_cast<int>(threadIdx.x) + it * kBlock;
                if (i >= kGSN) continue;
                const int cx = clampi(x0 - 2 + i % kGPW, 0, C.w - 1), cy = clampi(y0 - 2 + i / kGPW, 0, C.h - 1);
                const unsigned q = static_cast<unsigned>(cy) * static_cast<unsigned>(C.w) + static_cast<unsigned>(cx);
                if (need_a) {
                    const unsigned o = in_support(j, 1, 0, cx, cy) ? 4u * q : kNoLoad;
#pragma unroll
                    for (int c = 0; c < 3; ++c) sa[it][c] = a1(__builtin_amdgcn_raw_buffer_load_b16(pa[c], half_off(o), 0, 0));
                    sa[it][3] = ld_buf_w16(pa[3], half_off(o));
                }
                if (need_b) {
                    const unsigned o = in_support(j, 1, 1, cx, cy) ? 4u * q : kNoLoad;
#pragma unroll
                    for (int c = 0; c < 3; ++c) sb[it][c] = ld_buf(pb[c], o);
                    sb[it][3] = ld_buf_w16(pb[3], half_off(o));
                }
            }
            if (coarse) {
                const S6Level& D = j.lv[2];
                StageRGB<kG2W, kG2W, kG2H> la, lb;
                if (need_a) la.load(D.a, D.n, D.w, D.h, x0 / 2 - 2, y0 / 2 - 2);
                if (need_b) lb.load(D.b, D.n, D.w, D.h, x0 / 2 - 2, y0 / 2 - 2);
                if (need_a) la.store(&p2[0][0][0]);
                if (need_b) lb.store(&p2[1][0][0]);
                __syncthreads();
            }
            // Normalise the staged pixels, or fill them from the expanded level 2.
#pragma unroll
            for (int it = 0; it < kGSIt; ++it) {
                const int i = static_cast<int>(threadIdx.x) + it * kBlock;
                if (i >= kGSN) continue;
                const int cx = clampi(x0 - 2 + i % kGPW, 0, C.w - 1), cy = clampi(y0 - 2 + i / kGPW, 0, C.h - 1);
                const bool fill_a = need_a && !(sa[it][3] > 0), fill_b = need_b && !(sb[it][3] > 0);
                float ea[3] = {0, 0, 0}, eb[3] = {0, 0, 0};
                if (coarse && (fill_a || fill_b)) {
                    Expand e;
                    expand_setup(cx, cy, j.lv[2].w, j.lv[2].h, x0 / 2 - 2, y0 / 2 - 2, kG2W, kG2H, e);
                    if (fill_a) expand_rgb<kG2W>(&p2[0][0][0], e, ea);
                    if (fill_b) expand_rgb<kG2W>(&p2[1][0][0], e, eb);
                }
                // normalize_color divides each channel by the weight; one
                // correctly rounded reciprocal and three products differ from it
                // by at most an ulp, far below what the blend can show.
                if (need_a) {
                    const float r = sa[it][3] > 0 ? 1.0f / sa[it][3] : 0.0f;
                    float4 v;
                    v.x = sa[it][3] > 0 ? sa[it][0] * r : ea[0];
                    v.y = sa[it][3] > 0 ? sa[it][1] * r : ea[1];
                    v.z = sa[it][3] > 0 ? sa[it][2] * r : ea[2];
                    v.w = 0.0f;
                    pp[0][i / kGPW][i % kGPW] = v;
                }
                if (need_b) {
                    const float r = sb[it][3] > 0 ? 1.0f / sb[it][3] : 0.0f;
                    float4 v;
                    v.x = sb[it][3] > 0 ? sb[it][0] * r : eb[0];
                    v.y = sb[it][3] > 0 ? sb[it][1] * r : eb[1];
                    v.z = sb[it][3] > 0 ? sb[it][2] * r : eb[2];
                    v.w = 0.0f;
                    pp[1][i / kGPW][i % kGPW] = v;
                }
            }
            __syncthreads();
        };
        const int ex1 = min(x0 + kGW - 1, C.w - 1), ey1 = min(y0 + kGH - 1, C.h - 1);
        const int fx0 = 2 * x0 - 2, fx1 = 2 * ex1 + 2, fy0 = 2 * y0 - 2, fy1 = 2 * ey1 + 2;
        const int rx0 = -j.origin_x, ry0 = -j.origin_y, rx1 = rx0 + j.wa - 1, ry1 = ry0 + j.ha - 1;
        const bool geo_a = !(fx0 >= rx0 && fx1 <= rx1 && fy0 >= ry0 && fy1 <= ry1);
        const bool geo_b = !(fx0 >= j.bi[0] && fx1 <= j.bi[2] && fy0 >= j.bi[1] && fy1 <= j.bi[3]);
        stage(geo_a, geo_b);
        const bool late_a = __syncthreads_or(inside && wa != 1.0f) && !geo_a;
        const bool late_b = __syncthreads_or(inside && wb != 1.0f) && !geo_b;
        stage(late_a, late_b);
        if (inside) {
            // The footprint is all valid where the weight is 1 and all
            // missing where it is 0 (the weights are exact sums of taps).
            float fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
            const int ax0 = -j.origin_x, ay0 = -j.origin_y, ax1 = ax0 + j.wa - 1, ay1 = ay0 + j.ha - 1;
            float cx[5], cy[5], dx[5], dy[5];
            if (wa != 1.0f || wb != 1.0f) {
                stencil_axis(x, F.w, C.w, ax0, ax1, cx, dx);
                stencil_axis(y, F.h, C.h, ay0, ay1, cy, dy);
            }
            if (wa != 1.0f) {
                // A is valid on a rectangle of the canvas, so its fill term
                // is (cy cx^T - dy dx^T) P with dx, dy restricted to it.
                bool empty_x = true, empty_y = true;
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    empty_x &= dx[k] == 0.0f;
                    empty_y &= dy[k] == 0.0f;
                }
                stencil_apply(&pp[0][0][0], lx, ly, cx, cy, dx, dy, !(empty_x || empty_y), fa);
            }
            if (wb == 0.0f) {
                stencil_apply(&pp[1][0][0], lx, ly, cx, cy, dx, dy, false, fb);
            } else if (wb != 1.0f) {
                gauss1_b_border(&pp[1][0][0], lx, ly, x, y, x0, y0, F, C, cx, cy, fb);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                C.ga[c * C.n + p] = ra[c] + fa[c];
                C.gb[c * C.n + p] = rb[c] + fb[c];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------- push, mix, collapse
// Fine tiles of 64x16 pixels, four rows per thread (rows ty, ty + 4, ...),
// so every thread has four independent pixels' loads in flight; their 3x3
// coarse neighbourhoods span 34x10 coarse pixels, staged in LDS.
constexpr int kEW = 64, kEH = 16, kEPer = kEH / (kBlock / kEW), kCW = kEW / 2 + 2, kCH = kEH / 2 + 2;
static_assert(kEPer * (kBlock / kEW) == kEH, "expand tile");

// normalize_color + push_missing for level l >= 1, top-down (s6.cu:254-270,
// :399-410): the coarse level is final when it is expanded.
__global__ void __launch_bounds__(kBlock) s6_push(dev::JobSpan<S6Job> span, dev::Total work, int l) {
    const long long total = work.get();
    __shared__ float4 cc[2][kCH][kCW];
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk w = dev::tile_walk(total); w.t < w.end; w.t += w.step) {
        const long long t = w.t;
        ji = next_job_by(span, t, [l](const S6Job& j) -> const long long& { return j.pb[l]; }, ji, njobs);
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& L = j.lv[l];
        const unsigned local = static_cast<unsigned>(t - j.pb[l]);
        const int bx = j.pbox[l][0], by = j.pbox[l][1];
        const unsigned tx = (static_cast<unsigned>(j.pbox[l][2] - bx + 1) + kEW - 1) / kEW;
        const int x0 = bx + static_cast<int>(local % tx) * kEW, y0 = by + static_cast<int>(local / tx) * kEH;
        const bool coarse = l < j.top;
        const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1;
        const int x = x0 + static_cast<int>(threadIdx.x) % kEW;
        const unsigned nb = static_cast<unsigned>(4 * L.n);
        const __amdgpu_buffer_rsrc_t pa[4] = {buffer_of(L.a, nb), buffer_of(L.a + L.n, nb), buffer_of(L.a + 2 * L.n, nb),
                                              buffer_of(L.va, nb)};
        const __amdgpu_buffer_rsrc_t pb[4] = {buffer_of(L.b, nb), buffer_of(L.b + L.n, nb), buffer_of(L.b + 2 * L.n, nb),
                                              buffer_of(L.vb, nb)};
        float wa[kEPer], wb[kEPer], va[kEPer][3], vb[kEPer][3];
        bool missing = false;
#pragma unroll
        for (int k = 0; k < kEPer; ++k) {
            const int y = y0 + static_cast<int>(threadIdx.x) / kEW + k * (kBlock / kEW);
            const bool inside = x < L.w && y < L.h;
            const unsigned p = inside ? static_cast<unsigned>(y) * static_cast<unsigned>(L.w) + x : 0u;
            // Raw sums: 0 outside an image's support, where they are not stored.
            const unsigned oa = inside && in_support(j, l, 0, x, y) ? 4u * p : kNoLoad;
            const unsigned ob = inside && in_support(j, l, 1, x, y) ? 4u * p : kNoLoad;
            wa[k] = inside ? ld_buf(pa[3], oa) : 1.0f;
            wb[k] = inside ? ld_buf(pb[3], ob) : 1.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                va[k][c] = ld_buf(pa[c], oa);
                vb[k][c] = ld_buf(pb[c], ob);
            }
            missing |= !(wa[k] > 0) || !(wb[k] > 0);
        }
        // The coarse level is staged only when a pixel of the tile misses A or B.
        const bool staged = coarse && __syncthreads_or(missing);
        if (staged) {
            const S6Level& C = j.lv[l + 1];
            stage_rgb<kCW>(C.a, C.n, C.w, C.h, cx0, cy0, kCW, kCH, &cc[0][0][0]);
            stage_rgb<kCW>(C.b, C.n, C.w, C.h, cx0, cy0, kCW, kCH, &cc[1][0][0]);
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kEPer; ++k) {
            const int y = y0 + static_cast<int>(threadIdx.x) / kEW + k * (kBlock / kEW);
            if (!(x < L.w && y < L.h)) continue;
            const size_t p = static_cast<size_t>(y) * L.w + x;
            float ea[3] = {0, 0, 0}, eb[3] = {0, 0, 0};
            if (coarse && (!(wa[k] > 0) || !(wb[k] > 0))) {
                Expand e;
                expand_setup(x, y, j.lv[l + 1].w, j.lv[l + 1].h, cx0, cy0, kCW, kCH, e);
                if (!(wa[k] > 0)) expand_rgb<kCW>(&cc[0][0][0], e, ea);
                if (!(wb[k] > 0)) expand_rgb<kCW>(&cc[1][0][0], e, eb);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                L.a[c * L.n + p] = wa[k] > 0 ? va[k][c] / wa[k] : ea[c];
                L.b[c * L.n + p] = wb[k] > 0 ? vb[k][c] / wb[k] : eb[c];
            }
        }
        if (staged) __syncthreads();
    }
}

// mix_band + reconstruct (s6.cu:271-293) for level l, top-down; at level 0
// also finish (s6.cu:294-306): single-source pixels keep their warped colour
// (written by the warp) and only overlap pixels need the band mix, so a tile
// without overlap pixels stages nothing. Level 1 stores z - ga and z - gb, the
// two planes level 0 expands (one per overlap pixel, its owner's).
// kL1: level 1, whose ownership weights are 16-bit.
template <bool kL1 = false>
__global__ void __launch_bounds__(kBlock) s6_mix(dev::JobSpan<S6Job> span, dev::Total work, int l) {
    const long long total = work.get();
    __shared__ float4 cc[3][kCH][kCW];
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk w = dev::tile_walk(total); w.t < w.end; w.t += w.step) {
        const long long t = w.t;
        ji = next_job_by(span, t, [l](const S6Job& j) -> const long long& { return j.mb[l]; }, ji, njobs);
        const S6Job& j = dev::job_at(span, ji);
        const S6Level& L = j.lv[l];
        const unsigned local = static_cast<unsigned>(t - j.mb[l]);
        const int bx = l == 0 ? j.ox0 : j.zbox[l][0], by = l == 0 ? j.oy0 : j.zbox[l][1];
        const int bw = l == 0 ? j.ox1 - j.ox0 + 1 : j.zbox[l][2] - j.zbox[l][0] + 1;
        const unsigned tx = (static_cast<unsigned>(bw) + kEW - 1) / kEW;
        const int x0 = bx + static_cast<int>(local % tx) * kEW, y0 = by + static_cast<int>(local / tx) * kEH;
        const bool coarse = l < j.blend_top;
        const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1;
        const int x = x0 + static_cast<int>(threadIdx.x) % kEW;
        // Pass 1: every pixel's inputs, loads issued before any is used; at
        // level 0 A's source byte comes from an in-bounds position and is
        // kept only where A is valid (B likewise).
        float ga[kEPer][3], gb[kEPer][3], ua[kEPer], ub[kEPer];
        unsigned f[kEPer];
        if (l == 0) {
            // An overlap pixel's result is its owner's colour plus the
            // expansion of the owner's difference plane (the other pixels'
            // output came from the warp). Flags, both images' colours and (in
            // the staging below) both difference planes are loaded together
            // from in-bounds positions, one memory round trip per tile; what
            // a pixel does not own is dropped. A's position is clamped into
            // its rectangle (tiles start on the tile grid, left of or above
            // the overlap box); B's packed colour plane covers the canvas.
            unsigned ab[kEPer];  // A's bytes (R, G, B in bytes 0-2)
#pragma unroll
            for (int k = 0; k < kEPer; ++k) {
                const int y = y0 + static_cast<int>(threadIdx.x) / kEW + k * (kBlock / kEW);
                const bool inside = x < L.w && y < L.h;
                const unsigned p = inside ? static_cast<unsigned>(y) * static_cast<unsigned>(L.w) + x : 0u;
                f[k] = inside ? L.flags[p] : 0u;
                const int ax = clampi(x + j.origin_x, 0, j.wa - 1), ay = clampi(y + j.origin_y, 0, j.ha - 1);
                unsigned hi;
                load_rgb(j.a + static_cast<size_t>(ay) * j.pitch_a + 3 * ax, false, ab[k], hi);
                unpack_b(reinterpret_cast<const unsigned long long*>(L.b)[p], gb[k]);
            }
#pragma unroll
            for (int k = 0; k < kEPer; ++k) {
#pragma unroll
                for (int c = 0; c < 3; ++c) ga[k][c] = unit_byte(ab[k], c);
                ua[k] = (f[k] & kUA) ? 1.0f : 0.0f;
                ub[k] = (f[k] & kUB) ? 1.0f : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < kEPer && l > 0; ++k) {
            const int y = y0 + static_cast<int>(threadIdx.x) / kEW + k * (kBlock / kEW);
            const bool inside = x < L.w && y < L.h;
            const size_t p = inside ? static_cast<size_t>(y) * L.w + x : 0;
            {
                f[k] = inside ? (kVA | kVB) : 0u;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ga[k][c] = L.ga[c * L.n + p];
                    gb[k][c] = L.gb[c * L.n + p];
                }
                ua[k] = kL1 ? ld_w16(L.ua, static_cast<unsigned>(p)) : L.ua[p];
                ub[k] = kL1 ? ld_w16(L.ub, static_cast<unsigned>(p)) : L.ub[p];
            }
        }
        // The coarse level behind the tile; its loads go out with the
        // pixels' own.
        const bool staged = coarse;
        if (coarse) {
            const S6Level& C = j.lv[l + 1];
            StageRGB<kCW, kCW, kCH> s0, s1, s2;
            if (l > 0) {
                s0.load(C.ga, C.n, C.w, C.h, cx0, cy0);
                s1.load(C.gb, C.n, C.w, C.h, cx0, cy0);
                s2.load(C.z, C.n, C.w, C.h, cx0, cy0);
                s0.store(&cc[0][0][0]);
                s1.store(&cc[1][0][0]);
                s2.store(&cc[2][0][0]);
            } else {
                // Level 0 expands z1 - ga1 (kept in z) and z1 - gb1 (kept in a).
                s0.load(C.z, C.n, C.w, C.h, cx0, cy0);
                s1.load(C.a, C.n, C.w, C.h, cx0, cy0);
                s0.store(&cc[0][0][0]);
                s1.store(&cc[1][0][0]);
            }
            __syncthreads();
        }
        // Pass 2: band mix and collapse where both images are valid.
#pragma unroll
        for (int k = 0; k < kEPer; ++k) {
            const int y = y0 + static_cast<int>(threadIdx.x) / kEW + k * (kBlock / kEW);
            if (!(x < L.w && y < L.h)) continue;
            const size_t p = static_cast<size_t>(y) * L.w + x;
            float z[3];
            if (l == 0) {
                // An overlap pixel's ownership weights are 0 and 1, so its
                // band is the owner's level-0 colour less the owner's
                // expanded level-1 Gaussian, and the collapse adds the
                // expanded level-1 result: one expansion of
                // d = z1 - g1_owner, which the level-1 mix stored.
                if ((f[k] & (kVA | kVB)) != (kVA | kVB)) continue;  // written by the warp
                const bool own_a = f[k] & kUA;
                float e[3] = {0, 0, 0};
                if (coarse) {
                    Expand ex;
                    expand_setup(x, y, j.lv[1].w, j.lv[1].h, cx0, cy0, kCW, kCH, ex);
                    expand_rgb<kCW>(&cc[own_a ? 0 : 1][0][0], ex, e);
                }
                std::uint8_t* out = j.out + static_cast<size_t>(y) * j.out_pitch + 3 * x;
#pragma unroll
                for (int c = 0; c < 3; ++c) out[c] = quantize((own_a ? ga[k][c] : gb[k][c]) + e[c]);
                continue;
            }
            {
                float ea[3] = {0, 0, 0}, eb[3] = {0, 0, 0}, ez[3] = {0, 0, 0};
                if (coarse) {
                    Expand e;
                    expand_setup(x, y, j.lv[l + 1].w, j.lv[l + 1].h, cx0, cy0, kCW, kCH, e);
                    expand_rgb<kCW>(&cc[0][0][0], e, ea);
                    expand_rgb<kCW>(&cc[1][0][0], e, eb);
                    expand_rgb<kCW>(&cc[2][0][0], e, ez);
                }
                const float tot = ua[k] + ub[k];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float la = ga[k][c] - (coarse ? ea[c] : 0.0f);
                    const float lb = gb[k][c] - (coarse ? eb[c] : 0.0f);
                    const float m = tot > 0 ? fmaf(ua[k], la, ub[k] * lb) / tot : 0.0f;
                    z[c] = coarse ? m + ez[c] : m;
                }
            }
            if (l > 1) {
#pragma unroll
                for (int c = 0; c < 3; ++c) L.z[c * L.n + p] = z[c];
            } else {
                // Level 1 keeps what level 0 expands: z - ga in z, z - gb in a
                // (the pushed colour there is no longer read).
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    L.z[c * L.n + p] = z[c] - ga[k][c];
                    L.a[c * L.n + p] = z[c] - gb[k][c];
                }
            }
        }
        if (staged) __syncthreads();
    }
}

// A geometry the kernels can run: a canvas of addressable size that holds
// all of A (A sits at -origin, so the origin is never positive), non-empty
// inputs, a finite non-zero inverse map and tolerances in their ranges.
bool valid_geometry(const s6::Geometry& g) {
    const bool sizes = g.width > 0 && g.height > 0 && g.width_a > 0 && g.height_a > 0 && g.width_b > 0 &&
                       g.height_b > 0;
    if (!sizes || static_cast<std::size_t>(g.width) >
                      std::numeric_limits<std::size_t>::max() / static_cast<std::size_t>(g.height))
        return false;
    const bool holds_a = g.origin_x <= 0 && g.origin_y <= 0 &&
                         static_cast<long long>(g.origin_x) + g.width >= g.width_a &&
                         static_cast<long long>(g.origin_y) + g.height >= g.height_a;
    const bool finite_map = std::all_of(std::begin(g.inverse), std::end(g.inverse),
                                        [](double v) { return std::isfinite(v); });
    const bool nonzero_map = std::any_of(std::begin(g.inverse), std::end(g.inverse),
                                         [](double v) { return v != 0; });
    const bool tolerances = std::isfinite(g.denominator_tolerance) && g.denominator_tolerance > 0 &&
                            g.denominator_tolerance < 1 && std::isfinite(g.snap_tolerance) &&
                            g.snap_tolerance >= 0 && g.snap_tolerance < 0.5;
    return holds_a && finite_map && nonzero_map && tolerances;
}

bool valid_mode(s6::BlendMode m) { return s6::requested_reductions(m) >= 0; }

// Byte extent [lo, hi) of a buffer.
struct Extent {
    std::uintptr_t lo{0}, hi{0};
    bool meets(const Extent& o) const { return lo < o.hi && o.lo < hi; }
};

// Extent of a pitched image from its first byte to the last pixel byte of
// its last row; false for an unusable view or one that wraps the address space.
bool image_extent(const ImageView& v, PixelFormat format, Extent* e) {
    if (!v.data || v.format != format || v.width <= 0 || v.height <= 0 || v.pitch_bytes <= 0) return false;
    const std::size_t row = static_cast<std::size_t>(v.width) * (format == PixelFormat::RGB8 ? 3 : 1);
    const std::size_t pitch = static_cast<std::size_t>(v.pitch_bytes);
    const std::size_t top = std::numeric_limits<std::uintptr_t>::max();
    if (row > pitch) return false;
    const std::size_t below = static_cast<std::size_t>(v.height) - 1;  // rows after the first
    if (below && below > (top - row) / pitch) return false;
    const std::size_t span = below * pitch + row;
    const auto start = reinterpret_cast<std::uintptr_t>(v.data);
    if (start > top - span) return false;
    *e = {start, start + span};
    return true;
}

std::size_t align256(std::size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

namespace jobs {

// Pyramid layout (s6.cu:60-88): level 0 is the canvas, each level halves with
// ceil down to 1x1. Level 0: B's warped colour (packed, 8 B) and one flag byte;
// levels >= 1: A, B (RGB FP32), VA, VB, and on the blended levels UA, UB, Z;
// level 1 also keeps its Gaussian colour planes apart from the pushed ones,
// which the level-1 Gaussian reads while it writes.
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
    // Level-1 weights: 16 bits each (256 w, exact).
    const auto add16 = [&](std::size_t count, std::size_t& off) {
        cur = align256(cur);
        off = cur;
        cur += count * sizeof(std::uint16_t);
    };
    for (int l = 0; l < kS6MaxLevels; ++l) {
        S6LevelOffsets& o = L.lv[l];
        o.w = w;
        o.h = h;
        const std::size_t n = static_cast<std::size_t>(w) * h;
        add(n, l == 0 ? 2 : 3, o.b);  // level 0: B's packed colour, 8 B per pixel
        if (l == 0) {
            cur = align256(cur);
            o.flags = cur;
            cur += n;
        } else {
            add(n, 3, o.a);
            if (l == 1) {
                add16(n, o.va);
                add16(n, o.vb);
            } else {
                add(n, 1, o.va);
                add(n, 1, o.vb);
            }
            if (l <= requested) {
                if (l == 1) {
                    add16(n, o.ua);
                    add16(n, o.ub);
                } else {
                    add(n, 1, o.ua);
                    add(n, 1, o.ub);
                }
                add(n, 3, o.z);
                if (l == 1) {
                    add(n, 3, o.ga);
                    add(n, 3, o.gb);
                }
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

namespace {

// Canvas box of B's footprint: B's corners through the forward map (the
// inverse of the canvas -> B map, in long double), widened by 2 pixels. A
// canvas pixel outside it maps more than a pixel outside B, which its
// validity test (exact to ~1e-10 px) rejects; the whole canvas when the
// corners cannot be mapped.
void b_box(const s6::Geometry& g, int box[4], int inner[4]) {
    box[0] = 0;
    box[1] = 0;
    box[2] = g.width - 1;
    box[3] = g.height - 1;
    inner[0] = inner[1] = 0;
    inner[2] = inner[3] = -1;
    using R = long double;
    R m[9];
    for (int i = 0; i < 9; ++i) m[i] = g.inverse[i];
    const R c0 = m[4] * m[8] - m[5] * m[7], c1 = m[5] * m[6] - m[3] * m[8], c2 = m[3] * m[7] - m[4] * m[6];
    const R det = m[0] * c0 + m[1] * c1 + m[2] * c2;
    if (!(std::fabs(det) > 0) || !std::isfinite(det)) return;
    const R f[9] = {c0, m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                    c1, m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                    c2, m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
    R x0 = 0, y0 = 0, x1 = 0, y1 = 0, cx[4], cy[4];
    for (int k = 0; k < 4; ++k) {
        const R x = (k & 1) ? g.width_b - 1 : 0, y = (k & 2) ? g.height_b - 1 : 0;
        const R w = f[6] * x + f[7] * y + f[8];
        if (!(std::fabs(w) > 0) || !std::isfinite(w)) return;
        const R px = (f[0] * x + f[1] * y + f[2]) / w - g.origin_x;
        const R py = (f[3] * x + f[4] * y + f[5]) / w - g.origin_y;
        if (!std::isfinite(px) || !std::isfinite(py)) return;
        cx[k] = px;
        cy[k] = py;
        x0 = k ? std::min(x0, px) : px;
        x1 = k ? std::max(x1, px) : px;
        y0 = k ? std::min(y0, py) : py;
        y1 = k ? std::max(y1, py) : py;
    }
    // Inner box: B's corners 0 (0, 0), 1 (w-1, 0), 2 (0, h-1), 3 (w-1, h-1)
    // in their usual order on the canvas make a convex quad (the corners'
    // denominators share a sign) whose left edge lies left of max(x0, x2),
    // right edge right of min(x1, x3), top edge above max(y0, y1) and
    // bottom edge below min(y2, y3); the box between those, less 2 pixels,
    // maps well inside B. Any other order leaves it empty.
    if (cx[1] > cx[0] && cx[3] > cx[2] && cy[2] > cy[0] && cy[3] > cy[1]) {
        const R l = std::max(cx[0], cx[2]) + 2, r = std::min(cx[1], cx[3]) - 2;
        const R t = std::max(cy[0], cy[1]) + 2, b = std::min(cy[2], cy[3]) - 2;
        if (l < r && t < b && l >= 0 && t >= 0 && r <= g.width - 1 && b <= g.height - 1) {
            inner[0] = static_cast<int>(std::ceil(l));
            inner[1] = static_cast<int>(std::ceil(t));
            inner[2] = static_cast<int>(std::floor(r));
            inner[3] = static_cast<int>(std::floor(b));
        }
    }
    const auto clip = [](R v, int lo, int hi) {
        return static_cast<int>(std::max<R>(lo, std::min<R>(hi, v)));
    };
    box[0] = clip(std::floor(x0) - 2, 0, g.width - 1);
    box[1] = clip(std::floor(y0) - 2, 0, g.height - 1);
    box[2] = clip(std::ceil(x1) + 2, -1, g.width - 1);
    box[3] = clip(std::ceil(y1) + 2, -1, g.height - 1);
}

}  // namespace

namespace {

struct Box {
    int x0, y0, x1, y1;
    bool empty() const { return x1 < x0 || y1 < y0; }
};

constexpr Box kEmpty{0, 0, -1, -1};

Box clip(Box b, int w, int h) {
    if (b.empty()) return kEmpty;
    b.x0 = std::max(b.x0, 0);
    b.y0 = std::max(b.y0, 0);
    b.x1 = std::min(b.x1, w - 1);
    b.y1 = std::min(b.y1, h - 1);
    return b.empty() ? kEmpty : b;
}

Box unite(const Box& a, const Box& b) {
    if (a.empty()) return b;
    if (b.empty()) return a;
    return {std::min(a.x0, b.x0), std::min(a.y0, b.y0), std::max(a.x1, b.x1), std::max(a.y1, b.y1)};
}

Box grow(const Box& b, int m, int w, int h) {
    return b.empty() ? kEmpty : clip({b.x0 - m, b.y0 - m, b.x1 + m, b.y1 + m}, w, h);
}

// Coarse pixels an expansion reads for fine pixels in b: (q + d) >> 1 for
// |d| <= 2; reflected indices land inside the same span at the borders.
Box expand_need(const Box& b, int cw, int ch) {
    if (b.empty()) return kEmpty;
    return clip({(b.x0 - 2) >> 1, (b.y0 - 2) >> 1, (b.x1 + 2) >> 1, (b.y1 + 2) >> 1}, cw, ch);
}

// Fine pixels a 5x5 reduce reads for coarse pixels in b: 2q + d, |d| <= 2.
Box reduce_need(const Box& b, int fw, int fh) {
    if (b.empty()) return kEmpty;
    return clip({2 * b.x0 - 2, 2 * b.y0 - 2, 2 * b.x1 + 2, 2 * b.y1 + 2}, fw, fh);
}

void put(int (&dst)[4], const Box& b, int tw, int th) {
    if (b.empty()) {
        dst[0] = dst[1] = 0;
        dst[2] = dst[3] = -1;
        return;
    }
    dst[0] = b.x0 - b.x0 % tw;
    dst[1] = b.y0 - b.y0 % th;
    dst[2] = b.x1;
    dst[3] = b.y1;
}

}  // namespace

// Only overlap pixels take the band mix at level 0; every other pixel keeps
// its warped colour. Working back from the overlap box: the band mix at
// level l reads the Gaussians and the result of level l + 1 through an
// expansion (z regions), the Gaussian at level l + 1 reads level l through a
// reduce (g regions grow by the reduce footprint), the Gaussian into level 1
// fills its footprint from the pushed level 1 two pixels around it, which it
// pushes itself from the pushed level 2, and the push at each level reads the
// next one through an expansion (p regions). Everything outside these boxes
// is never read by anything that reaches the output, so those kernels skip
// it. Pulled values (raw sums and weights) are read by the push at each level
// over its p region, by the Gaussian into level 1 over its region and two
// pixels around (from which it pushes level 1 itself), and by the pull into
// the next level through a reduce (u regions). Only the warp covers the
// whole canvas.
static void s6_regions(S6Job* job) {
    S6Job& j = *job;
    for (int l = 0; l < kS6MaxLevels; ++l) {
        put(j.zbox[l], kEmpty, 1, 1);
        put(j.gbox[l], kEmpty, 1, 1);
        put(j.pbox[l], kEmpty, 1, 1);
        put(j.ubox[l], kEmpty, 1, 1);
    }
    if (j.top < 1 || j.blend_top < 1 || j.ox1 < j.ox0) return;
    const auto W = [&](int l) { return j.lv[l].w; };
    const auto Hh = [&](int l) { return j.lv[l].h; };
    Box z[kS6MaxLevels], g[kS6MaxLevels], p[kS6MaxLevels];
    z[0] = clip({j.ox0, j.oy0, j.ox1, j.oy1}, W(0), Hh(0));
    for (int l = 1; l <= j.blend_top; ++l) z[l] = expand_need(z[l - 1], W(l), Hh(l));
    g[j.blend_top] = z[j.blend_top];
    for (int l = j.blend_top - 1; l >= 1; --l) g[l] = unite(z[l], reduce_need(g[l + 1], W(l), Hh(l)));
    for (int l = 0; l <= j.top; ++l) p[l] = kEmpty;
    if (j.top >= 2) {
        p[2] = expand_need(grow(g[1], 2, W(1), Hh(1)), W(2), Hh(2));
        for (int l = 3; l <= j.top; ++l) p[l] = expand_need(p[l - 1], W(l), Hh(l));
    }
    Box u[kS6MaxLevels];
    for (int l = j.top; l >= 1; --l) {
        u[l] = l >= 2 ? p[l] : grow(g[1], 2, W(1), Hh(1));
        if (l < j.top) u[l] = unite(u[l], reduce_need(u[l + 1], W(l), Hh(l)));
    }
    for (int l = 1; l <= j.top; ++l) {
        if (l <= j.blend_top) {
            put(j.zbox[l], z[l], kEW, kEH);
            put(j.gbox[l], g[l], kRW, kRH);
        }
        if (l >= 2) put(j.pbox[l], p[l], kEW, kEH);
        put(j.ubox[l], u[l], kRW, kRH);
    }
}

// Level l support from level l - 1: X such that [2X - 2, 2X + 2] meets
// [s0, s1], within the level. Reflected footprint positions fold back inside
// the same clipped span, so reflection adds nothing.
static void s6_supports(S6Job* job) {
    S6Job& j = *job;
    for (int l = 0; l < kS6MaxLevels; ++l)
        for (int i = 0; i < 2; ++i) put(j.sup[l][i], kEmpty, 1, 1);
    if (j.top < 1) return;
    Box s[2] = {clip({-j.origin_x, -j.origin_y, -j.origin_x + j.wa - 1, -j.origin_y + j.ha - 1}, j.width, j.height),
                clip({j.bx0, j.by0, j.bx1, j.by1}, j.width, j.height)};
    for (int l = 0; l <= j.top; ++l) {
        for (int i = 0; i < 2; ++i) {
            if (l > 0 && !s[i].empty())
                s[i] = clip({(s[i].x0 - 1) >> 1, (s[i].y0 - 1) >> 1, (s[i].x1 + 2) >> 1, (s[i].y1 + 2) >> 1}, j.lv[l].w,
                            j.lv[l].h);
            put(j.sup[l][i], s[i], 1, 1);
        }
    }
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
    // Aliasing rules (src/s6/s6.cu:338-366): the inputs may share memory with
    // each other but nothing the render writes - panorama, mask, workspace
    // (when the mode uses one) and statistics - may touch any other buffer.
    Extent in_a, in_b, pano, union_mask;
    if (!image_extent(a, PixelFormat::RGB8, &in_a) || !image_extent(b, PixelFormat::RGB8, &in_b) ||
        !image_extent(output, PixelFormat::RGB8, &pano) || !image_extent(mask, PixelFormat::GRAY8, &union_mask))
        return Status::INVALID_ARGUMENT;
    const Extent images[4] = {in_a, in_b, pano, union_mask};
    const auto touches_image = [&](const Extent& e, int skip) {
        for (int i = 0; i < 4; ++i)
            if (i != skip && e.meets(images[i])) return true;
        return false;
    };
    if (touches_image(pano, 2) || touches_image(union_mask, 3)) return Status::INVALID_ARGUMENT;
    if (available < layout.bytes || (layout.bytes && !ws)) return Status::INSUFFICIENT_WORKSPACE;
    Extent arena{};
    const bool have_arena = ws && available;
    if (have_arena) {
        arena.lo = reinterpret_cast<std::uintptr_t>(ws);
        if (arena.lo > std::numeric_limits<std::uintptr_t>::max() - available) return Status::INVALID_ARGUMENT;
        arena.hi = arena.lo + available;
    }
    if (layout.bytes && (arena.lo % 256 || touches_image(arena, -1))) return Status::INVALID_ARGUMENT;
    if (stats) {
        const auto at = reinterpret_cast<std::uintptr_t>(stats);
        if (at % alignof(s6::RenderStatistics)) return Status::INVALID_ARGUMENT;
        const Extent counters{at, at + sizeof(s6::RenderStatistics)};
        // A supplied workspace counts even when the mode does not use it.
        if (touches_image(counters, -1) || (have_arena && counters.meets(arena))) return Status::INVALID_ARGUMENT;
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
    j.inv_wa = 1.0 / a.width;
    j.inv_wb = 1.0 / b.width;
    j.inv_hb = 1.0 / b.height;
    // The snap threshold is the tolerance itself over B's whole extent.
    j.tol_snap = g.snap_tolerance >= 16 * 2.2204460492503131e-16 * std::max(1.0, static_cast<double>(std::max(b.width, b.height)));
    std::memcpy(j.inverse, g.inverse, sizeof j.inverse);
    int box[4];
    b_box(g, box, j.bi);
    j.bx0 = box[0];
    j.by0 = box[1];
    j.bx1 = box[2];
    j.by1 = box[3];
    // Overlap box: A's canvas rectangle within B's box.
    j.ox0 = std::max(j.bx0, -g.origin_x);
    j.oy0 = std::max(j.by0, -g.origin_y);
    j.ox1 = std::min(j.bx1, -g.origin_x + g.width_a - 1);
    j.oy1 = std::min(j.by1, -g.origin_y + g.height_a - 1);
    if (j.ox1 < j.ox0 || j.oy1 < j.oy0) {
        j.ox0 = j.oy0 = 0;
        j.ox1 = j.oy1 = -1;
    } else {
        j.ox0 -= j.ox0 % kEW;
        j.oy0 -= j.oy0 % kEH;
    }
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
            L.b = f(o.b);
            if (l == 0) {
                L.flags = base + o.flags;
                continue;
            }
            L.a = f(o.a);
            L.va = f(o.va);
            L.vb = f(o.vb);
            L.ga = L.a;
            L.gb = L.b;
            if (l <= layout.blend_top) {
                L.ua = f(o.ua);
                L.ub = f(o.ub);
                L.z = f(o.z);
                if (l == 1) {
                    L.ga = f(o.ga);
                    L.gb = f(o.gb);
                }
            }
        }
    }
    s6_regions(&j);
    s6_supports(&j);
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
        t.warp += static_cast<long long>((j.width + kWarpSeg - 1) / kWarpSeg) * j.height;
        for (int l = 0; l < kS6MaxLevels; ++l) {
            j.rb[l] = t.pull[l];
            j.gb[l] = t.gauss[l];
            j.pb[l] = t.push[l];
            j.mb[l] = t.mix[l];
            if (j.top < l) continue;
            const auto in_box = [](const int (&b)[4], int tw, int th) -> long long {
                return b[2] < b[0] || b[3] < b[1] ? 0
                                                  : static_cast<long long>((b[2] - b[0] + tw) / tw) * ((b[3] - b[1] + th) / th);
            };
            // Level 0 of the band mix covers only the overlap box; the Gaussian,
            // the push and the band mix above it cover their dependency regions.
            const long long otiles = j.ox1 < j.ox0 ? 0
                                                   : ((j.ox1 - j.ox0 + kEW) / kEW) * ((j.oy1 - j.oy0 + kEH) / kEH);
            if (l >= 1) {
                t.pull[l] += in_box(j.ubox[l], kRW, kRH);
                if (l >= 2) t.push[l] += in_box(j.pbox[l], kEW, kEH);
            }
            if (j.blend_top >= l) {
                if (l >= 1) t.gauss[l] += in_box(j.gbox[l], kRW, kRH);
                t.mix[l] += l == 0 ? otiles : in_box(j.zbox[l], kEW, kEH);
            }
        }
        t.top = std::max(t.top, j.top);
        t.blend_top = std::max(t.blend_top, j.blend_top);
    }
    return t;
}

const char* probe_kernel_name(int kernel) {
    static const char* names[] = {"s6_clear_stats", "s6_warp", "s6_reduce_pull", "s6_push", "s6_reduce_gauss", "s6_mix"};
    return kernel >= 0 && kernel < LaunchProbe::kKernels ? names[kernel] : "?";
}

double s6_algorithmic_bytes(int kernel, int l, const S6Job& j, const s6::RenderStatistics& st) {
    const auto n = [&](int lv) { return static_cast<double>(j.lv[lv].w) * j.lv[lv].h; };
    const double P = static_cast<double>(j.width) * j.height;
    const double nA = static_cast<double>(j.wa) * j.ha, SB = static_cast<double>(j.wb) * j.hb;
    const double U = static_cast<double>(st.union_pixels), O = static_cast<double>(st.overlap_pixels);
    const double nB = U - nA + O;
    // Pixels of a region box within its level.
    const auto area = [&](const int (&b)[4], int lv) {
        const int x1 = std::min(b[2], j.lv[lv].w - 1), y1 = std::min(b[3], j.lv[lv].h - 1);
        return x1 < b[0] || y1 < b[1] ? 0.0 : static_cast<double>(x1 - b[0] + 1) * (y1 - b[1] + 1);
    };
    switch (kernel) {
        case LaunchProbe::kClear:
            return sizeof(s6::RenderStatistics);
        case LaunchProbe::kWarp:
            // A and B sources once; flags and mask (multiband) or the mask
            // (feather); B's colour where valid; every output byte the warp owns.
            if (j.top < 0) return 3 * nA + 3 * SB + P + 3 * P;
            return 3 * (nA - O) + 3 * SB + 2 * P + 8 * nB + 3 * (P - O);
        case LaunchProbe::kPull:
        case LaunchProbe::kGauss: {
            if (l > (kernel == LaunchProbe::kPull ? j.top : j.blend_top)) return 0;
            // Two colours and two weights per pixel, in and out, over the
            // region (reading four finer pixels per pixel).
            if (l >= 2) return 5 * 32 * area(kernel == LaunchProbe::kPull ? j.ubox[l] : j.gbox[l], l);
            // Pull from level 0: flags, A's bytes and B's colour in (B's
            // colour at the share of its valid pixels that its box holds
            // inside the footprint); two raw colours and two weights out,
            // and two ownership weights in the Gaussian region. Gaussian
            // into level 1: the raw sums and validity weights (pushed on the
            // fly) and the pushed level 2 in, two colours out, over its region.
            if (kernel == LaunchProbe::kPull) {
                const double u1 = area(j.ubox[1], 1);
                if (u1 == 0) return 0;
                const int fx0 = std::max(2 * j.ubox[1][0] - 2, 0), fy0 = std::max(2 * j.ubox[1][1] - 2, 0);
                const int fx1 = std::min(2 * j.ubox[1][2] + 2, j.width - 1), fy1 = std::min(2 * j.ubox[1][3] + 2, j.height - 1);
                const auto span = [](int a0, int a1, int b0, int b1) {
                    return std::max(0, std::min(a1, b1) - std::max(a0, b0) + 1);
                };
                const double foot = static_cast<double>(fx1 - fx0 + 1) * (fy1 - fy0 + 1);
                const double fa = static_cast<double>(span(fx0, fx1, -j.origin_x, j.wa - 1 - j.origin_x)) *
                                  span(fy0, fy1, -j.origin_y, j.ha - 1 - j.origin_y);
                const double bbox = static_cast<double>(j.bx1 - j.bx0 + 1) * (j.by1 - j.by0 + 1);
                const double fb = bbox > 0 ? static_cast<double>(span(fx0, fx1, j.bx0, j.bx1)) *
                                                 span(fy0, fy1, j.by0, j.by1) / bbox * nB
                                           : 0.0;
                return foot + 3 * fa + 8 * fb + 28 * u1 + 4 * area(j.gbox[1], 1);
            }
            return (32 + (j.top >= 2 ? 6 : 0) + 24) * area(j.gbox[1], 1);
        }
        case LaunchProbe::kPush:
            if (l > j.top || l < 2) return 0;
            return (56 + (l < j.top ? 6 : 0)) * area(j.pbox[l], l);
        case LaunchProbe::kMix:
            if (l > j.blend_top) return 0;
            // Gaussians and weights in, the result out (two planes at level 1),
            // the coarser Gaussians and result staged, over the region.
            if (l >= 1) return ((l == 1 ? 56 : 44) + (l < j.blend_top ? 9 : 0)) * area(j.zbox[l], l);
            {
                // Flags over the overlap box; A's bytes, B's colour, the two
                // level-1 difference planes under the overlap and the output.
                const double box = j.ox1 >= j.ox0 && j.oy1 >= j.oy0
                                       ? static_cast<double>(j.ox1 - j.ox0 + 1) * (j.oy1 - j.oy0 + 1) : 0.0;
                return box + 11 * O + 3 * O + (j.blend_top > 0 ? 6 * O : 0);
            }
        default:
            return 0;
    }
}

hipError_t launch_s6(const S6Job* table, int count, const S6Totals& t, const S6Job* single, hipStream_t stream,
                     const dev::Replay<S6Totals>* replay, LaunchProbe* probe) {
    dev::JobSpan<S6Job> span{table, count, single ? *single : S6Job{}, replay ? replay->count : nullptr};
    const S6Totals* d = replay ? replay->totals : nullptr;
    const auto total = [&](const long long* host, const long long* dev) { return dev::Total(*host, d ? dev : nullptr); };
    // Events around each launch when probed.
    const auto timed = [&](int kernel, int level, auto&& launch) {
        const bool on = probe && probe->count < LaunchProbe::kMax;
        if (on) {
            probe->kernel[probe->count] = kernel;
            probe->level[probe->count] = level;
            (void)hipEventRecord(probe->begin[probe->count], stream);
        }
        launch();
        if (on) (void)hipEventRecord(probe->end[probe->count++], stream);
    };
    const bool any_stats = single ? single->stats != nullptr : true;
    if (any_stats) timed(LaunchProbe::kClear, 0, [&] { s6_clear_stats<<<count, 64, 0, stream>>>(span); });
    if (t.warp > 0)
        timed(LaunchProbe::kWarp, 0, [&] {
            s6_warp<<<persistent_grid(t.warp, jobs::Stage::S6), kBlock, 0, stream>>>(span, total(&t.warp, d ? &d->warp : nullptr));
        });
    for (int l = 1; l <= t.top; ++l) {
        if (t.pull[l] <= 0) continue;
        const dev::Total w = total(&t.pull[l], d ? &d->pull[l] : nullptr);
        timed(LaunchProbe::kPull, l, [&] {
            if (l == 1) s6_reduce<2><<<persistent_grid(t.pull[l], jobs::Stage::S6), kBlock, 0, stream>>>(span, w, l);
            else if (l == 2) s6_reduce<0, true><<<persistent_grid(t.pull[l], jobs::Stage::S6), kBlock, 0, stream>>>(span, w, l);
            else s6_reduce<0><<<persistent_grid(t.pull[l], jobs::Stage::S6), kBlock, 0, stream>>>(span, w, l);
        });
    }
    for (int l = t.top; l >= 2; --l)  // level 1 is pushed inside s6_gauss1
        if (t.push[l] > 0)
            timed(LaunchProbe::kPush, l, [&] {
                s6_push<<<persistent_grid(t.push[l], jobs::Stage::S6), kBlock, 0, stream>>>(
                    span, total(&t.push[l], d ? &d->push[l] : nullptr), l);
            });
    for (int l = 1; l <= t.blend_top; ++l) {
        if (t.gauss[l] <= 0) continue;
        const dev::Total w = total(&t.gauss[l], d ? &d->gauss[l] : nullptr);
        timed(LaunchProbe::kGauss, l, [&] {
            if (l == 1) s6_gauss1<<<persistent_grid(t.gauss[l], jobs::Stage::S6), kBlock, 0, stream>>>(span, w);
            else if (l == 2) s6_reduce<1, true><<<persistent_grid(t.gauss[l], jobs::Stage::S6), kBlock, 0, stream>>>(span, w, l);
            else s6_reduce<1><<<persistent_grid(t.gauss[l], jobs::Stage::S6), kBlock, 0, stream>>>(span, w, l);
        });
    }
    for (int l = t.blend_top; l >= 0; --l)
        if (t.mix[l] > 0)
            timed(LaunchProbe::kMix, l, [&] {
                if (l == 1)
                    s6_mix<true><<<persistent_grid(t.mix[l], jobs::Stage::S6), kBlock, 0, stream>>>(
                        span, total(&t.mix[l], d ? &d->mix[l] : nullptr), l);
                else
                    s6_mix<<<persistent_grid(t.mix[l], jobs::Stage::S6), kBlock, 0, stream>>>(
                        span, total(&t.mix[l], d ? &d->mix[l] : nullptr), l);
            });
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
