// S2: single-level ORB detection for any number of images per launch.
// FAST-9/16 score -> strict 3x3 maxima inside the 31-px margin -> the 2K
// highest FAST scores (ties to the smaller row-major index) -> Harris response
// (integer sums, FP32 without contraction) -> the K highest responses (ties to
// the smaller index). Semantics: src/s2_s3/s2.cu:39-387 and
// tools/s2_s3/reference.py:50-147.
//
// MI355X layout: FAST and NMS run fused on LDS tiles with halos; the
// pre-selection is an ordered compaction with wave ballots; the top-K order
// is a rank count split over many workgroups instead of one O(n^2) block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

#include "stitch/jobs.hpp"

namespace stitch {
namespace {

using jobs::S2Job;
using jobs::SelectParams;
using jobs::kS2Chunk;
using jobs::kS2RankKeys;
using jobs::kS2RankSegment;
using jobs::kS2TileH;
using jobs::kS2TileW;

constexpr int kEdge = s2::kEdge;
constexpr int kHalo = 4;  // 3 for the FAST circle + 1 for the NMS neighbourhood
constexpr int kGrayW = kS2TileW + 2 * kHalo, kGrayH = kS2TileH + 2 * kHalo;
constexpr int kScoreW = kS2TileW + 2, kScoreH = kS2TileH + 2;
// Score rows in LDS: pitch 72 bytes, score column sx at byte sx + 3, so the
// tile's pixel x0 + 4t starts word t + 1 (NMS reads whole words).
constexpr int kScorePitch = kS2TileW + 8, kScoreOff = 3;

__constant__ int kCircleX[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
__constant__ int kCircleY[16] = {3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1, 0, 1, 2, 3};

__device__ __forceinline__ void s2_init_job(const S2Job& j) {
    if (!j.active) {
        if (threadIdx.x == 0) *j.d_count = 0;
        return;
    }
    for (int i = threadIdx.x; i < 256; i += blockDim.x) j.hist[i] = 0;
    for (int i = threadIdx.x; i < j.candidates; i += blockDim.x) j.ranks[i] = 0;
}

__global__ void s2_init(dev::JobSpan<S2Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s2_init_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

// OpenCV's FAST-9 early exit on the four compass pixels: a pixel can only be
// a corner when at least two of them are brighter than v + t or at least two
// darker than v - t. "At least two of four above a bound" is "the second
// largest above it": max(med3(a, b, c), min(max3(a, b, c), d)); likewise the
// second smallest is min(med3(a, b, c), max(min3(a, b, c), d)). Branch-free,
// about 11 VALU per pixel instead of 8 compare-and-count steps.
__device__ __forceinline__ bool fast_candidate(const std::uint8_t* g, int threshold) {
    const int v = g[0];
    const int a = g[3 * kGrayW], b = g[3], c = g[-3 * kGrayW], d = g[-3];
    const int hi3 = max(max(a, b), c), lo3 = min(min(a, b), c);
    const int mid3 = max(min(a, b), min(max(a, b), c));
    return max(mid3, min(hi3, d)) > v + threshold || min(mid3, max(lo3, d)) < v - threshold;
}

// cornerScore<16> on an LDS tile for a pixel that passed fast_candidate; g
// points at the centre pixel, row stride kGrayW.
__device__ int fast_score_lds(const std::uint8_t* g, int threshold) {
    const int v = g[0];
    int d[16];
    unsigned dark = 0, bright = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        d[k] = v - g[kCircleY[k] * kGrayW + kCircleX[k]];
        dark |= static_cast<unsigned>(d[k] > threshold) << k;
        bright |= static_cast<unsigned>(d[k] < -threshold) << k;
    }
    // The score exceeds the threshold exactly when 9 contiguous circle
    // pixels are all darker or all brighter; otherwise it is 0.
    const auto arc9 = [](unsigned m) {
        const unsigned x = m | (m << 16);
        unsigned y = x & (x >> 1);
        y &= y >> 2;
        y &= y >> 4;
        return (y & (x >> 8)) != 0;
    };
    if (!arc9(dark) && !arc9(bright)) return 0;
    // cornerScore: the best over the 16 arcs of 9 of min(d) and of -max(d).
    // Each arc is scanned directly (only d[] stays live: this path is rare,
    // and fewer registers keep the kernel's occupancy up).
    int best = INT_MIN;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        int lo = d[k], hi = d[k];
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            lo = min(lo, d[(k + i) & 15]);
            hi = max(hi, d[(k + i) & 15]);
        }
        best = max(best, max(lo, -hi));
    }
    return best > threshold ? best - 1 : 0;
}

// FAST + NMS + histogram on 64x16 tiles with a 4-pixel gray halo in LDS.
__global__ void __launch_bounds__(256) s2_fast_nms(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ __align__(4) std::uint8_t gray[kGrayH * kGrayW];
    __shared__ __align__(16) std::uint8_t score[kScoreH * kScorePitch];
    __shared__ std::uint32_t hist[256];
    const int tid = threadIdx.x;
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk tw = dev::tile_walk(total); tw.t < tw.end; tw.t += tw.step) {
        const long long t = tw.t;
        ji = dev::next_job(span, t, &S2Job::fast_begin, ji, njobs);
        const S2Job& j = dev::job_at(span, ji);
        // Job fields used per pixel, read once per tile.
        const int w = j.w, h = j.h, pitch = j.pitch, thr = j.threshold;
        const std::uint8_t* img = j.img;
        std::uint8_t* maxima = j.maxima;
        const unsigned local = static_cast<unsigned>(t - j.fast_begin), tiles_x = static_cast<unsigned>(j.tiles_x);
        const int x0 = static_cast<int>(local % tiles_x) * kS2TileW;
        const int y0 = static_cast<int>(local / tiles_x) * kS2TileH;
        hist[tid] = 0;
        constexpr int kWords = kGrayW / 4;
        static_assert(kGrayW % 4 == 0 && kHalo % 4 == 0 && kS2TileW % 4 == 0, "gray tile words");
        const bool words = ((reinterpret_cast<std::uintptr_t>(img) | static_cast<std::uintptr_t>(pitch)) & 3u) == 0;
        if (words && x0 >= kHalo && x0 - kHalo + kGrayW <= w && y0 >= kHalo && y0 - kHalo + kGrayH <= h) {
            // Interior tile: every word of the window lies in the image.
            for (int base = 0; base < kGrayH * kWords; base += 256) {
                const int i = base + tid;
                if (i < kGrayH * kWords)
                    reinterpret_cast<unsigned*>(gray)[i] = *reinterpret_cast<const unsigned*>(
                        img + static_cast<size_t>(y0 - kHalo + i / kWords) * pitch + x0 - kHalo + 4 * (i % kWords));
            }
        } else if (words) {
            // Four pixels per load (x0 - kHalo is a multiple of 4); pixels
            // outside the image read 0.
            for (int i = tid; i < kGrayH * kWords; i += 256) {
                const int gx = x0 - kHalo + 4 * (i % kWords), gy = y0 - kHalo + i / kWords;
                unsigned word = 0;
                if (gy >= 0 && gy < h) {
                    const std::uint8_t* row = img + static_cast<size_t>(gy) * pitch;
                    if (gx >= 0 && gx + 3 < w) {
                        word = *reinterpret_cast<const unsigned*>(row + gx);
                    } else {
                        for (int b = 0; b < 4; ++b)
                            if (gx + b >= 0 && gx + b < w) word |= static_cast<unsigned>(row[gx + b]) << (8 * b);
                    }
                }
                reinterpret_cast<unsigned*>(gray)[i] = word;
            }
        } else {
            for (int i = tid; i < kGrayH * kGrayW; i += 256) {
                const int gx = x0 - kHalo + i % kGrayW, gy = y0 - kHalo + i / kGrayW;
                gray[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? img[static_cast<size_t>(gy) * pitch + gx] : 0;
            }
        }
        __syncthreads();
        // Scores of the tile and its one-pixel ring. Trip counts are uniform;
        // the candidate test runs on every lane (indexes clamped into the
        // tile, conditions combined without branches) and the full score only
        // where it passed.
        // A tile whose scored pixels all lie 3 or more pixels inside the
        // image (most tiles) needs no per-pixel bounds test (a block-uniform
        // branch between two copies of the pass).
        const auto score_pass = [&](auto all_inside) {
            #pragma unroll 1
            for (int base = 0; base < kScoreH * kScoreW; base += 256) {
                const int i = base + tid, ic = min(i, kScoreH * kScoreW - 1);
                const int sx = ic % kScoreW, sy = ic / kScoreW;
                const int gx = x0 - 1 + sx, gy = y0 - 1 + sy;
                const std::uint8_t* g = &gray[(sy + kHalo - 1) * kGrayW + sx + kHalo - 1];
                const bool inside = decltype(all_inside)::value || ((gx >= 3) & (gx < w - 3) & (gy >= 3) & (gy < h - 3));
                int s = 0;
                if (inside & fast_candidate(g, thr)) s = fast_score_lds(g, thr);
                if (i < kScoreH * kScoreW) score[sy * kScorePitch + sx + kScoreOff] = static_cast<std::uint8_t>(s);
            }
        };
        if (x0 - 1 >= 3 && x0 + kS2TileW < w - 3 && y0 - 1 >= 3 && y0 + kS2TileH < h - 3)
            score_pass(std::true_type{});
        else
            score_pass(std::false_type{});
        __syncthreads();
        // Strict 3x3 maxima on strips of four pixels: one word of scores per
        // strip; where all four are 0 (most strips) the strip writes zeros,
        // otherwise its eight neighbouring words are read and each pixel is
        // tested against its eight neighbours.
        constexpr int kStripsX = kS2TileW / 4;
        static_assert(kStripsX * kS2TileH % 256 == 0, "NMS trips");
        const unsigned* sw = reinterpret_cast<const unsigned*>(score);
        constexpr int kRowWords = kScorePitch / 4;
        #pragma unroll 1
        for (int base = 0; base < kStripsX * kS2TileH; base += 256) {
            const int i = base + tid;
            const int t = i % kStripsX, ly = i / kStripsX;
            const int gx = x0 + 4 * t, gy = y0 + ly;
            const int o = (ly + 1) * kRowWords + t + 1;  // the strip's word in its score row
            const unsigned m = sw[o];
            unsigned kept = 0;
            if (m != 0) {
                const unsigned ml = sw[o - 1], mr = sw[o + 1];
                const unsigned u = sw[o - kRowWords], ul = sw[o - kRowWords - 1], ur = sw[o - kRowWords + 1];
                const unsigned d = sw[o + kRowWords], dl = sw[o + kRowWords - 1], dr = sw[o + kRowWords + 1];
                // Neighbours at x - 1 and x + 1 of the strip's four pixels.
                const unsigned l1 = __builtin_amdgcn_alignbyte(m, ml, 3), r1 = __builtin_amdgcn_alignbyte(mr, m, 1);
                const unsigned l0 = __builtin_amdgcn_alignbyte(u, ul, 3), r0 = __builtin_amdgcn_alignbyte(ur, u, 1);
                const unsigned l2 = __builtin_amdgcn_alignbyte(d, dl, 3), r2 = __builtin_amdgcn_alignbyte(dr, d, 1);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const unsigned sh = 8 * k;
                    const unsigned s = (m >> sh) & 0xffu;
                    const auto b = [sh](unsigned v) { return (v >> sh) & 0xffu; };
                    const bool keep = (s > 0) & (gx + k >= kEdge) & (gx + k < w - kEdge) & (gy >= kEdge) &
                                      (gy < h - kEdge) & (s > b(l0)) & (s > b(u)) & (s > b(r0)) & (s > b(l1)) &
                                      (s > b(r1)) & (s > b(l2)) & (s > b(d)) & (s > b(r2));
                    if (keep) {
                        kept |= s << sh;
                        atomicAdd(&hist[s], 1u);
                    }
                }
            }
            if (gy < h) {
                std::uint8_t* dst = maxima + static_cast<size_t>(gy) * w + gx;
                if (gx + 3 < w && (reinterpret_cast<std::uintptr_t>(dst) & 3u) == 0) {
                    *reinterpret_cast<unsigned*>(dst) = kept;
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (gx + k < w) dst[k] = static_cast<std::uint8_t>(kept >> (8 * k));
                }
            }
        }
        __syncthreads();
        if (hist[tid]) atomicAdd(&j.hist[tid], hist[tid]);
        __syncthreads();
    }
}

// One wave per job: suffix[s] = #maxima with score >= s; the boundary score
// of the 2K pre-selection is the largest s with suffix[s] >= 2K.
__device__ __forceinline__ void s2_select_job(const S2Job& j) {
    if (!j.active) return;
    __shared__ std::uint32_t suffix[257];
    const int lane = threadIdx.x;
    std::uint32_t h[4], sum = 0;
    for (int k = 0; k < 4; ++k) {
        const int s = 4 * lane + k;
        h[k] = s ? j.hist[s] : 0;
        sum += h[k];
    }
    // Inclusive suffix scan of the per-lane sums over lanes (lane 63 is the top).
    std::uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const std::uint32_t o = __shfl_down(incl, off, 64);
        if (lane + off < 64) incl += o;
    }
    std::uint32_t run = incl - sum;  // sum over higher lanes
    for (int k = 3; k >= 0; --k) {
        run += h[k];
        suffix[4 * lane + k] = run;
    }
    if (lane == 0) suffix[256] = 0;
    __syncthreads();
    const std::uint32_t cap = static_cast<std::uint32_t>(j.candidates);
    const std::uint32_t total = suffix[1];
    for (int k = 0; k < 4; ++k) {
        const int s = 4 * lane + k;
        SelectParams p{};
        bool writer = false;
        if (total <= cap) {
            writer = s == 1;
            p = {1u, total - suffix[1] + suffix[2], suffix[1] - suffix[2], total};
        } else if (s >= 1 && suffix[s] >= cap && suffix[s + 1] < cap) {
            writer = true;
            p = {static_cast<std::uint32_t>(s), suffix[s + 1], cap - suffix[s + 1], cap};
        }
        if (writer) {
            *j.params = p;
            *j.d_count = static_cast<int>(min(p.selected, static_cast<std::uint32_t>(j.K)));
        }
    }
}

__global__ void __launch_bounds__(64) s2_select(dev::JobSpan<S2Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s2_select_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

struct Flags {
    std::uint32_t above, tie;
};

// Counts (above, tie) over 16 consecutive pixels starting at p.
__device__ Flags count16(const std::uint8_t* maxima, long long pixels, long long p, std::uint32_t thr,
                         std::uint32_t bytes[4]) {
    Flags f{0, 0};
    if (p + 16 <= pixels) {
        const uint4 v = *reinterpret_cast<const uint4*>(maxima + p);
        bytes[0] = v.x; bytes[1] = v.y; bytes[2] = v.z; bytes[3] = v.w;
    } else {
        for (int w = 0; w < 4; ++w) {
            std::uint32_t word = 0;
            for (int b = 0; b < 4; ++b) {
                const long long q = p + 4 * w + b;
                if (q < pixels) word |= static_cast<std::uint32_t>(maxima[q]) << (8 * b);
            }
            bytes[w] = word;
        }
    }
    if ((bytes[0] | bytes[1] | bytes[2] | bytes[3]) == 0) return f;  // no maximum: the common case
    for (int i = 0; i < 16; ++i) {
        const std::uint32_t s = (bytes[i >> 2] >> (8 * (i & 3))) & 0xffu;
        f.above += s > thr;
        f.tie += (s && s == thr);
    }
    return f;
}

// Block-wide exclusive scan of (above, tie); returns the block totals.
__device__ Flags block_scan(Flags v, Flags* excl) {
    __shared__ Flags wave_tot[4];
    const int lane = dev::lane_id(), wave = threadIdx.x / 64;
    Flags incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const std::uint32_t a = __shfl_up(incl.above, off, 64), t = __shfl_up(incl.tie, off, 64);
        if (lane >= off) { incl.above += a; incl.tie += t; }
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    Flags base{0, 0}, total{0, 0};
    for (int w = 0; w < 4; ++w) {
        if (w < wave) { base.above += wave_tot[w].above; base.tie += wave_tot[w].tie; }
        total.above += wave_tot[w].above;
        total.tie += wave_tot[w].tie;
    }
    __syncthreads();
    excl->above = base.above + incl.above - v.above;
    excl->tie = base.tie + incl.tie - v.tie;
    return total;
}

// Block-wide totals of (above, tie).
__device__ Flags block_total(Flags v) {
    __shared__ Flags wave_tot[4];
    const int wave = threadIdx.x / 64;
    v.above = dev::wave_sum(v.above);
    v.tie = dev::wave_sum(v.tie);
    if (dev::lane_id() == 0) wave_tot[wave] = v;
    __syncthreads();
    Flags total{0, 0};
    for (int w = 0; w < 4; ++w) {
        total.above += wave_tot[w].above;
        total.tie += wave_tot[w].tie;
    }
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(256) s2_count(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk tw = dev::tile_walk(total); tw.t < tw.end; tw.t += tw.step) {
        const long long t = tw.t;
        ji = dev::next_job(span, t, &S2Job::chunk_begin, ji, njobs);
        const S2Job& j = dev::job_at(span, ji);
        const long long c = t - j.chunk_begin;
        const long long pixels = static_cast<long long>(j.w) * j.h;
        const std::uint32_t thr = j.params->threshold;
        std::uint32_t bytes[4];
        const Flags f = count16(j.maxima, pixels, c * kS2Chunk + threadIdx.x * 16, thr, bytes);
        const Flags tot = block_total(f);
        if (threadIdx.x == 0) {
            j.chunk_counts[2 * c] = tot.above;
            j.chunk_counts[2 * c + 1] = tot.tie;
        }
    }
}

// One workgroup per job: exclusive prefix over the job's chunk totals.
__device__ __forceinline__ void s2_scan_job(const S2Job& j) {
    if (!j.active) return;
    const int per = (j.chunks + 255) / 256;
    const int first = threadIdx.x * per;
    Flags sum{0, 0};
    for (int i = 0; i < per && first + i < j.chunks; ++i) {
        sum.above += j.chunk_counts[2 * (first + i)];
        sum.tie += j.chunk_counts[2 * (first + i) + 1];
    }
    Flags run;
    block_scan(sum, &run);
    for (int i = 0; i < per && first + i < j.chunks; ++i) {
        j.chunk_offsets[2 * (first + i)] = run.above;
        j.chunk_offsets[2 * (first + i) + 1] = run.tie;
        run.above += j.chunk_counts[2 * (first + i)];
        run.tie += j.chunk_counts[2 * (first + i) + 1];
    }
}

__global__ void __launch_bounds__(256) s2_scan(dev::JobSpan<S2Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s2_scan_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

// Writes the pixel index of every pre-selected maximum to its candidate slot:
// all points above the boundary score, then the first `ties` points at it.
__global__ void __launch_bounds__(256) s2_scatter(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    // A chunk that selects nothing (the count pass saw no maximum at or
    // above the boundary score) writes nothing.
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S2Job& jj = dev::lane_job(span, tk, &S2Job::chunk_begin, ji, njobs);
            const long long c = tk - jj.chunk_begin;
            return (jj.chunk_counts[2 * c] | jj.chunk_counts[2 * c + 1]) != 0;
        },
        [&](long long t) {
        ji = dev::next_job(span, t, &S2Job::chunk_begin, ji, njobs);
        const S2Job& j = dev::job_at(span, ji);
        const long long c = t - j.chunk_begin;
        const long long pixels = static_cast<long long>(j.w) * j.h;
        const SelectParams p = *j.params;
        const long long first = c * kS2Chunk + threadIdx.x * 16;
        std::uint32_t bytes[4];
        const Flags f = count16(j.maxima, pixels, first, p.threshold, bytes);
        Flags rank;
        block_scan(f, &rank);
        rank.above += j.chunk_offsets[2 * c];
        rank.tie += j.chunk_offsets[2 * c + 1];
        if (f.above | f.tie) {
            for (int i = 0; i < 16; ++i) {
                const std::uint32_t s = (bytes[i >> 2] >> (8 * (i & 3))) & 0xffu;
                const std::uint64_t px = static_cast<std::uint64_t>(first + i);
                if (s > p.threshold) j.keys[rank.above++] = px;
                else if (s && s == p.threshold) {
                    if (rank.tie < p.ties) j.keys[p.above + rank.tie] = px;
                    ++rank.tie;
                }
            }
        }
        });
}

__device__ std::uint32_t order_bits(float value) {
    const std::uint32_t u = __float_as_uint(value);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ float from_order_bits(std::uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// OpenCV HarrisResponses: 7x7 Sobel sums in integers, FP32 formula.
__device__ float harris_response(const std::uint8_t* image, int pitch, int x0, int y0, float k) {
    int a = 0, b = 0, c = 0;
    for (int i = 0; i < 7; ++i) {
        const std::uint8_t* row = image + static_cast<size_t>(y0 - 3 + i) * pitch + (x0 - 3);
#pragma unroll
        for (int jx = 0; jx < 7; ++jx) {
            const std::uint8_t* q = row + jx;
            const int ix = (q[1] - q[-1]) * 2 + (q[-pitch + 1] - q[-pitch - 1]) + (q[pitch + 1] - q[pitch - 1]);
            const int iy = (q[pitch] - q[-pitch]) * 2 + (q[pitch - 1] - q[-pitch - 1]) + (q[pitch + 1] - q[-pitch + 1]);
            a += ix * ix;
            b += iy * iy;
            c += ix * iy;
        }
    }
    const float fa = static_cast<float>(a), fb = static_cast<float>(b), fc = static_cast<float>(c);
    const float trace = fa + fb;
    const float det = fa * fb - fc * fc;
    const float penalty = (k * trace) * trace;
    return (det - penalty) * 0x1.bb9da2p-52f;
}

__global__ void __launch_bounds__(256) s2_harris(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    // Only the slots below the selected count hold keys (the later ones are
    // never read).
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S2Job& jj = dev::lane_job(span, tk, &S2Job::cand_begin, ji, njobs);
            const long long slot0 = (tk - jj.cand_begin) * 256;
            return slot0 < jj.candidates && slot0 < static_cast<long long>(jj.params->selected);
        },
        [&](long long t) {
            ji = dev::next_job(span, t, &S2Job::cand_begin, ji, njobs);
            const S2Job& j = dev::job_at(span, ji);
            const int slot = static_cast<int>(t - j.cand_begin) * 256 + threadIdx.x;
            if (slot >= j.candidates || static_cast<std::uint32_t>(slot) >= j.params->selected) return;
            const auto px = static_cast<std::uint32_t>(j.keys[slot]);
            const int x = static_cast<int>(px % static_cast<std::uint32_t>(j.w));
            const int y = static_cast<int>(px / static_cast<std::uint32_t>(j.w));
            const float r = harris_response(j.img, j.pitch, x, y, j.harris_k);
            j.keys[slot] = (static_cast<std::uint64_t>(order_bits(r)) << 32) | (0xffffffffu - px);
        });
}

// Sorts each kS2RankSegment-key segment of the selected candidates in place,
// descending, with a bitonic network in LDS (unused slots padded with 0,
// below every key: a key's high word is order_bits of a finite response,
// never 0). Keys are unique (the low word holds the pixel index), so after
// this pass a key's rank among all keys is its position in its own segment
// plus the number of greater keys in each other segment.
__global__ void __launch_bounds__(256) s2_sort(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ std::uint64_t seg[kS2RankSegment];
    int ji = 0;
    const int njobs = span.jobs();
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S2Job& jj = dev::lane_job(span, tk, &S2Job::sort_begin, ji, njobs);
            return (tk - jj.sort_begin) * kS2RankSegment < static_cast<long long>(jj.params->selected);
        },
        [&](long long t) {
            ji = dev::next_job(span, t, &S2Job::sort_begin, ji, njobs);
            const S2Job& j = dev::job_at(span, ji);
            const int seg0 = static_cast<int>(t - j.sort_begin) * kS2RankSegment;
            const int n = min(kS2RankSegment, static_cast<int>(j.params->selected) - seg0);
            // The network spans the smallest power of two holding the keys.
            int span_n = 2;
            while (span_n < n) span_n <<= 1;
            for (int i = threadIdx.x; i < span_n; i += 256) seg[i] = i < n ? j.keys[seg0 + i] : 0ull;
            __syncthreads();
            for (int k = 2; k <= span_n; k <<= 1) {
                for (int d = k >> 1; d > 0; d >>= 1) {
                    // Pair i compares elements lo and lo + d; blocks of k
                    // alternate direction so that every merge sees a bitonic
                    // sequence, and the final block (all of it) ends descending.
                    for (int i = threadIdx.x; i < span_n / 2; i += 256) {
                        const int lo = 2 * i - (i & (d - 1));
                        const bool descending = (lo & k) == 0;
                        const std::uint64_t x = seg[lo], y = seg[lo + d];
                        if ((x < y) == descending) {
                            seg[lo] = y;
                            seg[lo + d] = x;
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = threadIdx.x; i < n; i += 256) j.keys[seg0 + i] = seg[i];
            __syncthreads();
        });
}

// rank(key) = #keys greater, over segment-sorted keys. Tile = 256 keys x one
// sorted segment: the segment goes to LDS and each thread counts the
// segment's keys above its own by binary search (descending order: they are
// a prefix). Partial counts of the segments add up atomically.
__global__ void __launch_bounds__(256) s2_rank(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ std::uint64_t seg[kS2RankSegment];
    int ji = 0;
    const int njobs = span.jobs();
    const auto split = [](const S2Job& j, long long t, int& key0, int& seg0) {
        const unsigned local = static_cast<unsigned>(t - j.rank_begin);
        const int segs = (j.candidates + kS2RankSegment - 1) / kS2RankSegment;
        key0 = static_cast<int>(local / segs) * kS2RankKeys;
        seg0 = static_cast<int>(local % segs) * kS2RankSegment;
    };
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S2Job& jj = dev::lane_job(span, tk, &S2Job::rank_begin, ji, njobs);
            int key0, seg0;
            split(jj, tk, key0, seg0);
            const int selected = static_cast<int>(jj.params->selected);
            return key0 < selected && seg0 < selected;
        },
        [&](long long t) {
            ji = dev::next_job(span, t, &S2Job::rank_begin, ji, njobs);
            const S2Job& j = dev::job_at(span, ji);
            int key0, seg0;
            split(j, t, key0, seg0);
            const int selected = static_cast<int>(j.params->selected);
            const int n = min(kS2RankSegment, selected - seg0);
            for (int i = threadIdx.x; i < n; i += 256) seg[i] = j.keys[seg0 + i];
            __syncthreads();
            const int slot = key0 + threadIdx.x;
            if (slot < selected) {
                const std::uint64_t key = j.keys[slot];
                int lo = 0, hi = n;  // first position whose key is not above `key`
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (seg[mid] > key) lo = mid + 1;
                    else hi = mid;
                }
                if (lo) atomicAdd(&j.ranks[slot], static_cast<std::uint32_t>(lo));
            }
            __syncthreads();
        });
}

__global__ void __launch_bounds__(256) s2_write(dev::JobSpan<S2Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S2Job& jj = dev::lane_job(span, tk, &S2Job::cand_begin, ji, njobs);
            const long long slot0 = (tk - jj.cand_begin) * 256;
            return slot0 < jj.candidates && slot0 < static_cast<long long>(jj.params->selected);
        },
        [&](long long t) {
        ji = dev::next_job(span, t, &S2Job::cand_begin, ji, njobs);
        const S2Job& j = dev::job_at(span, ji);
        const int slot = static_cast<int>(t - j.cand_begin) * 256 + threadIdx.x;
        if (slot >= j.candidates || static_cast<std::uint32_t>(slot) >= j.params->selected) return;
        const std::uint32_t rank = j.ranks[slot];
        if (rank >= static_cast<std::uint32_t>(j.K)) return;
        const std::uint64_t key = j.keys[slot];
        const std::uint32_t px = 0xffffffffu - static_cast<std::uint32_t>(key);
        j.out_x[rank] = static_cast<float>(px % static_cast<std::uint32_t>(j.w));
        j.out_y[rank] = static_cast<float>(px / static_cast<std::uint32_t>(j.w));
        if (j.out_resp) j.out_resp[rank] = from_order_bits(static_cast<std::uint32_t>(key >> 32));
        });
}

bool valid_image_shape(int w, int h) { return w > 0 && h > 0 && w <= INT_MAX / h; }

bool valid_config(const s2::Config& c) {
    return c.fast_threshold >= 1 && c.fast_threshold <= 254 && c.max_keypoints >= 0 &&
           c.max_keypoints <= INT_MAX / s2::kCandidateFactor && std::isfinite(c.harris_k) && c.harris_k >= 0.0f;
}

std::size_t align256(std::size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

namespace jobs {

// Offsets: maxima, hist, params, chunk_counts, chunk_offsets, keys, ranks.
std::size_t s2_layout_bytes(int w, int h, int K, std::size_t off[7]) {
    const std::size_t pixels = static_cast<std::size_t>(w) * h;
    const std::size_t chunks = (pixels + kS2Chunk - 1) / kS2Chunk;
    const std::size_t cands = static_cast<std::size_t>(s2::kCandidateFactor) * K;
    const std::size_t sizes[7] = {pixels + 16, 256 * 4, sizeof(SelectParams), 8 * chunks, 8 * chunks, 8 * cands,
                                  4 * cands};
    std::size_t cur = 0;
    for (int i = 0; i < 7; ++i) {
        cur = align256(cur);
        off[i] = cur;
        cur += sizes[i];
    }
    return cur;
}

Status make_s2_job(const ImageView& image, const s2::Config& cfg, const Keypoints& out, void* ws,
                   std::size_t ws_bytes, S2Job* job) {
    if (!image.data || image.format != PixelFormat::GRAY8 || !valid_image_shape(image.width, image.height) ||
        image.pitch_bytes < image.width || !valid_config(cfg) || !out.d_count ||
        reinterpret_cast<std::uintptr_t>(out.d_count) % alignof(int) || out.capacity < cfg.max_keypoints ||
        (cfg.max_keypoints > 0 && (!out.x || !out.y)))
        return Status::INVALID_ARGUMENT;
    S2Job j{};
    j.img = image.data;
    j.w = image.width;
    j.h = image.height;
    j.pitch = image.pitch_bytes;
    j.threshold = cfg.fast_threshold;
    j.K = cfg.max_keypoints;
    j.candidates = s2::kCandidateFactor * cfg.max_keypoints;
    j.harris_k = cfg.harris_k;
    j.out_x = out.x;
    j.out_y = out.y;
    j.out_resp = out.response;
    j.d_count = out.d_count;
    j.active = cfg.max_keypoints > 0 && image.width > 2 * s2::kEdge && image.height > 2 * s2::kEdge;
    if (j.active) {
        std::size_t off[7];
        const std::size_t need = s2_layout_bytes(j.w, j.h, j.K, off);
        if (!ws || reinterpret_cast<std::uintptr_t>(ws) % 256) return Status::INVALID_ARGUMENT;
        if (ws_bytes < need) return Status::INSUFFICIENT_WORKSPACE;
        auto* base = static_cast<std::uint8_t*>(ws);
        j.maxima = base + off[0];
        j.hist = reinterpret_cast<std::uint32_t*>(base + off[1]);
        j.params = reinterpret_cast<SelectParams*>(base + off[2]);
        j.chunk_counts = reinterpret_cast<std::uint32_t*>(base + off[3]);
        j.chunk_offsets = reinterpret_cast<std::uint32_t*>(base + off[4]);
        j.keys = reinterpret_cast<std::uint64_t*>(base + off[5]);
        j.ranks = reinterpret_cast<std::uint32_t*>(base + off[6]);
        j.tiles_x = (j.w + kS2TileW - 1) / kS2TileW;
        j.tiles = j.tiles_x * ((j.h + kS2TileH - 1) / kS2TileH);
        j.chunks = static_cast<int>((static_cast<long long>(j.w) * j.h + kS2Chunk - 1) / kS2Chunk);
    }
    *job = j;
    return Status::OK;
}

S2Totals s2_prefix(S2Job* js, int n) {
    S2Totals t{n, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        S2Job& j = js[i];
        j.fast_begin = t.fast;
        j.chunk_begin = t.chunks;
        j.cand_begin = t.cands;
        j.sort_begin = t.sort;
        j.rank_begin = t.rank;
        if (!j.active) continue;
        t.fast += j.tiles;
        t.chunks += j.chunks;
        t.cands += (j.candidates + 255) / 256;
        const long long segments = (j.candidates + kS2RankSegment - 1) / kS2RankSegment;
        t.sort += segments;
        t.rank += static_cast<long long>((j.candidates + kS2RankKeys - 1) / kS2RankKeys) * segments;
    }
    return t;
}

hipError_t launch_s2(const S2Job* table, int count, const S2Totals& t, const S2Job* single, hipStream_t stream,
                     const dev::Replay<S2Totals>* replay) {
    dev::JobSpan<S2Job> span{table, count, single ? *single : S2Job{}, replay ? replay->count : nullptr};
    const auto total = [&](long long S2Totals::*m) {
        return dev::Total(t.*m, replay ? &(replay->totals->*m) : nullptr);
    };
    s2_init<<<count, 256, 0, stream>>>(span);
    if (t.fast > 0) {
        s2_fast_nms<<<persistent_grid(t.fast, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::fast));
        s2_select<<<count, 64, 0, stream>>>(span);
        s2_count<<<persistent_grid(t.chunks, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::chunks));
        s2_scan<<<count, 256, 0, stream>>>(span);
        s2_scatter<<<persistent_grid(t.chunks, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::chunks));
        s2_harris<<<persistent_grid(t.cands, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::cands));
        s2_sort<<<persistent_grid(t.sort, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::sort));
        s2_rank<<<persistent_grid(t.rank, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::rank));
        s2_write<<<persistent_grid(t.cands, jobs::Stage::S2), 256, 0, stream>>>(span, total(&S2Totals::cands));
    }
    return hipGetLastError();
}

}  // namespace jobs

namespace s2 {

Status s2_workspace_bytes(int width, int height, const Config& config, std::size_t* bytes) {
    if (!bytes) return Status::INVALID_ARGUMENT;
    *bytes = 0;
    if (!valid_config(config) || !valid_image_shape(width, height)) return Status::INVALID_ARGUMENT;
    if (config.max_keypoints == 0) return Status::OK;
    std::size_t off[7];
    *bytes = jobs::s2_layout_bytes(width, height, config.max_keypoints, off);
    return Status::OK;
}

Status s2_detect(const ImageView& image, const Config& config, Keypoints output, void* workspace,
                 std::size_t workspace_bytes, hipStream_t stream) {
    jobs::S2Job job{};
    const Status s = jobs::make_s2_job(image, config, output, workspace, workspace_bytes, &job);
    if (s != Status::OK) return s;
    const jobs::S2Totals t = jobs::s2_prefix(&job, 1);
    return jobs::launch_s2(nullptr, 1, t, &job, stream) == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

}  // namespace s2
}  // namespace stitch
