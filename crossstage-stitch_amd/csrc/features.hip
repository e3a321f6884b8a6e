// Controller observations (scene and gate features) on the GPU.
//
// Every statistic is an FP64 sum whose value depends on its summation order,
// and the reference's order is part of the feature definition: each of N
// threads (128, or 64 for the patches) sums the elements i = t, t + N, ...
// in increasing i, then the N partials are added pairwise in a shared-memory
// tree (t += t + N/2, then N/4, ... 1). The kernels here keep that partition
// and tree (block_sum) so the values can be checked against the CPU oracle.
// Semantics: src/controller/features.cu:11-171 and
// src/controller/relational_features.cu:6-190.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "stitch/features.hpp"

namespace stitch::features {
namespace {

constexpr int kThreads = 128;
constexpr int kPatchThreads = 64;
constexpr int kShifts = 49;  // 7 x 7 shifts of 1/8 of the image
constexpr int kPatches = 16;  // 4 x 4 patch centres

// Gray level in [0, 1] (the controller's own conversion, in FP64).
__device__ double gray_of(const ImageView& im, int x, int y) {
    const std::uint8_t* p =
        im.data + static_cast<std::size_t>(y) * im.pitch_bytes + static_cast<std::size_t>(x) * (im.format == PixelFormat::RGB8 ? 3 : 1);
    if (im.format == PixelFormat::GRAY8) return p[0] / 255.0;
    return (9798.0 * p[0] + 19235.0 * p[1] + 3735.0 * p[2]) / (32768.0 * 255.0);
}

// Bilinear gray at (x, y) >= 0: truncated corner, right/bottom neighbour
// clamped to the image.
__device__ double bilinear(const ImageView& im, double x, double y) {
    const int ix = static_cast<int>(x), iy = static_cast<int>(y);
    const int jx = min(ix + 1, im.width - 1), jy = min(iy + 1, im.height - 1);
    const double dx = x - ix, dy = y - iy;
    return (1 - dy) * ((1 - dx) * gray_of(im, ix, iy) + dx * gray_of(im, jx, iy)) +
           dy * ((1 - dx) * gray_of(im, ix, jy) + dx * gray_of(im, jx, jy));
}

// Pairwise tree over the first N threads' values (threads beyond N, in a
// wider block, take no part); every thread gets the total.
template <int N>
__device__ double block_sum(double v, double* scratch) {
    const int t = static_cast<int>(threadIdx.x);
    if (t < N) scratch[t] = v;
    __syncthreads();
    for (int step = N / 2; step > 0; step /= 2) {
        if (t < step) scratch[t] += scratch[t + step];
        __syncthreads();
    }
    const double r = scratch[0];
    __syncthreads();
    return r;
}

__device__ double ncc(double a, double b, double aa, double bb, double ab, double n) {
    const double va = fmax(0.0, aa - a * a / n), vb = fmax(0.0, bb - b * b / n);
    return fmin(1.0, fmax(-1.0, (ab - a * b / n) / fmax(1e-12, sqrt(va * vb))));
}

struct Mat3 {
    double v[9];
};

// B position of A's (x, y) through the inverse homography, and its gray.
__device__ bool gray_through(const ImageView& b, const Mat3& m, double x, double y, double* value) {
    const double z = m.v[6] * x + m.v[7] * y + m.v[8];
    if (!isfinite(z) || fabs(z) < 1e-12) return false;
    const double u = (m.v[0] * x + m.v[1] * y + m.v[2]) / z, v = (m.v[3] * x + m.v[4] * y + m.v[5]) / z;
    if (!isfinite(u) || !isfinite(v) || u < 0 || v < 0 || u > b.width - 1 || v > b.height - 1) return false;
    *value = bilinear(b, u, v);
    return true;
}

__global__ void fill_zero(double* p, int n) {
    for (int i = static_cast<int>(threadIdx.x); i < n; i += static_cast<int>(blockDim.x)) p[i] = 0.0;
}

// Blocks 0 and 1: the mean absolute gradient of A and B. Blocks 2-9: per
// image (A then B) and quadrant (row-major), the mean, standard deviation
// and gradient of the gray levels (features.cu:34-61). Pixels are read by
// all 1024 threads, 1024 at a time, into LDS; threads 0..127 then add them
// in the order of one thread per partial walking i = t, t + 128, ...
constexpr int kSceneThreads = 1024;

__global__ void __launch_bounds__(kSceneThreads) scene_stats(ImageView a, ImageView b, double* out) {
    __shared__ double value[kSceneThreads], right[kSceneThreads], down[kSceneThreads], scratch[kThreads];
    __shared__ unsigned char kind[kSceneThreads];  // bit 0: pixel, bit 1: right neighbour, bit 2: lower neighbour
    const int t = static_cast<int>(threadIdx.x);
    const int region = static_cast<int>(blockIdx.x);
    const bool whole = region < 2;
    const int image = whole ? region : (region - 2) / 4, quad = whole ? 0 : (region - 2) % 4;
    const ImageView& im = image ? b : a;
    const int x0 = whole ? 0 : (quad % 2) * im.width / 2, y0 = whole ? 0 : (quad / 2) * im.height / 2;
    const int x1 = whole ? im.width : (quad % 2 + 1) * im.width / 2;
    const int y1 = whole ? im.height : (quad / 2 + 1) * im.height / 2;
    const int w = x1 - x0, h = y1 - y0, count = w * h;
    double sum = 0, sq = 0, grad = 0;
    for (int round = 0; round < count; round += kSceneThreads) {
        {
            const int i = round + t;
            unsigned char k = 0;
            if (i < count) {
                const int x = x0 + i % w, y = y0 + i / w;
                const double v = gray_of(im, x, y);
                value[t] = v;
                k = 1;
                if (x + 1 < x1) {
                    right[t] = fabs(v - gray_of(im, x + 1, y));
                    k |= 2;
                }
                if (y + 1 < y1) {
                    down[t] = fabs(v - gray_of(im, x, y + 1));
                    k |= 4;
                }
            }
            kind[t] = k;
        }
        __syncthreads();
        if (t < kThreads)
            for (int j = t; j < kSceneThreads; j += kThreads) {
                const unsigned char k = kind[j];
                if (!k) continue;
                const double v = value[j];
                sum += v;
                sq += v * v;
                if (k & 2) grad += right[j];
                if (k & 4) grad += down[j];
            }
        __syncthreads();
    }
    sum = block_sum<kThreads>(sum, scratch);
    sq = block_sum<kThreads>(sq, scratch);
    grad = block_sum<kThreads>(grad, scratch);
    if (t == 0) {
        const double pixels = static_cast<double>(count), edges = static_cast<double>(h * (w - 1) + w * (h - 1));
        const double g = edges != 0 ? grad / edges : 0.0;
        if (whole) {
            out[region] = g;
        } else {
            const double mean = sum / pixels;
            double* o = out + 2 + image * 12 + quad * 3;
            o[0] = mean;
            o[1] = sqrt(fmax(0.0, sq / pixels - mean * mean));
            o[2] = g;
        }
    }
}

// One block per shift (dx, dy) in {-3..3}/8 of the image: NCC of A's 32x32
// grid against B's grid shifted by it, and the share of grid points inside B
// (relational_features.cu:27-43).
__global__ void __launch_bounds__(kThreads) pair_shifts(ImageView a, ImageView b, double* workspace) {
    __shared__ double scratch[kThreads];
    const int shift = static_cast<int>(blockIdx.x);
    const double dx = (shift % 7 - 3) / 8.0, dy = (shift / 7 - 3) / 8.0;
    double n = 0, sa = 0, sb = 0, aa = 0, bb = 0, ab = 0;
    for (int i = static_cast<int>(threadIdx.x); i < 1024; i += kThreads) {
        const double x = (i % 32) / 31.0, y = (i / 32) / 31.0, u = x + dx, v = y + dy;
        if (u < 0 || u > 1 || v < 0 || v > 1) continue;
        const double av = bilinear(a, x * (a.width - 1), y * (a.height - 1));
        const double bv = bilinear(b, u * (b.width - 1), v * (b.height - 1));
        ++n;
        sa += av;
        sb += bv;
        aa += av * av;
        bb += bv * bv;
        ab += av * bv;
    }
    n = block_sum<kThreads>(n, scratch);
    sa = block_sum<kThreads>(sa, scratch);
    sb = block_sum<kThreads>(sb, scratch);
    aa = block_sum<kThreads>(aa, scratch);
    bb = block_sum<kThreads>(bb, scratch);
    ab = block_sum<kThreads>(ab, scratch);
    if (threadIdx.x == 0) {
        workspace[shift] = ncc(sa, sb, aa, bb, ab, n);
        workspace[kShifts + shift] = n / 1024.0;
    }
}

// Peak NCC (first maximum), its margin over the best other shift, the peak's
// shift, the zero-shift NCC and the peak's overlap (relational_features.cu:44-50).
__global__ void pair_peak(const double* workspace, double* out) {
    int best = 0;
    for (int s = 1; s < kShifts; ++s)
        if (workspace[s] > workspace[best]) best = s;
    double second = -2;
    for (int s = 0; s < kShifts; ++s)
        if (s != best) second = fmax(second, workspace[s]);
    out[0] = workspace[best];
    out[1] = workspace[best] - second;
    out[2] = (best % 7 - 3) / 8.0;
    out[3] = (best / 7 - 3) / 8.0;
    out[4] = workspace[kShifts / 2];
    out[5] = workspace[kShifts + best];
}

// A 64x64 grid over A through the inverse homography into B: overlap share
// and photometric agreement (features.cu:65-91). The samples are evaluated
// by all 1024 threads, 1024 at a time, into LDS; then threads 0..127 add
// them to their partial sums in the same order as one thread per partial
// walking i = t, t + 128, ... would (the sums are unchanged).
constexpr int kPhotoThreads = 1024;
constexpr int kPhotoSamples = 4096;

__global__ void __launch_bounds__(kPhotoThreads) photo_consistency(ImageView a, ImageView b, Mat3 inverse, double* out) {
    __shared__ double sample_a[kPhotoThreads], sample_b[kPhotoThreads], scratch[kThreads];
    __shared__ bool sample_ok[kPhotoThreads];
    const int t = static_cast<int>(threadIdx.x);
    double n = 0, sum_a = 0, sum_b = 0, sq_a = 0, sq_b = 0, cross = 0, abs_error = 0, sq_error = 0;
    for (int round = 0; round < kPhotoSamples; round += kPhotoThreads) {
        {
            const int i = round + t;
            const double x = (i % 64) * (a.width - 1.0) / 63.0, y = (i / 64) * (a.height - 1.0) / 63.0;
            const double* m = inverse.v;
            const double z = m[6] * x + m[7] * y + m[8];
            bool ok = !(fabs(z) < 1e-12);
            double u = 0, v = 0;
            if (ok) {
                u = (m[0] * x + m[1] * y + m[2]) / z;
                v = (m[3] * x + m[4] * y + m[5]) / z;
                ok = !(u < 0 || v < 0 || u > b.width - 1 || v > b.height - 1 || !isfinite(u) || !isfinite(v));
            }
            sample_ok[t] = ok;
            if (ok) {
                sample_a[t] = bilinear(a, fmax(0.0, x), fmax(0.0, y));
                sample_b[t] = bilinear(b, u, v);
            }
        }
        __syncthreads();
        if (t < kThreads)
            for (int k = t; k < kPhotoThreads; k += kThreads) {
                if (!sample_ok[k]) continue;
                const double av = sample_a[k], bv = sample_b[k], d = av - bv;
                n += 1;
                sum_a += av;
                sum_b += bv;
                sq_a += av * av;
                sq_b += bv * bv;
                cross += av * bv;
                abs_error += fabs(d);
                sq_error += d * d;
            }
        __syncthreads();
    }
    // Threads 128..1023 hold zeros and take no part in the tree below.
    n = block_sum<kThreads>(n, scratch);
    sum_a = block_sum<kThreads>(sum_a, scratch);
    sum_b = block_sum<kThreads>(sum_b, scratch);
    sq_a = block_sum<kThreads>(sq_a, scratch);
    sq_b = block_sum<kThreads>(sq_b, scratch);
    cross = block_sum<kThreads>(cross, scratch);
    abs_error = block_sum<kThreads>(abs_error, scratch);
    sq_error = block_sum<kThreads>(sq_error, scratch);
    if (t == 0) {
        const double den = fmax(1.0, n), ma = sum_a / den, mb = sum_b / den;
        const double va = fmax(0.0, sq_a / den - ma * ma), vb = fmax(0.0, sq_b / den - mb * mb);
        const double sd = sqrt(va) + sqrt(vb);
        out[10] = n / 4096.0;
        out[11] = abs_error / den;
        out[12] = sqrt(sq_error / den);
        out[13] = fmin(1.0, fmax(-1.0, (cross / den - ma * mb) / fmax(1e-12, sqrt(va * vb))));
        out[14] = fabs(ma - mb);
        out[15] = sqrt(va);
        out[16] = sqrt(vb);
        out[17] = fmin(1e6, abs_error / den / fmax(1e-6, sd));
    }
}

// Inliers binned on a 4x4 grid of each image: occupied share, largest bin
// share and normalised entropy (features.cu:92-114). Bins use the FP32
// keypoint coordinates as the reference does.
__global__ void __launch_bounds__(kThreads) inlier_spread(Keypoints a, Keypoints b, Matches m, const std::uint8_t* mask,
                                                          int count, int wa, int ha, int wb, int hb, double* out) {
    __shared__ int hist[32];
    if (threadIdx.x < 32) hist[threadIdx.x] = 0;
    __syncthreads();
    for (int i = static_cast<int>(threadIdx.x); i < count; i += kThreads) {
        if (!mask[i]) continue;
        const int ia = m.idx_a[i], ib = m.idx_b[i];
        if (ia < 0 || ib < 0 || ia >= a.capacity || ib >= b.capacity) continue;
        const int xa = max(0, min(3, static_cast<int>(4 * a.x[ia] / wa)));
        const int ya = max(0, min(3, static_cast<int>(4 * a.y[ia] / ha)));
        const int xb = max(0, min(3, static_cast<int>(4 * b.x[ib] / wb)));
        const int yb = max(0, min(3, static_cast<int>(4 * b.y[ib] / hb)));
        atomicAdd(&hist[ya * 4 + xa], 1);
        atomicAdd(&hist[16 + yb * 4 + xb], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int image = 0; image < 2; ++image) {
            const int* hi = hist + image * 16;
            int total = 0, occupied = 0, largest = 0;
            for (int k = 0; k < 16; ++k) {
                total += hi[k];
                occupied += hi[k] > 0;
                largest = max(largest, hi[k]);
            }
            double entropy = 0;
            if (total)
                for (int k = 0; k < 16; ++k)
                    if (hi[k]) {
                        const double p = static_cast<double>(hi[k]) / total;
                        entropy -= p * log(p) / log(16.0);
                    }
            out[18 + image] = occupied / 16.0;
            out[20 + image] = total ? static_cast<double>(largest) / total : 0.0;
            out[22 + image] = entropy;
        }
    }
}

// Residual, regional, match-distance and forward-ratio statistics
// (relational_features.cu:99-153): integer histograms by atomics, FP64 sums
// in the fixed partition and tree.
__global__ void __launch_bounds__(kThreads) alignment_statistics(Keypoints a, Keypoints b, Matches m,
                                                                 const std::uint8_t* mask, HomographyResult h,
                                                                 ScaleInfo sa, ScaleInfo sb, const s4::Top2* forward,
                                                                 double* out, unsigned groups) {
    __shared__ int residual_hist[257], distance_hist[257], counts[16];
    __shared__ double regional[16][kThreads], scratch[kThreads];
    const int t = static_cast<int>(threadIdx.x);
    for (int k = t; k < 257; k += kThreads) residual_hist[k] = distance_hist[k] = 0;
    if (t < 16) counts[t] = 0;
    for (int k = 0; k < 16; ++k) regional[k][t] = 0;
    __syncthreads();
    int inliers = 0;
    double distance_sum = 0;
    for (int i = t; i < h.num_matches; i += kThreads) {
        const int ia = m.idx_a[i], ib = m.idx_b[i];
        if (ia < 0 || ib < 0 || ia >= a.capacity || ib >= b.capacity) continue;
        if (groups & kMatchingRelations) {
            const double distance = fmin(256.0, fmax(0.0, static_cast<double>(m.distance[i])));
            distance_sum += distance;
            atomicAdd(&distance_hist[static_cast<int>(distance)], 1);
        }
        if (!(groups & kGeometryRelations) || !mask[i]) continue;
        const double x = (b.x[ib] + 0.5) / sb.sx - 0.5, y = (b.y[ib] + 0.5) / sb.sy - 0.5;
        const double z = h.H[6] * x + h.H[7] * y + h.H[8];
        if (!isfinite(z) || fabs(z) < 1e-12) continue;
        const double u = (h.H[0] * x + h.H[1] * y + h.H[2]) / z, v = (h.H[3] * x + h.H[4] * y + h.H[5]) / z;
        const double e = hypot(u - ((a.x[ia] + 0.5) / sa.sx - 0.5), v - ((a.y[ia] + 0.5) / sa.sy - 0.5));
        if (!isfinite(e)) continue;
        atomicAdd(&residual_hist[static_cast<int>(fmin(32.0, e) * 8)], 1);
        ++inliers;
        const int tile = max(0, min(3, static_cast<int>(a.x[ia] * 4 / sa.work_w))) +
                         4 * max(0, min(3, static_cast<int>(a.y[ia] * 4 / sa.work_h)));
        atomicAdd(&counts[tile], 1);
        regional[tile][t] += fmin(1e6, e);
    }
    distance_sum = block_sum<kThreads>(distance_sum, scratch);
    inliers = static_cast<int>(block_sum<kThreads>(inliers, scratch));
    double maximum = 0, sum = 0, sq = 0;
    int regions = 0;
    if (groups & kGeometryRelations)
        for (int k = 0; k < 16; ++k) {
            const double total = block_sum<kThreads>(regional[k][t], scratch);
            if (t == 0 && counts[k]) {
                const double v = total / counts[k];
                maximum = fmax(maximum, v);
                sum += v;
                sq += v * v;
                ++regions;
            }
        }
    double ratio_sum = 0;
    int ratios = 0, ambiguous = 0;
    if (forward && (groups & kMatchingRelations))
        for (int k = t; k < a.capacity; k += kThreads) {
            const s4::Top2 p = forward[k];
            if (p.first_index < 0 || p.second_index < 0 || !isfinite(p.second_distance)) continue;
            const double ratio = p.second_distance > 0 ? p.first_distance / p.second_distance : 1.0;
            ratio_sum += ratio;
            ++ratios;
            ambiguous += ratio >= 0.8;
        }
    ratio_sum = block_sum<kThreads>(ratio_sum, scratch);
    ratios = static_cast<int>(block_sum<kThreads>(ratios, scratch));
    ambiguous = static_cast<int>(block_sum<kThreads>(ambiguous, scratch));
    if (t == 0) {
        int seen = 0, p50 = 256, p90 = 256, dp90 = 256;
        for (int k = 0; k < 257; ++k) {
            seen += residual_hist[k];
            if (seen >= ceil(0.5 * inliers) && p50 == 256) p50 = k;
            if (seen >= ceil(0.9 * inliers)) {
                p90 = k;
                break;
            }
        }
        seen = 0;
        for (int k = 0; k < 257; ++k) {
            seen += distance_hist[k];
            if (seen >= ceil(0.9 * h.num_matches)) {
                dp90 = k;
                break;
            }
        }
        out[0] = inliers ? p50 / 8.0 : 32;
        out[1] = inliers ? p90 / 8.0 : 32;
        out[2] = maximum;
        out[3] = regions ? sqrt(fmax(0.0, sq / regions - (sum / regions) * (sum / regions))) : 0;
        out[4] = distance_sum / fmax(1.0, static_cast<double>(h.num_matches)) / 256.0;
        out[5] = dp90 / 256.0;
        out[6] = ratios ? ratio_sum / ratios : 1;
        out[7] = ratios ? static_cast<double>(ambiguous) / ratios : 1;
    }
}

// 7x7 patches at the centres of a 4x4 grid over A against B through the
// inverse homography: SSIM-like score, edge error and cosine, normalised
// MAE, and whether all 49 points were usable (relational_features.cu:67-89).
__global__ void __launch_bounds__(kPatchThreads) patch_similarity(ImageView a, ImageView b, Mat3 inverse,
                                                                  double* workspace) {
    __shared__ double scratch[kPatchThreads];
    const int patch = static_cast<int>(blockIdx.x);
    const double cx = (patch % 4 + 0.5) * a.width / 4.0, cy = (patch / 4 + 0.5) * a.height / 4.0;
    double n = 0, sa = 0, sb = 0, aa = 0, bb = 0, ab = 0, mae = 0, edge_error = 0, dot = 0, ga2 = 0, gb2 = 0;
    for (int i = static_cast<int>(threadIdx.x); i < 49; i += kPatchThreads) {
        const int dx = i % 7 - 3, dy = i / 7 - 3;
        const double x = cx + dx, y = cy + dy;
        double bv, bx, by;
        if (x < 0 || y < 0 || x + 1 > a.width - 1 || y + 1 > a.height - 1 || !gray_through(b, inverse, x, y, &bv) ||
            !gray_through(b, inverse, x + 1, y, &bx) || !gray_through(b, inverse, x, y + 1, &by))
            continue;
        const double av = bilinear(a, x, y), ax = bilinear(a, x + 1, y) - av, ay = bilinear(a, x, y + 1) - av;
        bx -= bv;
        by -= bv;
        ++n;
        sa += av;
        sb += bv;
        aa += av * av;
        bb += bv * bv;
        ab += av * bv;
        mae += fabs(av - bv);
        edge_error += fabs(ax - bx) + fabs(ay - by);
        dot += ax * bx + ay * by;
        ga2 += ax * ax + ay * ay;
        gb2 += bx * bx + by * by;
    }
    n = block_sum<kPatchThreads>(n, scratch);
    sa = block_sum<kPatchThreads>(sa, scratch);
    sb = block_sum<kPatchThreads>(sb, scratch);
    aa = block_sum<kPatchThreads>(aa, scratch);
    bb = block_sum<kPatchThreads>(bb, scratch);
    ab = block_sum<kPatchThreads>(ab, scratch);
    mae = block_sum<kPatchThreads>(mae, scratch);
    edge_error = block_sum<kPatchThreads>(edge_error, scratch);
    dot = block_sum<kPatchThreads>(dot, scratch);
    ga2 = block_sum<kPatchThreads>(ga2, scratch);
    gb2 = block_sum<kPatchThreads>(gb2, scratch);
    if (threadIdx.x == 0) {
        const double den = fmax(n, 1.0), ma = sa / den, mb = sb / den;
        const double va = fmax(0.0, aa / den - ma * ma), vb = fmax(0.0, bb / den - mb * mb);
        double* w = workspace + patch * 5;
        w[0] = ((2 * ma * mb + 0.0001) * (2 * (ab / den - ma * mb) + 0.0009)) /
               ((ma * ma + mb * mb + 0.0001) * (va + vb + 0.0009));
        w[1] = edge_error / (2 * den);
        w[2] = dot / fmax(1e-12, sqrt(ga2 * gb2));
        w[3] = fmin(1e6, mae / den / fmax(1e-6, sqrt(va) + sqrt(vb)));
        w[4] = n == 49;
    }
}

// Over the fully usable patches: mean and minimum score, mean edge error and
// cosine, usable share, and the 90th percentile normalised MAE
// (relational_features.cu:90-98).
__global__ void patch_summary(const double* workspace, double* out) {
    int n = 0;
    double sum = 0, low = 1, edge = 0, cosine = 0, errors[kPatches];
    for (int p = 0; p < kPatches; ++p) {
        const double* w = workspace + p * 5;
        if (!w[4]) continue;
        errors[n++] = w[3];
        sum += w[0];
        low = fmin(low, w[0]);
        edge += w[1];
        cosine += w[2];
    }
    for (int i = 1; i < n; ++i) {  // insertion sort, ascending
        const double v = errors[i];
        int k = i;
        while (k && errors[k - 1] > v) {
            errors[k] = errors[k - 1];
            --k;
        }
        errors[k] = v;
    }
    out[8] = n ? sum / n : -1;
    out[9] = n ? low : -1;
    out[10] = n ? edge / n : 1;
    out[11] = n ? cosine / n : 0;
    out[12] = n / 16.0;
    out[13] = n ? errors[static_cast<int>(ceil(0.9 * n)) - 1] : 1e6;
}

// Adjugate over the determinant (features.cu:115-124).
bool invert(const double* h, Mat3* out) {
    const double d = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                     h[2] * (h[3] * h[7] - h[4] * h[6]);
    if (!std::isfinite(d) || std::abs(d) < 1e-18) return false;
    const double adj[9] = {h[4] * h[8] - h[5] * h[7], h[2] * h[7] - h[1] * h[8], h[1] * h[5] - h[2] * h[4],
                           h[5] * h[6] - h[3] * h[8], h[0] * h[8] - h[2] * h[6], h[2] * h[3] - h[0] * h[5],
                           h[3] * h[7] - h[4] * h[6], h[1] * h[6] - h[0] * h[7], h[0] * h[4] - h[1] * h[3]};
    for (int i = 0; i < 9; ++i) {
        out->v[i] = adj[i] / d;
        if (!std::isfinite(out->v[i])) return false;
    }
    return true;
}

bool usable(const ImageView& im) {
    return im.data && im.width >= 2 && im.height >= 2 && im.pitch_bytes > 0 &&
           (im.format == PixelFormat::GRAY8 || im.format == PixelFormat::RGB8) &&
           im.pitch_bytes >= im.width * (im.format == PixelFormat::RGB8 ? 3 : 1);
}

Status launched() { return hipGetLastError() == hipSuccess ? Status::OK : Status::CUDA_ERROR; }

}  // namespace

Status extract_scene(const ImageView& a, const ImageView& b, double* device, double* host, hipStream_t stream,
                     bool relations) {
    if (!a.data || !b.data || !device || !host || a.width < 2 || b.width < 2 || a.height < 2 || b.height < 2 ||
        a.format != PixelFormat::GRAY8 || b.format != PixelFormat::GRAY8)
        return Status::INVALID_ARGUMENT;
    scene_stats<<<10, kSceneThreads, 0, stream>>>(a, b, device);
    fill_zero<<<1, 64, 0, stream>>>(device + 26, kPairRelations);
    if (launched() != Status::OK) return Status::CUDA_ERROR;
    if (relations) {
        if (!usable(a) || !usable(b)) return Status::CUDA_ERROR;
        pair_shifts<<<kShifts, kThreads, 0, stream>>>(a, b, device + kGateCount);
        pair_peak<<<1, 1, 0, stream>>>(device + kGateCount, device + 26);
        if (launched() != Status::OK) return Status::CUDA_ERROR;
    }
    if (hipMemcpyAsync(host, device, kSceneCount * sizeof(double), hipMemcpyDeviceToHost, stream) != hipSuccess)
        return Status::CUDA_ERROR;
    return Status::OK;
}

Status extract_gate(const ImageView& a, const ImageView& b, const Keypoints& ka, const Keypoints& kb,
                    const Matches& matches, const std::uint8_t* mask, const HomographyResult& h, const ScaleInfo& sa,
                    const ScaleInfo& sb, double* device, double* host, hipStream_t stream, const s4::Top2* forward,
                    unsigned relation_mask) {
    Mat3 inv{};
    if (!a.data || !b.data || !device || !host || !mask || h.status != Status::OK || h.num_matches < 4 ||
        h.num_matches > matches.capacity || !invert(h.H, &inv))
        return Status::INVALID_ARGUMENT;
    // Alignment-level features from the homography alone (features.cu:140-158).
    double base[10] = {h.num_matches / 4000.0, h.num_inliers / 4000.0, h.inlier_ratio, h.inlier_coverage,
                       h.reproj_residual, h.reproj_residual / std::max(sa.sx, sa.sy), 0, 0, 0, 0};
    double x[4], y[4], zmin = 1e300, zmax = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double u = (i == 1 || i == 2) ? b.width - 1.0 : 0.0, v = i >= 2 ? b.height - 1.0 : 0.0;
        const double z = h.H[6] * u + h.H[7] * v + h.H[8];
        if (!std::isfinite(z) || std::abs(z) < 1e-12) return Status::INVALID_ARGUMENT;
        x[i] = (h.H[0] * u + h.H[1] * v + h.H[2]) / z;
        y[i] = (h.H[3] * u + h.H[4] * v + h.H[5]) / z;
        if (!std::isfinite(x[i]) || !std::isfinite(y[i])) return Status::INVALID_ARGUMENT;
        zmin = std::min(zmin, std::abs(z));
        zmax = std::max(zmax, std::abs(z));
    }
    double area = 0.0;
    for (int i = 0; i < 4; ++i) area += x[i] * y[(i + 1) % 4] - y[i] * x[(i + 1) % 4];
    base[6] = std::min(1e6, std::abs(area) / 2.0 / (static_cast<double>(b.width) * b.height));
    base[7] = std::min(1e6, (*std::max_element(x, x + 4) - *std::min_element(x, x + 4)) / a.width);
    base[8] = std::min(1e6, (*std::max_element(y, y + 4) - *std::min_element(y, y + 4)) / a.height);
    base[9] = std::min(1e6, zmax / zmin);
    // The alignment relations take the reference's input limits; inputs
    // they reject fail the whole call (relational_features.cu:169-176).
    if (relation_mask) {
        const bool ok = usable(a) && usable(b) && matches.idx_a && matches.idx_b && matches.distance && ka.x &&
                        ka.y && kb.x && kb.y && ka.capacity >= 1 && ka.capacity <= 4000 && kb.capacity >= 1 &&
                        kb.capacity <= 4000 && std::isfinite(sa.sx) && std::isfinite(sa.sy) &&
                        std::isfinite(sb.sx) && std::isfinite(sb.sy) && sa.sx > 0 && sa.sy > 0 && sb.sx > 0 &&
                        sb.sy > 0 && sa.work_w > 0 && sa.work_h > 0 && relation_mask <= 7;
        if (!ok) return Status::CUDA_ERROR;
    }
    photo_consistency<<<1, kPhotoThreads, 0, stream>>>(a, b, inv, device);
    inlier_spread<<<1, kThreads, 0, stream>>>(ka, kb, matches, mask, h.num_matches, sa.work_w, sa.work_h, sb.work_w,
                                              sb.work_h, device);
    fill_zero<<<1, 64, 0, stream>>>(device + 24, kAlignmentRelations);
    if (relation_mask & (kMatchingRelations | kGeometryRelations))
        alignment_statistics<<<1, kThreads, 0, stream>>>(ka, kb, matches, mask, h, sa, sb, forward, device + 24,
                                                         relation_mask);
    if (relation_mask & kPhotometricRelations) {
        patch_similarity<<<kPatches, kPatchThreads, 0, stream>>>(a, b, inv, device + kGateCount);
        patch_summary<<<1, 1, 0, stream>>>(device + kGateCount, device + 24);
    }
    if (launched() != Status::OK) return Status::CUDA_ERROR;
    std::copy(base, base + 10, host);
    if (hipMemcpyAsync(host + 10, device + 10, (kGateCount - 10) * sizeof(double), hipMemcpyDeviceToHost, stream) !=
        hipSuccess)
        return Status::CUDA_ERROR;
    return Status::OK;
}

}  // namespace stitch::features
