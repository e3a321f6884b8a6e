// Multi-level detection: FAST + Harris top-K and rBRIEF on an image pyramid.
//
// The reference's S2/S3 are single-level (src/s2_s3/s2.hpp:1, s3.hpp); this
// is a build-defined extension. Level l is S1 of the input at scale
// S / f^l (area resampling from the original, so every level is as exact as
// level 0), detected and described with the single-level stages, and its
// keypoints are mapped back to level-0 work coordinates. Keypoints per level
// follow OpenCV's ORB split: a geometric series with ratio 1/f whose first
// L-1 terms are rounded and whose last level takes the remainder.
#include <hip/hip_runtime.h>

#include <cmath>

#include "stitch/pyramid.hpp"

namespace stitch::pyramid {
namespace {

constexpr int kAppendThreads = 256;
constexpr int kMinLevelSide = 2 * s2::kEdge + 2;  // smaller levels hold no detectable corner

std::size_t align256(std::size_t n) { return (n + 255) & ~std::size_t{255}; }

__global__ void clear_total(int* total, int* out_count) {
    if (threadIdx.x == 0) {
        *total = 0;
        *out_count = 0;
    }
}

// Appends one level's keypoints and descriptors at the running total, in
// level-0 coordinates: x0 = ((x + 0.5) / sx_l) * sx_0 - 0.5 in FP64.
__global__ void __launch_bounds__(kAppendThreads)
    append_level(Keypoints level, const std::uint8_t* level_desc, Keypoints out, std::uint8_t* out_desc,
                 int out_stride, std::uint8_t* out_levels, int level_index, double sx_l, double sy_l, double sx_0,
                 double sy_0, int* total) {
    __shared__ int base, n;
    if (threadIdx.x == 0) {
        base = *total;
        n = min(max(*level.d_count, 0), level.capacity);
        n = max(0, min(n, out.capacity - base));
    }
    __syncthreads();
    for (int i = static_cast<int>(threadIdx.x); i < n; i += kAppendThreads) {
        const int o = base + i;
        out.x[o] = static_cast<float>((static_cast<double>(level.x[i]) + 0.5) / sx_l * sx_0 - 0.5);
        out.y[o] = static_cast<float>((static_cast<double>(level.y[i]) + 0.5) / sy_l * sy_0 - 0.5);
        if (out.response) out.response[o] = level.response[i];
        if (out_levels) out_levels[o] = static_cast<std::uint8_t>(level_index);
    }
    // Descriptor rows, one dword per thread step.
    for (int w = static_cast<int>(threadIdx.x); w < n * 8; w += kAppendThreads) {
        const int row = w / 8, word = w % 8;
        reinterpret_cast<std::uint32_t*>(out_desc + static_cast<std::size_t>(base + row) * out_stride)[word] =
            reinterpret_cast<const std::uint32_t*>(level_desc + static_cast<std::size_t>(row) * 32)[word];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *total = base + n;
        *out.d_count = base + n;
    }
}

struct Layout {
    std::size_t gray, kp, desc, s2, s3, total, bytes;
    int pitch, height;
    std::size_t s2_bytes, s3_bytes;
};

Status layout(int w, int h, double S, const Config& c, Layout* out) {
    ScaleInfo s0{};
    if (make_scale_info(w, h, S, &s0) != Status::OK) return Status::INVALID_ARGUMENT;
    Layout L{};
    L.pitch = static_cast<int>(align256(static_cast<std::size_t>(s0.work_w)));
    L.height = s0.work_h;
    if (s2::s2_workspace_bytes(s0.work_w, s0.work_h, {c.fast_threshold, c.max_keypoints, c.harris_k}, &L.s2_bytes) !=
            Status::OK ||
        s3::s3_workspace_bytes(s0.work_w, s0.work_h, &L.s3_bytes) != Status::OK)
        return Status::INVALID_ARGUMENT;
    const std::size_t k = static_cast<std::size_t>(std::max(c.max_keypoints, 1));
    std::size_t cur = 0;
    auto take = [&](std::size_t n) {
        const std::size_t at = cur;
        cur = align256(cur + n);
        return at;
    };
    L.gray = take(static_cast<std::size_t>(L.pitch) * L.height);
    L.kp = take(3 * k * sizeof(float) + sizeof(int));
    L.desc = take(k * 32);
    L.s2 = take(L.s2_bytes);
    L.s3 = take(L.s3_bytes);
    L.total = take(sizeof(int));
    L.bytes = cur;
    *out = L;
    return Status::OK;
}

bool valid(const Config& c) {
    return c.levels >= 1 && c.levels <= kMaxLevels && std::isfinite(c.scale_factor) && c.scale_factor > 1.0 &&
           c.scale_factor <= 2.0 && c.fast_threshold >= 1 && c.fast_threshold <= 254 && c.max_keypoints >= 0 &&
           std::isfinite(c.harris_k) && c.harris_k >= 0;
}

}  // namespace

Status level_counts(const Config& c, int* counts) {
    if (!counts || !valid(c)) return Status::INVALID_ARGUMENT;
    const double ratio = 1.0 / c.scale_factor;
    double per_level = c.max_keypoints * (1.0 - ratio) / (1.0 - std::pow(ratio, c.levels));
    int sum = 0;
    for (int l = 0; l + 1 < c.levels; ++l) {
        counts[l] = static_cast<int>(std::lrint(per_level));
        sum += counts[l];
        per_level *= ratio;
    }
    counts[c.levels - 1] = std::max(c.max_keypoints - sum, 0);
    return Status::OK;
}

Status level_scale(double S, const Config& c, int level, double* scale) {
    if (!scale || !valid(c) || level < 0 || level >= c.levels || !(S > 0.0) || S > 1.0)
        return Status::INVALID_ARGUMENT;
    *scale = S / std::pow(c.scale_factor, level);
    return Status::OK;
}

Status workspace_bytes(int orig_w, int orig_h, double S, const Config& c, std::size_t* bytes) {
    if (!bytes || !valid(c)) return Status::INVALID_ARGUMENT;
    Layout L{};
    const Status s = layout(orig_w, orig_h, S, c, &L);
    if (s == Status::OK) *bytes = L.bytes;
    return s;
}

Status detect_describe(const ImageView& rgb, double S, const Config& c, Keypoints out, Descriptors out_desc,
                       std::uint8_t* out_levels, ScaleInfo* level0, void* ws, std::size_t ws_bytes,
                       hipStream_t stream) {
    if (!valid(c) || !rgb.data || rgb.format != PixelFormat::RGB8 || !out.x || !out.y || !out.d_count ||
        out.capacity < c.max_keypoints || !out_desc.data || out_desc.type != DescriptorType::BINARY_U8 ||
        out_desc.dim != 32 || out_desc.row_stride_bytes < 32 || out_desc.row_stride_bytes % 4 ||
        reinterpret_cast<std::uintptr_t>(out_desc.data) % 4)
        return Status::INVALID_ARGUMENT;
    Layout L{};
    if (layout(rgb.width, rgb.height, S, c, &L) != Status::OK) return Status::INVALID_ARGUMENT;
    if (!ws || ws_bytes < L.bytes) return Status::INSUFFICIENT_WORKSPACE;
    if (reinterpret_cast<std::uintptr_t>(ws) % 256) return Status::INVALID_ARGUMENT;
    auto* base = static_cast<std::uint8_t*>(ws);
    int counts[kMaxLevels];
    level_counts(c, counts);
    ScaleInfo s0{};
    make_scale_info(rgb.width, rgb.height, S, &s0);
    if (level0) *level0 = s0;
    int* total = reinterpret_cast<int*>(base + L.total);
    clear_total<<<1, 64, 0, stream>>>(total, out.d_count);
    if (hipGetLastError() != hipSuccess) return Status::CUDA_ERROR;
    const std::size_t k = static_cast<std::size_t>(std::max(c.max_keypoints, 1));
    Keypoints tmp{};
    tmp.x = reinterpret_cast<float*>(base + L.kp);
    tmp.y = tmp.x + k;
    tmp.response = tmp.y + k;
    tmp.d_count = reinterpret_cast<int*>(tmp.response + k);
    Descriptors tmp_desc{DescriptorType::BINARY_U8, 32, 32, base + L.desc};
    for (int l = 0; l < c.levels; ++l) {
        double scale = 0;
        ScaleInfo sl{};
        level_scale(S, c, l, &scale);
        if (counts[l] == 0 || make_scale_info(rgb.width, rgb.height, scale, &sl) != Status::OK ||
            sl.work_w < kMinLevelSide || sl.work_h < kMinLevelSide)
            continue;
        ImageView gray{base + L.gray, sl.work_w, sl.work_h, L.pitch, PixelFormat::GRAY8};
        tmp.capacity = counts[l];
        Status s = s1::s1_preprocess_single(rgb, scale, &gray, &sl, stream);
        if (s == Status::OK)
            s = s2::s2_detect(gray, {c.fast_threshold, counts[l], c.harris_k}, tmp, base + L.s2, L.s2_bytes, stream);
        if (s == Status::OK) s = s3::s3_describe(gray, tmp, tmp_desc, base + L.s3, L.s3_bytes, stream);
        if (s != Status::OK) return s;
        append_level<<<1, kAppendThreads, 0, stream>>>(tmp, base + L.desc, out,
                                                       static_cast<std::uint8_t*>(out_desc.data),
                                                       out_desc.row_stride_bytes, out_levels, l, sl.sx, sl.sy, s0.sx,
                                                       s0.sy, total);
        if (hipGetLastError() != hipSuccess) return Status::CUDA_ERROR;
    }
    return Status::OK;
}

}  // namespace stitch::pyramid
