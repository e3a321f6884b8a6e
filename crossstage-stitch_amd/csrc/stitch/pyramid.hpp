// Multi-level detection and description (an image pyramid over S1-S3).
// Build-defined extension: the reference detects on one level only
// (src/s2_s3/s2.hpp:1). Level l is S1 of the input at scale S / f^l; each
// level runs the single-level S2 and S3 with its share of K; keypoints come
// back in level-0 work coordinates, ordered by level, then as S2 ranks them.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "stitch/stages.hpp"

namespace stitch::pyramid {

constexpr int kMaxLevels = 16;

struct Config {
    int levels;           // 1..16
    double scale_factor;  // f in (1, 2]
    int fast_threshold;   // as s2::Config
    int max_keypoints;    // K over all levels
    float harris_k;
};

// Keypoints per level: K(1 - 1/f) / (1 - f^-L) * f^-l rounded to nearest
// even for the first L-1 levels, the remainder (at least 0) for the last.
Status level_counts(const Config& config, int* counts);
// S / f^level.
Status level_scale(double S, const Config& config, int level, double* scale);
Status workspace_bytes(int orig_w, int orig_h, double S, const Config& config, std::size_t* bytes);
// rgb: device RGB8. Output keypoints (capacity >= K), descriptors (32-byte
// rows), optional per-keypoint level (K bytes) and the level-0 scale info.
// Levels smaller than 64 px on a side are skipped.
Status detect_describe(const ImageView& rgb, double S, const Config& config, Keypoints keypoints,
                       Descriptors descriptors, std::uint8_t* levels, ScaleInfo* level0, void* workspace,
                       std::size_t workspace_bytes, hipStream_t stream);

}  // namespace stitch::pyramid
