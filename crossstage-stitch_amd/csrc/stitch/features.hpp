// Controller observations: scene features of a pair (before matching) and
// gate features of an attempt (after S5), as the reference's runner collects
// them for its features CSV and learned policies. Observations only: nothing
// here changes S1-S6. Layout and semantics: src/controller/features.hpp:10-32,
// features.cu:11-171, relational_features.hpp:8-24, relational_features.cu:6-190.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>

#include "stitch/stages.hpp"

namespace stitch::features {

constexpr int kSceneCount = 32;          // 2 gradients, 2 x 4 quadrants x (mean, std, gradient), 6 pair relations
constexpr int kGateCount = 38;           // 10 alignment, 8 photometric, 6 inlier spread, 14 alignment relations
constexpr int kPairRelations = 6;
constexpr int kAlignmentRelations = 14;
constexpr int kRelationWorkspace = 98;   // 49 shifts x (NCC, overlap); 16 patches x 5 fit in it
constexpr int kDeviceDoubles = kGateCount + kRelationWorkspace;
constexpr unsigned kMatchingRelations = 1, kGeometryRelations = 2, kPhotometricRelations = 4;

using SceneFeatures = std::array<double, kSceneCount>;
using GateFeatures = std::array<double, kGateCount>;

// Scene features of two GRAY8 images (the reference collects them at S = 0.25).
// `device` holds kDeviceDoubles doubles; `host` (kSceneCount doubles, pinned)
// is written by an async copy on `stream` - valid after the stream is synced.
Status extract_scene(const ImageView& a, const ImageView& b, double* device, double* host, hipStream_t stream,
                     bool relations);

// Gate features of a finished S5: the original RGB8 (or GRAY8) images, the
// attempt's keypoints, matches, inlier mask and homography (host copy), and
// optionally S4's forward top-2 view. Invalid inputs (no usable homography,
// fewer than 4 matches) return INVALID_ARGUMENT and enqueue nothing; inputs
// the alignment relations reject (keypoint capacity above 4000, as in the
// reference) return CUDA_ERROR, as the reference's extract_gate does.
Status extract_gate(const ImageView& a, const ImageView& b, const Keypoints& ka, const Keypoints& kb,
                    const Matches& matches, const std::uint8_t* mask, const HomographyResult& h, const ScaleInfo& sa,
                    const ScaleInfo& sb, double* device, double* host, hipStream_t stream, const s4::Top2* forward,
                    unsigned relation_mask);

}  // namespace stitch::features
