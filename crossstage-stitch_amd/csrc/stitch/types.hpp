// Plain-data stage contracts shared by S1..S6, the runner and the C ABI.
// Same names, fields and meaning as the reference's src/common/types.hpp:13-77
// and src/common/status.hpp:6-25 so stage callers port over unchanged; the
// stream type is HIP's. The caller allocates every buffer; stages never
// allocate and never synchronize.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace stitch {

enum class Status : int {
    OK = 0,
    INVALID_ARGUMENT,
    INSUFFICIENT_WORKSPACE,
    CUDA_ERROR,  // a HIP runtime or launch error; the name is kept for drop-in callers
    NOT_ENOUGH_POINTS,
    DEGENERATE,
};

inline const char* status_name(Status s) {
    switch (s) {
        case Status::OK: return "OK";
        case Status::INVALID_ARGUMENT: return "INVALID_ARGUMENT";
        case Status::INSUFFICIENT_WORKSPACE: return "INSUFFICIENT_WORKSPACE";
        case Status::CUDA_ERROR: return "CUDA_ERROR";
        case Status::NOT_ENOUGH_POINTS: return "NOT_ENOUGH_POINTS";
        case Status::DEGENERATE: return "DEGENERATE";
    }
    return "UNKNOWN";
}

enum class PixelFormat : int { RGB8 = 0, GRAY8 = 1 };

// Device image, interleaved, row pitch in bytes.
struct ImageView {
    std::uint8_t* data;
    int width;
    int height;
    int pitch_bytes;
    PixelFormat format;
};

// S1 output for one image (host). D maps original to S1 coordinates.
struct ScaleInfo {
    int orig_w, orig_h;
    int work_w, work_h;
    double sx, sy;
    double D[9];
};

// Device SoA keypoints in S1 coordinates; valid rows [0, *d_count).
struct Keypoints {
    float* x;
    float* y;
    float* response;
    int capacity;
    int* d_count;
};

enum class DescriptorType : int { BINARY_U8 = 0, FLOAT32 = 1 };

struct Descriptors {
    DescriptorType type;
    int dim;
    int row_stride_bytes;
    void* data;
};

// Device SoA matches: index into A keypoints, into B keypoints, distance.
struct Matches {
    std::int32_t* idx_a;
    std::int32_t* idx_b;
    float* distance;
    int capacity;
    int* d_count;
};

// S5 output. H maps B -> A in original coordinates; H_s in S1 coordinates.
struct HomographyResult {
    double H[9];
    double H_s[9];
    int num_matches;
    int num_inliers;
    double inlier_ratio;
    double inlier_coverage;
    double reproj_residual;
    Status status;
};

// Per-stage times in ms (HIP events on the pipeline stream).
struct StageTimings {
    double t_h2d, t_s1, t_s2, t_s3, t_s4, t_s5, t_s6;
    double t_controller, t_d2h, t_e2e;
};

}  // namespace stitch
