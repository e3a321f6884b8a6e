// Strip stitching: n frames in a row, each adjacent pair overlapping, onto
// one canvas. No counterpart in the reference, which stitches pairs (its
// large-frame stress driver, harness/stress_multires.cu:81-200, runs each
// adjacent pair on its own canvas of at most 16,777,216 pixels).
//
// S1-S3 run once per frame (a frame is B of one pair and A of the next),
// S4 and S5 once per adjacent pair, then the chain canvas is laid out on the
// host from the n - 1 homographies and S6 renders every frame onto it in
// one pass. Two host synchronizations per strip: the homographies, and the
// finished canvas. Device memory for frames, features and matches is
// allocated at creation; the canvas and the render workspace grow on demand
// and are kept.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "stitch/runner.hpp"
#include "stitch/stages.hpp"

namespace stitch::strip {

struct StripRecord {
    Status status{Status::OK};
    const char* fail_stage{""};  // "", "s1".."s6"
    int fail_pair{-1};           // the adjacent pair of an S4 or S5 failure
    int canvas_width{0}, canvas_height{0}, origin_x{0}, origin_y{0};
    std::uint64_t union_pixels{0}, overlap_pixels{0};
    double t_features{0}, t_match{0}, t_estimate{0}, t_geometry{0}, t_render{0}, t_total{0};  // ms
};

class StripStitcher {
public:
    StripStitcher(int max_frames, int max_w, int max_h, int max_k, int max_r);
    ~StripStitcher();
    StripStitcher(const StripStitcher&) = delete;
    StripStitcher& operator=(const StripStitcher&) = delete;

    // frames: n device RGB8 views (2 <= n <= max_frames, each within
    // max_w x max_h). pairs (n - 1 entries, host) receive each adjacent
    // pair's S5 result, pairs[k] mapping frame k + 1 to frame k; keypoints
    // (n entries, host, may be null) each frame's keypoint count.
    void run(int n, const ImageView* frames, const harness::HarnessConfig& config, StripRecord* record,
             HomographyResult* pairs, int* keypoints);

    const ImageView& canvas() const { return canvas_; }
    const ImageView& mask() const { return mask_; }
    const std::vector<s6::Geometry>& geometry() const { return geometry_; }
    hipStream_t stream() const { return stream_; }

private:
    void check(hipError_t e, const char* what);
    void* alloc(std::size_t bytes);
    void release();
    void grow_outputs(int width, int height, std::size_t workspace_bytes);

    int max_frames_, max_w_, max_h_, max_k_, max_r_;
    hipStream_t stream_{};
    hipEvent_t ev_[6]{};
    std::vector<void*> owned_;
    std::vector<ImageView> gray_;
    std::vector<Keypoints> keypoints_;
    std::vector<Descriptors> descriptors_;
    std::vector<Matches> matches_;
    std::vector<std::uint8_t*> inliers_;
    HomographyResult* pinned_h_{};
    int* pinned_counts_{};
    s6::RenderStatistics* stats_{};
    s6::RenderStatistics* pinned_stats_{};
    void *ws_s2_{}, *ws_s3_{}, *ws_s4_{}, *ws_s5_{};
    std::size_t bytes_s2_{}, bytes_s3_{}, bytes_s4_{}, bytes_s5_{};
    // Grown on demand.
    ImageView canvas_{}, mask_{};
    void* ws_s6_{};
    std::size_t bytes_s6_{}, canvas_bytes_{}, mask_bytes_{};
    std::vector<s6::Geometry> geometry_;
};

}  // namespace stitch::strip
