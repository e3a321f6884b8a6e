// Strip stitching: n frames in a row, each adjacent pair overlapping, onto
// one canvas. No counterpart in the reference, which stitches pairs (its
// large-frame stress driver, harness/stress_multires.cu:81-200, runs each
// adjacent pair on its own canvas of at most 16,777,216 pixels).
//
// S1-S5 of the n - 1 adjacent pairs run as one batched phase (the batched
// form's feature engine: one launch per stage kernel over all pairs, from a
// captured graph), so each pair's result is the plain runner's bit for bit;
// then the chain canvas is laid out on the host from the n - 1 homographies
// and S6 renders every frame onto it in one pass. Two host synchronizations
// per strip: the homographies, and the finished canvas. Pair buffers are
// carved from one pool at creation; the canvas and the render workspace grow
// on demand and are kept.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "stitch/batch.hpp"
#include "stitch/runner.hpp"
#include "stitch/stages.hpp"

namespace stitch::strip {

struct StripRecord {
    Status status{Status::OK};
    const char* fail_stage{""};  // "", "s1".."s6"
    int fail_pair{-1};           // the adjacent pair of an S4 or S5 failure
    int canvas_width{0}, canvas_height{0}, origin_x{0}, origin_y{0};
    std::uint64_t union_pixels{0}, overlap_pixels{0};
    double t_pairs{0}, t_geometry{0}, t_render{0}, t_total{0};  // ms: S1-S5 phase, canvas, S6, wall
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
    void release();
    void grow_outputs(int width, int height, std::size_t workspace_bytes);

    int max_frames_, max_w_, max_h_, max_k_, max_r_;
    hipStream_t stream_{};
    hipEvent_t ev_[2]{};
    batch::DevicePool pool_;
    batch::FeatureEngine features_;
    std::vector<batch::PairBuffers> pairs_;
    s6::RenderStatistics* stats_{};
    s6::RenderStatistics* pinned_stats_{};
    // Grown on demand.
    ImageView canvas_{}, mask_{};
    void* ws_s6_{};
    std::size_t bytes_s6_{}, canvas_bytes_{}, mask_bytes_{};
    std::vector<s6::Geometry> geometry_;
};

}  // namespace stitch::strip
