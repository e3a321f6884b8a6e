// Plain per-pair pipeline runner: preallocated device memory, one HIP stream,
// HIP event timing, zero allocations while running. Mirrors the reference's
// harness/pipeline_runner.hpp:26-166 (initialize, query_memory, upload_pair,
// execute_run with its two host synchronizations per pair, and
// execute_adaptive_run driven by the controller with S5 state reuse) and
// harness/harness_types.hpp:14-106 (HarnessConfig, RunRecord).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>

#include "stitch/controller.hpp"
#include "stitch/features.hpp"
#include "stitch/stages.hpp"

namespace stitch::harness {

struct HarnessConfig {
    std::string config_id;
    double S{0.5};
    int K{1000};
    int R{256};
    s6::BlendMode blend_mode{s6::BlendMode::Feather};
    int fast_threshold{20};
    float harris_k{0.04f};
    double ratio_threshold{0.8};
    bool cross_check{false};
    int msac_batch_size{64};
    int msac_refinements{1};
    double msac_threshold{3.0};
    std::uint64_t seed{0};
    // Multi-level detection (an extension; 1 = the reference's single level):
    // S2/S3 run on an S1 pyramid with this many levels and scale factor.
    int pyramid_levels{1};
    double pyramid_scale_factor{1.2};
};

struct RunRecord {
    std::string config_id;
    std::uint64_t seed{0};
    double S{0.5};
    int K{1000};
    int R{256};
    int blend_mode{0};
    double t_s1{0}, t_s2{0}, t_s3{0}, t_s4{0}, t_s5{0}, t_s6{0};
    double t_controller{0}, t_e2e{0}, t_probe{0}, t_stream_wait{0};
    double probe_gradient_a{0}, probe_gradient_b{0};
    int attempt_count{0};
    int num_kp_a{0}, num_kp_b{0}, num_matches{0}, num_inliers{0};
    double inlier_ratio{0}, inlier_coverage{0}, reproj_residual{0};
    double H[9]{1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::uint64_t canvas_pixels{0}, overlap_pixels{0};
    int canvas_width{0}, canvas_height{0};
    int num_levels{0};
    std::size_t peak_workspace_bytes{0};
    int cache_hit{0};
    std::string status{"OK"};
    std::string fail_stage{};
    // Controller observations, filled when feature collection is on
    // (harness/harness_types.hpp:72-76).
    bool features_present{false};
    double t_features{0};
    features::SceneFeatures scene_features{};
    features::GateFeatures gate_features{};
};

// Bounded per-attempt trace of the adaptive run (harness/attempt_trace.hpp:10-31):
// knobs, the controller's decision, the S5 result and the stage times of each
// attempt, without allocation in the timed run.
struct AttemptRecord {
    controller::Knobs knobs{};
    controller::Decision decision{};
    HomographyResult homography{};
    int previous_R{0};
    std::uint64_t generation{0};
    double t_s1{0}, t_s2{0}, t_s3{0}, t_s4{0}, t_s5{0}, t_attempt{0};
    const char* restart_from{"none"};
    const char* fail_stage{""};
    Status execution_status{Status::OK};
    features::GateFeatures features{};  // gate observations given to the controller
    double t_features{0};
    bool has_decision{false};
    bool has_feedback{false};    // S5 ran; homography holds its result
    bool has_homography{false};  // ... and that result has a usable H
};

struct AttemptTrace {
    static constexpr int capacity = 16;
    AttemptRecord records[capacity]{};
    int count{0};
};

struct PipelineMemoryPlan {
    std::size_t device_bytes = 0;
    std::size_t workspace_bytes = 0;
    std::size_t s6_workspace_bytes = 0;
};

class PipelineRunner {
public:
    PipelineRunner();
    ~PipelineRunner();
    PipelineRunner(const PipelineRunner&) = delete;
    PipelineRunner& operator=(const PipelineRunner&) = delete;

    void initialize(int max_width, int max_height, int max_k = 4000, int max_r = 2048, int canvas_width = 4096,
                    int canvas_height = 4096, s6::BlendMode workspace_mode = s6::BlendMode::Multiband5,
                    std::size_t device_reserve = 256u * 1024u * 1024u);
    static PipelineMemoryPlan query_memory(int width, int height, int k, int r, int canvas_width, int canvas_height,
                                           s6::BlendMode mode);
    void synchronize();
    // Host RGB8 (w*h*3, unpitched); returns once the device copies are ready.
    void upload_pair(const std::uint8_t* a, int wa, int ha, const std::uint8_t* b, int wb, int hb);
    // Uses device-resident RGB8 images in place (no copy); they must outlive the runs.
    void use_device_pair(const ImageView& a, const ImageView& b);
    void set_reuse_enabled(bool enabled) { reuse_enabled_ = enabled; }
    void set_probe_enabled(bool enabled) { probe_enabled_ = enabled; }
    // Scene features at S = 0.25 once each pair is on the device, gate
    // features after every successful S5 of execute_run (the reference's
    // set_collect_features, harness/pipeline_runner.hpp:56).
    void set_collect_features(bool enabled) { collect_features_ = enabled; }
    bool collect_features() const { return collect_features_; }

    void execute_run(const HarnessConfig& config, RunRecord* record);
    void execute_adaptive_run(controller::Controller& controller, std::uint64_t seed, RunRecord* record);

    hipStream_t stream() const { return stream_; }
    const ImageView& canvas() const { return canvas_; }
    const ImageView& canvas_mask() const { return canvas_mask_; }
    const Keypoints& keypoints_a() const { return kp_a_; }
    const Keypoints& keypoints_b() const { return kp_b_; }
    const Descriptors& descriptors_a() const { return desc_a_; }
    const Descriptors& descriptors_b() const { return desc_b_; }
    const Matches& matches() const { return matches_; }
    const std::uint8_t* inlier_mask() const { return inlier_mask_; }
    const HomographyResult& homography() const { return *pinned_h_; }
    double last_gpu_span_ms() const { return last_gpu_span_ms_; }
    double last_fixed_gpu_span_ms() const { return last_gpu_span_ms_; }
    const AttemptTrace& attempt_trace() const noexcept { return attempt_trace_; }

private:
    AttemptTrace attempt_trace_{};
    void check(hipError_t e, const char* what);
    void collect_scene();
    bool reuse_enabled_{true};
    bool collect_features_{false};
    features::SceneFeatures collected_scene_{};
    double* d_features_{nullptr};
    double* pinned_features_{nullptr};
    bool probe_enabled_{true};
    hipStream_t stream_{nullptr};
    hipEvent_t ev_[8]{};
    int wa_{0}, ha_{0}, wb_{0}, hb_{0};
    ImageView img_a_{}, img_b_{};          // owned upload buffers
    ImageView use_a_{}, use_b_{};          // what the stages read (owned or external)
    ImageView gray_a_{}, gray_b_{};
    Keypoints kp_a_{}, kp_b_{};
    Descriptors desc_a_{}, desc_b_{};
    Matches matches_{};
    std::uint8_t* inlier_mask_{nullptr};
    HomographyResult* pinned_h_{nullptr};
    ImageView canvas_{}, canvas_mask_{};
    s6::RenderStatistics* stats_{nullptr};
    s6::RenderStatistics* pinned_stats_{nullptr};
    int* pinned_counts_{nullptr};
    void* ws_s2_a_{nullptr};
    void* ws_s2_b_{nullptr};
    void* ws_s3_{nullptr};
    void* ws_s4_{nullptr};
    void* ws_s5_{nullptr};
    void* ws_s6_{nullptr};
    void* ws_probe_{nullptr};
    void* ws_pyramid_{nullptr};
    std::size_t bytes_pyramid_{0};
    double* pinned_grad_{nullptr};
    std::size_t bytes_s2_{0}, bytes_s3_{0}, bytes_s4_{0}, bytes_s5_{0}, bytes_s6_{0};
    bool initialized_{false};
    int cap_w_{0}, cap_h_{0}, canvas_cap_w_{0}, canvas_cap_h_{0}, max_k_{0}, max_r_{0};
    double last_gpu_span_ms_{-1};
};

const char* blend_mode_name(s6::BlendMode m);

}  // namespace stitch::harness
