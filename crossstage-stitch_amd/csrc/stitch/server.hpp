// Served form: a request loop with continuous batching.
//
// Requests arrive over time. Each one is admitted when its device memory
// fits the HBM budget and a slot is free, then walks the reference's adaptive
// flow (harness/pipeline_runner.cu:531-885): controller probe (S1 at the probe
// scale + gradient), begin_pair, attempts with S5 hypothesis reuse between
// rungs, self-checks, and S6 for the accepted homography. Every phase of
// every live request is batched with whatever else is ready at that moment:
// whenever the feature engine (S1-S5) or the render engine (S6) goes idle it
// takes all ready work, so newly arrived requests join the next launch while
// earlier ones are still on the GPU, and the two engines overlap on their own
// streams. Pair buffers (gray images, keypoints, descriptors, matches, S5
// state) and canvases stay device-resident in one pool for the life of the
// request. Kernels treat each job independently, so a request's result does
// not depend on what it was batched with.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "stitch/batch.hpp"
#include "stitch/controller.hpp"

namespace stitch::serve {

struct ServerConfig {
    std::size_t hbm_budget{64ull << 30};  // device pool for pair buffers and canvases
    int max_batch{64};                    // tasks per feature / render launch
    int max_live{256};                    // admitted requests at once
    bool adaptive{true};                  // legacy-v1 controller; false: the request's fixed knobs
    bool use_graphs{true};
    double canvas_factor{4.0};            // canvas pixels reserved at admission, per input pixel
    // Controller policy document (controller-policy-v2/v3/v4); empty: the
    // built-in legacy-v1 ladder. Learned policies get their scene features
    // after the batched probe and their gate features after each batched S5,
    // extracted per request on the server's observation stream.
    std::string policy_json{};
};

struct Request {
    std::uint64_t id{0};
    ImageView a{}, b{};            // device RGB8, resident for the request's life
    double arrival_ms{0};          // offset from the start of the trace
    std::uint64_t seed{0};         // S5 seed
    batch::PairSpec fixed{};       // knobs when the server is not adaptive
    bool keep_output{false};       // keep canvas / mask until release(id)
};

struct Completion {
    std::uint64_t id{0};
    std::string status{"OK"};
    std::string fail_stage;
    double arrival_ms{0}, admitted_ms{0}, done_ms{0};
    int attempts{0};
    controller::Knobs knobs{};
    int num_kp_a{0}, num_kp_b{0};
    HomographyResult h{};
    s6::Geometry geometry{};
    s6::RenderStatistics stats{};
    ImageView canvas{}, mask{};  // valid until release(id) when keep_output
    std::size_t reserved_bytes{0};  // budget held from admission until completion (or release)
    double latency_ms() const { return done_ms - arrival_ms; }
};

struct ServeStats {
    int completed{0}, failed{0};
    double makespan_ms{0};
    double p50_ms{0}, p99_ms{0}, mean_ms{0};
    int feature_launches{0}, render_launches{0};
    double mean_feature_batch{0}, mean_render_batch{0};
    std::size_t peak_pool_bytes{0};
    double feature_busy_ms{0}, render_busy_ms{0};
};

class Server {
public:
    Server();
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    void initialize(const ServerConfig& config);
    // Replays an arrival trace in real time (requests sorted by arrival) and
    // returns completions in completion order.
    void run_trace(const Request* requests, int n, std::vector<Completion>* out, ServeStats* stats);
    void release(std::uint64_t id);  // frees a kept output
    batch::DevicePool& pool() { return pool_; }

private:
    struct Live;
    ServerConfig config_{};
    controller::ControllerPolicy policy_{};
    bool learned_{false};
    hipStream_t obs_stream_{nullptr};
    double* obs_device_{nullptr};  // features::kDeviceDoubles
    double* obs_host_{nullptr};    // features::kGateCount, pinned
    int max_k_{0}, max_r_{0};
    batch::DevicePool pool_;
    batch::FeatureEngine features_;
    batch::RenderEngine render_;
    std::vector<std::unique_ptr<Live>> kept_;
};

}  // namespace stitch::serve
