// Batched form: ragged batches of pairs run as one launch per stage kernel.
//
// Every stage already walks a device job table with per-kernel tile
// prefixes, so a batch of pairs with uneven resolutions, keypoint caps and
// hypothesis counts is one persistent launch per kernel. The launch sequence
// of each phase is captured once into a hipGraph whose kernels read their
// job counts and tile totals from device memory; a batch is one upload of
// the job tables plus one graph replay.
//
// Two engines split the pipeline where the host must step in: the feature
// engine runs S1 (and the controller probe gradient), S2, S3, S4 and S5; the
// canvas size depends on the homography, so the render engine (S6) starts
// after the host has turned each H into a canvas geometry
// (s6::prepare_geometry, host long double, like the reference's
// harness/pipeline_runner.cu:434-446). Each engine has its own stream, so a
// render batch overlaps the next feature batch.
//
// All device memory comes from one DevicePool allocated at set-up; nothing
// is allocated while batches run.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "stitch/stages.hpp"

namespace stitch::jobs {
struct LaunchProbe;
}

namespace stitch::batch {

// ------------------------------------------------------------ device pool
// One hipMalloc at set-up; first-fit ranges, 256-B aligned, handed out and
// returned by the host. Frees coalesce with their neighbours.
class DevicePool {
public:
    DevicePool() = default;
    ~DevicePool();
    DevicePool(const DevicePool&) = delete;
    DevicePool& operator=(const DevicePool&) = delete;

    void initialize(std::size_t bytes);
    void* allocate(std::size_t bytes);  // nullptr when no free range fits
    void release(void* p);
    void reset();
    std::size_t capacity() const { return size_; }
    std::size_t in_use() const { return in_use_; }
    std::size_t peak() const { return peak_; }

private:
    std::uint8_t* base_{nullptr};
    std::size_t size_{0}, in_use_{0}, peak_{0};
    std::map<std::size_t, std::size_t> free_;  // offset -> length
    std::map<std::size_t, std::size_t> used_;  // offset -> length
};

// --------------------------------------------------------- pair buffers
// Device-resident state of one pair between stages and across controller
// attempts: gray images, keypoints, descriptors, matches, inlier mask, S5
// state (kept for hypothesis reuse) and the stage workspaces.
struct PairCapacity {
    int width, height;  // max over A and B at S = 1
    int K, R;
};

struct PairBuffers {
    PairCapacity cap{};
    ImageView gray_a{}, gray_b{};
    Keypoints kp_a{}, kp_b{};
    Descriptors desc_a{}, desc_b{};
    Matches matches{};
    std::uint8_t* inlier_mask{nullptr};
    unsigned long long* grad_accum{nullptr};  // [2]
    double* grad{nullptr};                    // [2]
    void* ws_s2_a{nullptr};
    void* ws_s2_b{nullptr};
    void* ws_s3_a{nullptr};
    void* ws_s3_b{nullptr};
    void* ws_s4{nullptr};
    void* ws_s5{nullptr};
    std::size_t bytes_s2{0}, bytes_s3{0}, bytes_s4{0}, bytes_s5{0};
    void* block{nullptr};  // the pool range holding all of the above
    std::size_t block_bytes{0};
};

// Bytes one pair's buffers take in the pool (ERR_CAPACITY on bad sizes).
std::size_t pair_bytes(const PairCapacity& cap);
// Carves the buffers out of the pool; false when the pool has no room.
bool allocate_pair(DevicePool& pool, const PairCapacity& cap, PairBuffers* out);
void release_pair(DevicePool& pool, PairBuffers* b);

// Per-render device memory: canvas, mask, S6 workspace, statistics.
struct RenderBuffers {
    ImageView canvas{}, mask{};
    void* ws{nullptr};
    std::size_t ws_bytes{0};
    s6::RenderStatistics* stats{nullptr};
    void* block{nullptr};
    std::size_t block_bytes{0};
};
std::size_t render_bytes(const s6::Geometry& g, s6::BlendMode mode);
bool allocate_render(DevicePool& pool, const s6::Geometry& g, s6::BlendMode mode, RenderBuffers* out);
void release_render(DevicePool& pool, RenderBuffers* r);

// ----------------------------------------------------------- feature engine
struct FeatureTask {
    ImageView a{}, b{};  // device RGB8
    PairBuffers* buffers{nullptr};
    bool probe{false};   // controller probe: S1 at scale S and the gradient only
    double S{1.0};
    bool run_s1{true}, run_s2s3{true}, run_s4{true}, run_s5{true};
    s2::Config s2{20, 1000, 0.04f};
    s4::Config s4{0.8, false};
    s5::Config s5{256, 0, 64, 1, 3.0, 0, 1};
};

struct FeatureResult {
    Status status{Status::OK};  // host-side job errors (nothing was launched for this task)
    const char* fail_stage{""};
    ScaleInfo sa{}, sb{};
    int num_kp_a{0}, num_kp_b{0};
    double grad_a{0}, grad_b{0};
    HomographyResult h{};
};

// Device-to-device copy of one small result (gathered into one D2H).
struct CopyOp {
    const void* src;
    void* dst;
    int bytes;
};
hipError_t launch_copy_list(const CopyOp* ops, const int* count_dev, int capacity, hipStream_t stream);

class FeatureEngine {
public:
    FeatureEngine() = default;
    ~FeatureEngine();
    FeatureEngine(const FeatureEngine&) = delete;
    FeatureEngine& operator=(const FeatureEngine&) = delete;

    void initialize(int max_tasks, bool use_graph);
    // Builds the job tables, uploads them and enqueues the phase. Tasks whose
    // jobs cannot be built get a failed result and run nothing.
    void submit(const FeatureTask* tasks, int n);
    bool ready();  // non-blocking
    void wait();
    const std::vector<FeatureResult>& results() const { return results_; }
    hipStream_t stream() const { return stream_; }
    float last_ms() const;  // device time of the last phase
    void set_use_graph(bool on) { use_graph_ = on; }
    void invalidate();  // drop the captured graph (launch shapes changed)

private:
    struct Header;
    void launch(bool replay);
    void capture();
    int max_tasks_{0};
    bool use_graph_{true};
    hipStream_t stream_{nullptr};
    hipEvent_t ev_begin_{nullptr}, ev_end_{nullptr};
    hipGraph_t graph_{nullptr};
    hipGraphExec_t exec_{nullptr};
    std::uint8_t* dev_{nullptr};     // tables + header + results
    std::uint8_t* staging_{nullptr}; // pinned mirror of the tables + header
    std::uint8_t* out_{nullptr};     // pinned results
    std::size_t table_bytes_{0}, out_bytes_{0};
    std::size_t off_s1_{0}, off_grad_{0}, off_s2_{0}, off_s3_{0}, off_s4_{0}, off_s5_{0}, off_copy_{0}, off_out_{0};
    std::vector<FeatureResult> results_;
    std::vector<int> out_slot_;
    bool pending_{false};
};

// ------------------------------------------------------------ render engine
struct RenderTask {
    ImageView a{}, b{};  // device RGB8
    s6::Geometry geometry{};
    s6::BlendMode mode{s6::BlendMode::Feather};
    RenderBuffers* buffers{nullptr};
};

struct RenderResult {
    Status status{Status::OK};
    s6::RenderStatistics stats{};
};

class RenderEngine {
public:
    RenderEngine() = default;
    ~RenderEngine();
    RenderEngine(const RenderEngine&) = delete;
    RenderEngine& operator=(const RenderEngine&) = delete;

    void initialize(int max_tasks, bool use_graph);
    void submit(const RenderTask* tasks, int n);
    bool ready();
    void wait();
    const std::vector<RenderResult>& results() const { return results_; }
    hipStream_t stream() const { return stream_; }
    float last_ms() const;
    void set_use_graph(bool on) { use_graph_ = on; }
    void invalidate();
    // Per-kernel timing: while on, renders launch directly (no graph) with
    // events around every kernel; each completed render adds its launches'
    // durations and compulsory bytes. Turning it on clears the sums.
    struct ProbeStat {
        const char* name;
        int kernel, level;
        long long launches;
        double ms, bytes;
    };
    void set_probe(bool on);
    const std::vector<ProbeStat>& probe_stats() const { return probe_stats_; }

private:
    struct Header;
    void read_probe();
    bool probe_on_{false};
    jobs::LaunchProbe* probe_{nullptr};  // owned; created on first use
    std::vector<ProbeStat> probe_stats_;
    hipGraphExec_t graph_for(int top, int blend_top);
    int max_tasks_{0};
    bool use_graph_{true};
    hipStream_t stream_{nullptr};
    hipEvent_t ev_begin_{nullptr}, ev_end_{nullptr};
    std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs_;  // key: top * 64 + blend_top + 1
    std::uint8_t* dev_{nullptr};
    std::uint8_t* staging_{nullptr};
    std::uint8_t* out_{nullptr};
    std::size_t table_bytes_{0};
    std::size_t off_s6_{0}, off_copy_{0}, off_out_{0};
    std::vector<RenderResult> results_;
    std::vector<int> out_slot_;
    bool pending_{false};
};

// ------------------------------------------------------------- batch runner
struct PairSpec {
    ImageView a{}, b{};  // device RGB8, pitched
    double S{1.0};
    int K{1000};
    int R{256};
    s6::BlendMode blend{s6::BlendMode::Multiband5};
    int fast_threshold{20};
    float harris_k{0.04f};
    double ratio{0.8};
    bool cross_check{false};
    int msac_batch_size{64};
    int msac_refinements{1};
    double msac_threshold{3.0};
    std::uint64_t seed{0};
};

struct PairOutcome {
    Status status{Status::OK};
    std::string fail_stage;
    FeatureResult features{};
    s6::Geometry geometry{};
    RenderResult render{};
    const PairBuffers* buffers{nullptr};      // device views, valid until the next run
    const RenderBuffers* render_buffers{nullptr};
};

struct BatchTimings {
    double feature_ms{0}, render_ms{0}, host_ms{0}, total_ms{0};
};

struct Tuning {
    int s4_segment{128};
    int workgroups_per_cu{8};  // every stage's persistent grids, unless stage_workgroups overrides it
    bool graphs{true};
    // Per stage S1..S6: workgroups per CU in eighths of the default (0: workgroups_per_cu).
    int stage_workgroups[6]{0, 0, 0, 0, 0, 0};
    int s1_tile_rows{64};  // output rows per S1 tile: 16, 32 or 64
};

class BatchRunner {
public:
    BatchRunner() = default;
    // The pool is split into two slots (consecutive batches alternate between
    // them), each split between pair buffers (pair_share of it) and canvases
    // with their pyramids, so a batch's S1-S5 never touch memory that the
    // previous batch's S6 or its caller may still be using.
    void initialize(int max_pairs, std::size_t pool_bytes, bool use_graphs = true, double pair_share = 0.25);
    // One batch: features for every pair, one host step (geometry), render.
    void run(const PairSpec* pairs, int n, std::vector<PairOutcome>* out, BatchTimings* timings = nullptr);
    // Stream of batches: returns once this batch's S6 is enqueued; its S1-S5
    // ran on the feature stream while the previous batch's S6 was still on
    // the render stream. The previous batch completes inside the call and
    // its outcomes go to *previous (when given). A batch's device outputs
    // stay valid until the second submit after its own.
    void submit(const PairSpec* pairs, int n, std::vector<PairOutcome>* previous = nullptr,
                BatchTimings* timings = nullptr);
    // Completes the last submitted batch and returns its outcomes.
    void finish(std::vector<PairOutcome>* out, BatchTimings* timings = nullptr);
    // Pending batch count (0..1 between calls).
    bool pending() const { return slots_[cur_ ^ 1].pending; }
    // Times the tunable launch parameters on this batch and keeps the fastest
    // (results never depend on them). Reads / writes a JSON cache when a path is given.
    Tuning autotune(const PairSpec* pairs, int n, const std::string& cache_path = "");
    void apply(const Tuning& t);
    const Tuning& tuning() const { return tuning_; }
    void set_probe(bool on);
    // Both slots' render engines, summed per (kernel, level).
    const std::vector<RenderEngine::ProbeStat>& probe_stats();
    // Summed over both slots.
    void pool_use(std::size_t* capacity, std::size_t* in_use, std::size_t* peak) const;

private:
    struct Slot {
        DevicePool pair_pool, render_pool;
        std::vector<PairBuffers> pairs;
        std::vector<RenderBuffers> renders;
        std::vector<PairOutcome> out;
        std::vector<int> owner;
        bool pending{false};
    };
    void complete(int slot);
    FeatureEngine features_;
    // One render engine (stream, job tables, graphs) per slot: a batch's S6
    // is enqueued before the previous batch's S6 has finished, so the render
    // work of consecutive batches runs back to back (or overlapped) with no
    // idle gap for the host's geometry and table preparation.
    RenderEngine render_[2];
    std::vector<RenderEngine::ProbeStat> probe_merged_;
    Slot slots_[2];
    int cur_{0};
    double last_render_ms_{0};
    Tuning tuning_{};
    int max_pairs_{0};
};

// Task for a pair with fixed knobs (every stage runs).
FeatureTask fixed_task(const PairSpec& p, PairBuffers* buffers);

}  // namespace stitch::batch
