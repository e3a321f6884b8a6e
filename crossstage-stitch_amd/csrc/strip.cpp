// Strip stitching: S1-S5 of every adjacent pair in one batched phase, the
// chain canvas on the host, one chain render.
#include "stitch/strip.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

namespace stitch::strip {
namespace {

constexpr int kRowPadding = 64;  // bytes past each row of the canvas and mask

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

}  // namespace

void StripStitcher::check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP ") + what + ": " + hipGetErrorString(e));
}

StripStitcher::StripStitcher(int max_frames, int max_w, int max_h, int max_k, int max_r)
    : max_frames_(max_frames), max_w_(max_w), max_h_(max_h), max_k_(max_k), max_r_(max_r) {
    if (max_frames < 2 || max_frames > s6::kMaxChainFrames || max_w < 8 || max_h < 8 || max_k < 1 || max_r < 1)
        throw std::invalid_argument("strip capacities");
    try {
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
        for (hipEvent_t& e : ev_) check(hipEventCreate(&e), "event");
        // One pair's buffers per adjacent pair, at the largest frame size, K and R.
        const batch::PairCapacity cap{max_w, max_h, max_k, max_r};
        const std::size_t each = batch::pair_bytes(cap);
        const int tasks = max_frames - 1;
        pool_.initialize(each * tasks + (std::size_t{1} << 20));
        pairs_.assign(tasks, batch::PairBuffers{});
        for (batch::PairBuffers& b : pairs_)
            if (!batch::allocate_pair(pool_, cap, &b)) throw std::runtime_error("strip pair buffers");
        features_.initialize(tasks, true);
        check(hipMalloc(reinterpret_cast<void**>(&stats_), sizeof(s6::RenderStatistics)), "hipMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&pinned_stats_), sizeof(s6::RenderStatistics),
                            hipHostMallocDefault),
              "pin");
    } catch (...) {
        release();
        throw;
    }
}

StripStitcher::~StripStitcher() { release(); }

void StripStitcher::release() {
    if (stream_) hipStreamSynchronize(stream_);
    for (batch::PairBuffers& b : pairs_) batch::release_pair(pool_, &b);
    pairs_.clear();
    if (stats_) hipFree(stats_);
    if (canvas_.data) hipFree(canvas_.data);
    if (mask_.data) hipFree(mask_.data);
    if (ws_s6_) hipFree(ws_s6_);
    stats_ = nullptr;
    canvas_.data = mask_.data = nullptr;
    ws_s6_ = nullptr;
    if (pinned_stats_) hipHostFree(pinned_stats_);
    pinned_stats_ = nullptr;
    for (hipEvent_t& e : ev_)
        if (e) {
            hipEventDestroy(e);
            e = nullptr;
        }
    if (stream_) hipStreamDestroy(stream_);
    stream_ = nullptr;
}

void StripStitcher::grow_outputs(int width, int height, std::size_t workspace_bytes) {
    const std::size_t rgb = static_cast<std::size_t>(width * 3 + kRowPadding) * height;
    const std::size_t gray = static_cast<std::size_t>(width + kRowPadding) * height;
    // The stream may still read the old buffers.
    if (rgb > canvas_bytes_ || gray > mask_bytes_ || workspace_bytes > bytes_s6_) check(hipStreamSynchronize(stream_), "sync");
    if (rgb > canvas_bytes_) {
        if (canvas_.data) check(hipFree(canvas_.data), "hipFree");
        canvas_.data = nullptr;
        canvas_bytes_ = 0;
        check(hipMalloc(reinterpret_cast<void**>(&canvas_.data), rgb), "hipMalloc");
        canvas_bytes_ = rgb;
    }
    if (gray > mask_bytes_) {
        if (mask_.data) check(hipFree(mask_.data), "hipFree");
        mask_.data = nullptr;
        mask_bytes_ = 0;
        check(hipMalloc(reinterpret_cast<void**>(&mask_.data), gray), "hipMalloc");
        mask_bytes_ = gray;
    }
    if (workspace_bytes > bytes_s6_) {
        if (ws_s6_) check(hipFree(ws_s6_), "hipFree");
        ws_s6_ = nullptr;
        bytes_s6_ = 0;
        check(hipMalloc(&ws_s6_, workspace_bytes), "hipMalloc");
        bytes_s6_ = workspace_bytes;
    }
    canvas_.width = mask_.width = width;
    canvas_.height = mask_.height = height;
    canvas_.pitch_bytes = width * 3 + kRowPadding;
    mask_.pitch_bytes = width + kRowPadding;
    canvas_.format = PixelFormat::RGB8;
    mask_.format = PixelFormat::GRAY8;
}

void StripStitcher::run(int n, const ImageView* frames, const harness::HarnessConfig& cfg, StripRecord* rec,
                        HomographyResult* pairs, int* keypoints) {
    *rec = StripRecord{};
    auto fail = [&](Status s, const char* stage, int pair = -1) {
        rec->status = s;
        rec->fail_stage = stage;
        rec->fail_pair = pair;
    };
    if (n < 2 || n > max_frames_ || !frames || !pairs || cfg.K < 1 || cfg.K > max_k_ || cfg.R < 1 || cfg.R > max_r_ ||
        cfg.pyramid_levels > 1)
        return fail(Status::INVALID_ARGUMENT, "args");
    for (int k = 0; k < n; ++k)
        if (frames[k].width > max_w_ || frames[k].height > max_h_ || frames[k].format != PixelFormat::RGB8)
            return fail(Status::INVALID_ARGUMENT, "args");
    const auto t0 = Clock::now();
    // S1-S5 of pair k = (frame k, frame k + 1), every pair in one phase.
    std::vector<batch::FeatureTask> tasks(n - 1);
    for (int k = 0; k + 1 < n; ++k) {
        batch::FeatureTask& t = tasks[k];
        t.a = frames[k];
        t.b = frames[k + 1];
        t.buffers = &pairs_[k];
        t.S = cfg.S;
        t.s2 = {cfg.fast_threshold, cfg.K, cfg.harris_k};
        t.s4 = {cfg.ratio_threshold, cfg.cross_check};
        t.s5 = {cfg.R, 0, cfg.msac_batch_size, cfg.msac_refinements, cfg.msac_threshold, cfg.seed, 1};
    }
    features_.submit(tasks.data(), n - 1);
    // Host synchronization 1: the homographies lay out the canvas.
    features_.wait();
    rec->t_pairs = features_.last_ms();
    const std::vector<batch::FeatureResult>& res = features_.results();
    for (int k = 0; k + 1 < n; ++k) pairs[k] = res[k].h;
    if (keypoints) {
        for (int k = 0; k + 1 < n; ++k) keypoints[k] = res[k].num_kp_a;
        keypoints[n - 1] = res[n - 2].num_kp_b;
    }
    for (int k = 0; k + 1 < n; ++k) {
        if (res[k].status != Status::OK) {
            rec->t_total = ms_since(t0);
            return fail(res[k].status, res[k].fail_stage, k);
        }
        if (res[k].h.status != Status::OK) {
            rec->t_total = ms_since(t0);
            return fail(res[k].h.status, "s5", k);
        }
    }
    const auto tg = Clock::now();
    std::vector<int> widths(n), heights(n);
    std::vector<double> h_adj(9 * (n - 1));
    for (int k = 0; k < n; ++k) {
        widths[k] = frames[k].width;
        heights[k] = frames[k].height;
        if (k + 1 < n) std::copy_n(res[k].h.H, 9, h_adj.begin() + 9 * k);
    }
    s6::GeometryPolicy policy{};
    policy.max_canvas_pixels = s6::kChainCanvasPixels;
    geometry_.assign(n, s6::Geometry{});
    Status s = s6::prepare_chain_geometry(n, widths.data(), heights.data(), h_adj.data(), policy, geometry_.data());
    std::size_t need = 0;
    if (s == Status::OK) s = s6::chain_workspace_bytes(n, geometry_.data(), cfg.blend_mode, &need);
    rec->t_geometry = ms_since(tg);
    if (s != Status::OK) {
        rec->t_total = ms_since(t0);
        return fail(s, "s6");
    }
    grow_outputs(geometry_[0].width, geometry_[0].height, need);
    check(hipEventRecord(ev_[0], stream_), "event");
    s = s6::render_chain(n, frames, geometry_.data(), cfg.blend_mode, canvas_, mask_, ws_s6_, bytes_s6_, stream_,
                         stats_);
    if (s != Status::OK) {
        rec->t_total = ms_since(t0);
        return fail(s, "s6");
    }
    check(hipEventRecord(ev_[1], stream_), "event");
    check(hipMemcpyAsync(pinned_stats_, stats_, sizeof(*stats_), hipMemcpyDeviceToHost, stream_), "d2h");
    // Host synchronization 2: the canvas is ready.
    check(hipStreamSynchronize(stream_), "sync");
    float ms = 0;
    check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "elapsed");
    rec->t_render = ms;
    rec->t_total = ms_since(t0);
    rec->canvas_width = geometry_[0].width;
    rec->canvas_height = geometry_[0].height;
    rec->origin_x = geometry_[0].origin_x;
    rec->origin_y = geometry_[0].origin_y;
    rec->union_pixels = pinned_stats_->union_pixels;
    rec->overlap_pixels = pinned_stats_->overlap_pixels;
}

}  // namespace stitch::strip
