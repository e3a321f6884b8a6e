// Strip stitching over the stage APIs: features once per frame, matching and
// estimation per adjacent pair, one chain render.
#include "stitch/strip.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

namespace stitch::strip {
namespace {

constexpr int kRowPadding = 64;  // bytes past each row of the images kept here

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

}  // namespace

void StripStitcher::check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP ") + what + ": " + hipGetErrorString(e));
}

void* StripStitcher::alloc(std::size_t bytes) {
    void* p = nullptr;
    check(hipMalloc(&p, std::max<std::size_t>(bytes, 1)), "hipMalloc");
    owned_.push_back(p);
    return p;
}

StripStitcher::StripStitcher(int max_frames, int max_w, int max_h, int max_k, int max_r)
    : max_frames_(max_frames), max_w_(max_w), max_h_(max_h), max_k_(max_k), max_r_(max_r) {
    if (max_frames < 2 || max_frames > s6::kMaxChainFrames || max_w < 8 || max_h < 8 || max_k < 1 || max_r < 1)
        throw std::invalid_argument("strip capacities");
    try {
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
        for (hipEvent_t& e : ev_) check(hipEventCreate(&e), "event");
        for (int k = 0; k < max_frames; ++k) {
            ImageView g{};
            g.pitch_bytes = max_w + kRowPadding;
            g.format = PixelFormat::GRAY8;
            g.data = static_cast<std::uint8_t*>(alloc(static_cast<std::size_t>(g.pitch_bytes) * max_h));
            gray_.push_back(g);
            Keypoints kp{};
            kp.capacity = max_k;
            kp.x = static_cast<float*>(alloc(max_k * sizeof(float)));
            kp.y = static_cast<float*>(alloc(max_k * sizeof(float)));
            kp.response = static_cast<float*>(alloc(max_k * sizeof(float)));
            kp.d_count = static_cast<int*>(alloc(sizeof(int)));
            keypoints_.push_back(kp);
            Descriptors d{};
            d.type = DescriptorType::BINARY_U8;
            d.dim = 32;
            d.row_stride_bytes = 32;
            d.data = alloc(static_cast<std::size_t>(max_k) * 32);
            descriptors_.push_back(d);
            if (k + 1 < max_frames) {
                Matches m{};
                m.capacity = max_k;
                m.idx_a = static_cast<std::int32_t*>(alloc(max_k * sizeof(std::int32_t)));
                m.idx_b = static_cast<std::int32_t*>(alloc(max_k * sizeof(std::int32_t)));
                m.distance = static_cast<float*>(alloc(max_k * sizeof(float)));
                m.d_count = static_cast<int*>(alloc(sizeof(int)));
                matches_.push_back(m);
                inliers_.push_back(static_cast<std::uint8_t*>(alloc(max_k)));
            }
        }
        stats_ = static_cast<s6::RenderStatistics*>(alloc(sizeof(s6::RenderStatistics)));
        if (s2::s2_workspace_bytes(max_w, max_h, {20, max_k, .04f}, &bytes_s2_) != Status::OK ||
            s3::s3_workspace_bytes(max_w, max_h, &bytes_s3_) != Status::OK ||
            s4::s4_workspace_bytes(max_k, max_k, {0.8, true}, &bytes_s4_) != Status::OK ||
            s5::s5_workspace_bytes({max_r, 0, 64, 1, 3.0, 0, 1}, &bytes_s5_) != Status::OK)
            throw std::invalid_argument("strip capacities");
        ws_s2_ = alloc(bytes_s2_);
        ws_s3_ = alloc(bytes_s3_);
        ws_s4_ = alloc(bytes_s4_);
        ws_s5_ = alloc(bytes_s5_);
        check(hipHostMalloc(reinterpret_cast<void**>(&pinned_h_), sizeof(HomographyResult) * (max_frames - 1),
                            hipHostMallocDefault),
              "pin");
        check(hipHostMalloc(reinterpret_cast<void**>(&pinned_counts_), sizeof(int) * max_frames, hipHostMallocDefault),
              "pin");
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
    for (void* p : owned_) hipFree(p);
    owned_.clear();
    if (canvas_.data) hipFree(canvas_.data);
    if (mask_.data) hipFree(mask_.data);
    if (ws_s6_) hipFree(ws_s6_);
    canvas_.data = mask_.data = nullptr;
    ws_s6_ = nullptr;
    if (pinned_h_) hipHostFree(pinned_h_);
    if (pinned_counts_) hipHostFree(pinned_counts_);
    if (pinned_stats_) hipHostFree(pinned_stats_);
    pinned_h_ = nullptr;
    pinned_counts_ = nullptr;
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
    std::vector<ScaleInfo> scale(n);
    check(hipEventRecord(ev_[0], stream_), "event");
    const s2::Config c2{cfg.fast_threshold, cfg.K, cfg.harris_k};
    for (int k = 0; k < n; ++k) {
        check(hipMemsetAsync(keypoints_[k].d_count, 0, sizeof(int), stream_), "memset");
        if (make_scale_info(frames[k].width, frames[k].height, cfg.S, &scale[k]) != Status::OK)
            return fail(Status::INVALID_ARGUMENT, "s1");
        gray_[k].width = scale[k].work_w;
        gray_[k].height = scale[k].work_h;
        Status s = s1::s1_preprocess_single(frames[k], cfg.S, &gray_[k], &scale[k], stream_);
        if (s != Status::OK) return fail(s, "s1");
        Keypoints kp = keypoints_[k];
        kp.capacity = cfg.K;
        if ((s = s2::s2_detect(gray_[k], c2, kp, ws_s2_, bytes_s2_, stream_)) != Status::OK) return fail(s, "s2");
        if ((s = s3::s3_describe(gray_[k], kp, descriptors_[k], ws_s3_, bytes_s3_, stream_)) != Status::OK)
            return fail(s, "s3");
    }
    check(hipEventRecord(ev_[1], stream_), "event");
    for (int k = 0; k + 1 < n; ++k) {
        Keypoints a = keypoints_[k], b = keypoints_[k + 1];
        a.capacity = b.capacity = cfg.K;
        check(hipMemsetAsync(matches_[k].d_count, 0, sizeof(int), stream_), "memset");
        const Status s = s4::s4_match(a, descriptors_[k], b, descriptors_[k + 1], {cfg.ratio_threshold, cfg.cross_check},
                                      matches_[k], ws_s4_, bytes_s4_, stream_);
        if (s != Status::OK) return fail(s, "s4", k);
    }
    check(hipEventRecord(ev_[2], stream_), "event");
    for (int k = 0; k + 1 < n; ++k) {
        Keypoints a = keypoints_[k], b = keypoints_[k + 1];
        a.capacity = b.capacity = cfg.K;
        const Status s = s5::estimate(a, b, matches_[k], scale[k], scale[k + 1],
                                      {cfg.R, 0, cfg.msac_batch_size, cfg.msac_refinements, cfg.msac_threshold,
                                       cfg.seed, 1},
                                      inliers_[k], pinned_h_ + k, ws_s5_, bytes_s5_, stream_);
        if (s != Status::OK) return fail(s, "s5", k);
    }
    for (int k = 0; k < n; ++k)
        check(hipMemcpyAsync(pinned_counts_ + k, keypoints_[k].d_count, sizeof(int), hipMemcpyDeviceToHost, stream_),
              "d2h");
    check(hipEventRecord(ev_[3], stream_), "event");
    // Host synchronization 1: the homographies lay out the canvas.
    check(hipStreamSynchronize(stream_), "sync");
    float ms = 0;
    check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "elapsed");
    rec->t_features = ms;
    check(hipEventElapsedTime(&ms, ev_[1], ev_[2]), "elapsed");
    rec->t_match = ms;
    check(hipEventElapsedTime(&ms, ev_[2], ev_[3]), "elapsed");
    rec->t_estimate = ms;
    std::copy_n(pinned_h_, n - 1, pairs);
    if (keypoints) std::copy_n(pinned_counts_, n, keypoints);
    for (int k = 0; k + 1 < n; ++k)
        if (pinned_h_[k].status != Status::OK) {
            rec->t_total = ms_since(t0);
            return fail(pinned_h_[k].status, "s5", k);
        }
    const auto tg = Clock::now();
    std::vector<int> widths(n), heights(n);
    std::vector<double> h_adj(9 * (n - 1));
    for (int k = 0; k < n; ++k) {
        widths[k] = frames[k].width;
        heights[k] = frames[k].height;
        if (k + 1 < n) std::copy_n(pinned_h_[k].H, 9, h_adj.begin() + 9 * k);
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
    check(hipEventRecord(ev_[4], stream_), "event");
    s = s6::render_chain(n, frames, geometry_.data(), cfg.blend_mode, canvas_, mask_, ws_s6_, bytes_s6_, stream_,
                         stats_);
    if (s != Status::OK) {
        rec->t_total = ms_since(t0);
        return fail(s, "s6");
    }
    check(hipEventRecord(ev_[5], stream_), "event");
    check(hipMemcpyAsync(pinned_stats_, stats_, sizeof(*stats_), hipMemcpyDeviceToHost, stream_), "d2h");
    // Host synchronization 2: the canvas is ready.
    check(hipStreamSynchronize(stream_), "sync");
    check(hipEventElapsedTime(&ms, ev_[4], ev_[5]), "elapsed");
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
