// Plain per-pair pipeline runner (see stitch/runner.hpp).
// Flow of the reference's harness/pipeline_runner.cu:110-885.
#include "stitch/runner.hpp"

#include "stitch/pyramid.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>

namespace stitch::harness {
namespace {

constexpr int kPadding = 64;

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

struct DeviceError : std::runtime_error {
    hipError_t code;
    DeviceError(hipError_t e, const char* what) : std::runtime_error(what), code(e) {}
};

}  // namespace

const char* blend_mode_name(s6::BlendMode m) {
    switch (m) {
        case s6::BlendMode::Feather: return "feather";
        case s6::BlendMode::Multiband3: return "multiband_3";
        case s6::BlendMode::Multiband5: return "multiband_5";
        case s6::BlendMode::Multiband6: return "multiband_6";
        case s6::BlendMode::Multiband7: return "multiband_7";
    }
    return "unknown";
}

void PipelineRunner::check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(e, what);
}

PipelineRunner::PipelineRunner() {
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
    for (auto& e : ev_) check(hipEventCreate(&e), "event");
}

PipelineRunner::~PipelineRunner() {
    if (stream_) hipStreamSynchronize(stream_);
    for (auto& e : ev_)
        if (e) hipEventDestroy(e);
    for (void* p : {static_cast<void*>(img_a_.data), static_cast<void*>(img_b_.data), static_cast<void*>(gray_a_.data),
                    static_cast<void*>(gray_b_.data), static_cast<void*>(kp_a_.x), static_cast<void*>(kp_a_.y),
                    static_cast<void*>(kp_a_.response), static_cast<void*>(kp_a_.d_count),
                    static_cast<void*>(kp_b_.x), static_cast<void*>(kp_b_.y), static_cast<void*>(kp_b_.response),
                    static_cast<void*>(kp_b_.d_count), desc_a_.data, desc_b_.data,
                    static_cast<void*>(matches_.idx_a), static_cast<void*>(matches_.idx_b),
                    static_cast<void*>(matches_.distance), static_cast<void*>(matches_.d_count),
                    static_cast<void*>(inlier_mask_), static_cast<void*>(canvas_.data),
                    static_cast<void*>(canvas_mask_.data), static_cast<void*>(stats_), ws_s2_a_, ws_s2_b_, ws_s3_,
                    ws_s4_, ws_s5_, ws_s6_, ws_probe_, static_cast<void*>(d_features_), ws_pyramid_})
        if (p) hipFree(p);
    if (pinned_features_) hipHostFree(pinned_features_);
    if (pinned_h_) hipHostFree(pinned_h_);
    if (pinned_stats_) hipHostFree(pinned_stats_);
    if (pinned_counts_) hipHostFree(pinned_counts_);
    if (pinned_grad_) hipHostFree(pinned_grad_);
    if (stream_) hipStreamDestroy(stream_);
}

PipelineMemoryPlan PipelineRunner::query_memory(int w, int h, int k, int r, int cw, int ch, s6::BlendMode mode) {
    using namespace admission;
    dimensions_fit(w, h, w, h, cw, ch);
    if (k <= 0 || r <= 0 || multiply(cw, ch) > s6::GeometryPolicy{}.max_canvas_pixels) throw Error("ERR_CAPACITY");
    PipelineMemoryPlan plan{};
    plan.device_bytes = add(multiply(2, image_bytes(w, h, 3)), multiply(2, image_bytes(w, h, 1)));
    plan.device_bytes = add(plan.device_bytes, add(image_bytes(cw, ch, 3), image_bytes(cw, ch, 1)));
    plan.device_bytes = add(plan.device_bytes, add(multiply(k, 6 * sizeof(float) + 64 + 3 * sizeof(int) + 1),
                                                   3 * sizeof(int) + sizeof(s6::RenderStatistics) + 32));
    plan.device_bytes = add(plan.device_bytes, features::kDeviceDoubles * sizeof(double));
    {
        std::size_t pyramid_bytes = 0;
        if (pyramid::workspace_bytes(w, h, 1.0, {pyramid::kMaxLevels, 1.2, 20, k, .04f}, &pyramid_bytes) != Status::OK)
            throw Error("ERR_CAPACITY");
        plan.device_bytes = add(plan.device_bytes, pyramid_bytes);
    }
    auto workspace = [&](std::size_t bytes, int count = 1) {
        plan.workspace_bytes = add(plan.workspace_bytes, multiply(bytes, count));
        if (bytes) plan.device_bytes = add(plan.device_bytes, multiply(add(bytes, 256), count));
    };
    auto require = [](Status s) {
        if (s != Status::OK) throw Error("ERR_CAPACITY");
    };
    std::size_t bytes = 0;
    require(s2::s2_workspace_bytes(w, h, {20, k, .04f}, &bytes));
    workspace(bytes, 2);
    require(s3::s3_workspace_bytes(w, h, &bytes));
    workspace(bytes);
    require(s4::s4_workspace_bytes(k, k, {.8, true}, &bytes));
    workspace(bytes);
    require(s5::s5_workspace_bytes({r, 0, 64, 1, 3., 0, 1}, &bytes));
    workspace(bytes);
    s6::Geometry g{};
    g.width = cw;
    g.height = ch;
    g.width_a = g.width_b = w;
    g.height_a = g.height_b = h;
    g.inverse[0] = g.inverse[4] = g.inverse[8] = 1;
    require(s6::s6_workspace_bytes(g, mode, &plan.s6_workspace_bytes));
    workspace(plan.s6_workspace_bytes);
    return plan;
}

void PipelineRunner::initialize(int max_w, int max_h, int max_k, int max_r, int cw, int ch, s6::BlendMode mode,
                                std::size_t reserve) {
    if (initialized_) return;
    const PipelineMemoryPlan plan = query_memory(max_w, max_h, max_k, max_r, cw, ch, mode);
    std::size_t free_bytes = 0, total_bytes = 0;
    check(hipMemGetInfo(&free_bytes, &total_bytes), "meminfo");
    admission::fit(plan.device_bytes, free_bytes, reserve, "ERR_DEVICE_OOM");
    cap_w_ = max_w;
    cap_h_ = max_h;
    canvas_cap_w_ = cw;
    canvas_cap_h_ = ch;
    max_k_ = max_k;
    max_r_ = max_r;
    auto alloc = [&](std::size_t n) {
        void* p = nullptr;
        check(hipMalloc(&p, std::max<std::size_t>(n, 1)), "hipMalloc");
        return p;
    };
    auto image = [&](int w, int h, int channels, PixelFormat f) {
        ImageView v{};
        v.width = w;
        v.height = h;
        v.pitch_bytes = w * channels + kPadding;
        v.format = f;
        v.data = static_cast<std::uint8_t*>(alloc(static_cast<std::size_t>(v.pitch_bytes) * h));
        return v;
    };
    img_a_ = image(max_w, max_h, 3, PixelFormat::RGB8);
    img_b_ = image(max_w, max_h, 3, PixelFormat::RGB8);
    gray_a_ = image(max_w, max_h, 1, PixelFormat::GRAY8);
    gray_b_ = image(max_w, max_h, 1, PixelFormat::GRAY8);
    for (Keypoints* k : {&kp_a_, &kp_b_}) {
        k->capacity = max_k;
        k->x = static_cast<float*>(alloc(max_k * sizeof(float)));
        k->y = static_cast<float*>(alloc(max_k * sizeof(float)));
        k->response = static_cast<float*>(alloc(max_k * sizeof(float)));
        k->d_count = static_cast<int*>(alloc(sizeof(int)));
    }
    for (Descriptors* d : {&desc_a_, &desc_b_}) {
        d->type = DescriptorType::BINARY_U8;
        d->dim = 32;
        d->row_stride_bytes = 32;
        d->data = alloc(static_cast<std::size_t>(max_k) * 32);
    }
    matches_.capacity = max_k;
    matches_.idx_a = static_cast<std::int32_t*>(alloc(max_k * sizeof(std::int32_t)));
    matches_.idx_b = static_cast<std::int32_t*>(alloc(max_k * sizeof(std::int32_t)));
    matches_.distance = static_cast<float*>(alloc(max_k * sizeof(float)));
    matches_.d_count = static_cast<int*>(alloc(sizeof(int)));
    inlier_mask_ = static_cast<std::uint8_t*>(alloc(max_k));
    check(hipHostMalloc(reinterpret_cast<void**>(&pinned_h_), sizeof(HomographyResult), hipHostMallocDefault), "pin");
    check(hipHostMalloc(reinterpret_cast<void**>(&pinned_stats_), sizeof(s6::RenderStatistics), hipHostMallocDefault),
          "pin");
    check(hipHostMalloc(reinterpret_cast<void**>(&pinned_counts_), 4 * sizeof(int), hipHostMallocDefault), "pin");
    check(hipHostMalloc(reinterpret_cast<void**>(&pinned_grad_), 2 * sizeof(double), hipHostMallocDefault), "pin");
    canvas_ = image(cw, ch, 3, PixelFormat::RGB8);
    canvas_mask_ = image(cw, ch, 1, PixelFormat::GRAY8);
    stats_ = static_cast<s6::RenderStatistics*>(alloc(sizeof(s6::RenderStatistics)));
    if (s2::s2_workspace_bytes(max_w, max_h, {20, max_k, .04f}, &bytes_s2_) != Status::OK ||
        s3::s3_workspace_bytes(max_w, max_h, &bytes_s3_) != Status::OK ||
        s4::s4_workspace_bytes(max_k, max_k, {0.8, true}, &bytes_s4_) != Status::OK ||
        s5::s5_workspace_bytes({max_r, 0, 64, 1, 3.0, 0, 1}, &bytes_s5_) != Status::OK)
        throw admission::Error("ERR_CAPACITY");
    ws_s2_a_ = alloc(bytes_s2_);
    ws_s2_b_ = alloc(bytes_s2_);
    ws_s3_ = alloc(bytes_s3_);
    ws_s4_ = alloc(bytes_s4_);
    ws_s5_ = alloc(bytes_s5_);
    bytes_s6_ = plan.s6_workspace_bytes;
    if (bytes_s6_) ws_s6_ = alloc(bytes_s6_);
    ws_probe_ = alloc(32);
    if (pyramid::workspace_bytes(max_w, max_h, 1.0, {pyramid::kMaxLevels, 1.2, 20, max_k, .04f}, &bytes_pyramid_) !=
        Status::OK)
        throw admission::Error("ERR_CAPACITY");
    ws_pyramid_ = alloc(bytes_pyramid_);
    d_features_ = static_cast<double*>(alloc(features::kDeviceDoubles * sizeof(double)));
    check(hipHostMalloc(reinterpret_cast<void**>(&pinned_features_), features::kGateCount * sizeof(double),
                        hipHostMallocDefault),
          "pin");
    use_a_ = img_a_;
    use_b_ = img_b_;
    initialized_ = true;
}

void PipelineRunner::synchronize() { check(hipStreamSynchronize(stream_), "sync"); }

void PipelineRunner::upload_pair(const std::uint8_t* a, int wa, int ha, const std::uint8_t* b, int wb, int hb) {
    if (!initialized_) throw admission::Error("ERR_CAPACITY");
    admission::dimensions_fit(wa, ha, wb, hb, cap_w_, cap_h_);
    wa_ = wa;
    ha_ = ha;
    wb_ = wb;
    hb_ = hb;
    use_a_ = img_a_;
    use_b_ = img_b_;
    use_a_.width = wa;
    use_a_.height = ha;
    use_b_.width = wb;
    use_b_.height = hb;
    check(hipMemcpy2DAsync(use_a_.data, use_a_.pitch_bytes, a, wa * 3, wa * 3, ha, hipMemcpyHostToDevice, stream_), "h2d");
    check(hipMemcpy2DAsync(use_b_.data, use_b_.pitch_bytes, b, wb * 3, wb * 3, hb, hipMemcpyHostToDevice, stream_), "h2d");
    check(hipStreamSynchronize(stream_), "sync");
    if (collect_features_) collect_scene();
}

// Scene observations of the pair on the device: S1 at S = 0.25 into the gray
// buffers, then the scene features (harness/pipeline_runner.cu:295-305).
void PipelineRunner::collect_scene() {
    ScaleInfo a{}, b{};
    make_scale_info(wa_, ha_, 0.25, &a);
    make_scale_info(wb_, hb_, 0.25, &b);
    gray_a_.width = a.work_w;
    gray_a_.height = a.work_h;
    gray_b_.width = b.work_w;
    gray_b_.height = b.work_h;
    if (s1::s1_preprocess(use_a_, use_b_, 0.25, &gray_a_, &gray_b_, &a, &b, stream_) != Status::OK ||
        features::extract_scene(gray_a_, gray_b_, d_features_, pinned_features_, stream_, true) != Status::OK)
        throw std::runtime_error("Scene feature collection failed");
    check(hipStreamSynchronize(stream_), "sync");
    std::copy_n(pinned_features_, features::kSceneCount, collected_scene_.begin());
}

void PipelineRunner::use_device_pair(const ImageView& a, const ImageView& b) {
    if (!initialized_) throw admission::Error("ERR_CAPACITY");
    admission::dimensions_fit(a.width, a.height, b.width, b.height, cap_w_, cap_h_);
    use_a_ = a;
    use_b_ = b;
    wa_ = a.width;
    ha_ = a.height;
    wb_ = b.width;
    hb_ = b.height;
    if (collect_features_) collect_scene();
}

void PipelineRunner::execute_run(const HarnessConfig& cfg, RunRecord* rec) {
    last_gpu_span_ms_ = -1;
    rec->config_id = cfg.config_id;
    rec->seed = cfg.seed;
    rec->S = cfg.S;
    rec->K = cfg.K;
    rec->R = cfg.R;
    rec->blend_mode = static_cast<int>(cfg.blend_mode);
    rec->attempt_count = 1;
    if (collect_features_) rec->scene_features = collected_scene_;
    auto fail = [&](const char* status, const char* stage) {
        rec->status = status;
        rec->fail_stage = stage;
    };
    check(hipMemsetAsync(kp_a_.d_count, 0, sizeof(int), stream_), "memset");
    check(hipMemsetAsync(kp_b_.d_count, 0, sizeof(int), stream_), "memset");
    check(hipMemsetAsync(matches_.d_count, 0, sizeof(int), stream_), "memset");
    ScaleInfo sa{}, sb{};
    make_scale_info(wa_, ha_, cfg.S, &sa);
    make_scale_info(wb_, hb_, cfg.S, &sb);
    gray_a_.width = sa.work_w;
    gray_a_.height = sa.work_h;
    gray_b_.width = sb.work_w;
    gray_b_.height = sb.work_h;
    const auto t0 = Clock::now();
    check(hipEventRecord(ev_[0], stream_), "event");
    Status s = s1::s1_preprocess(use_a_, use_b_, cfg.S, &gray_a_, &gray_b_, &sa, &sb, stream_);
    if (s != Status::OK) return fail(status_name(s), "s1");
    check(hipEventRecord(ev_[1], stream_), "event");
    Keypoints ka = kp_a_, kb = kp_b_;
    if (cfg.K >= 0 && cfg.K <= kp_a_.capacity) ka.capacity = kb.capacity = cfg.K;
    if (cfg.pyramid_levels > 1) {
        // Pyramid detection and description (S2 time covers both; S3 is 0).
        const pyramid::Config pc{cfg.pyramid_levels, cfg.pyramid_scale_factor, cfg.fast_threshold, cfg.K, cfg.harris_k};
        Status sp = pyramid::detect_describe(use_a_, cfg.S, pc, ka, desc_a_, nullptr, nullptr, ws_pyramid_,
                                             bytes_pyramid_, stream_);
        if (sp == Status::OK)
            sp = pyramid::detect_describe(use_b_, cfg.S, pc, kb, desc_b_, nullptr, nullptr, ws_pyramid_,
                                          bytes_pyramid_, stream_);
        if (sp != Status::OK) return fail(status_name(sp), "s2");
        check(hipEventRecord(ev_[2], stream_), "event");
        check(hipEventRecord(ev_[3], stream_), "event");
    } else {
        const s2::Config c2{cfg.fast_threshold, cfg.K, cfg.harris_k};
        Status sa2 = s2::s2_detect(gray_a_, c2, ka, ws_s2_a_, bytes_s2_, stream_);
        Status sb2 = s2::s2_detect(gray_b_, c2, kb, ws_s2_b_, bytes_s2_, stream_);
        if (sa2 != Status::OK || sb2 != Status::OK) return fail(status_name(sa2 != Status::OK ? sa2 : sb2), "s2");
        check(hipEventRecord(ev_[2], stream_), "event");
        Status sa3 = s3::s3_describe(gray_a_, ka, desc_a_, ws_s3_, bytes_s3_, stream_);
        Status sb3 = s3::s3_describe(gray_b_, kb, desc_b_, ws_s3_, bytes_s3_, stream_);
        if (sa3 != Status::OK || sb3 != Status::OK) return fail(status_name(sa3 != Status::OK ? sa3 : sb3), "s3");
        check(hipEventRecord(ev_[3], stream_), "event");
    }
    s = s4::s4_match(ka, desc_a_, kb, desc_b_, {cfg.ratio_threshold, cfg.cross_check}, matches_, ws_s4_, bytes_s4_,
                     stream_);
    if (s != Status::OK) return fail(status_name(s), "s4");
    check(hipEventRecord(ev_[4], stream_), "event");
    s = s5::estimate(ka, kb, matches_, sa, sb,
                     {cfg.R, 0, cfg.msac_batch_size, cfg.msac_refinements, cfg.msac_threshold, cfg.seed, 1},
                     inlier_mask_, pinned_h_, ws_s5_, bytes_s5_, stream_);
    if (s != Status::OK) return fail(status_name(s), "s5");
    check(hipEventRecord(ev_[5], stream_), "event");
    check(hipMemcpyAsync(pinned_counts_, kp_a_.d_count, sizeof(int), hipMemcpyDeviceToHost, stream_), "d2h");
    check(hipMemcpyAsync(pinned_counts_ + 1, kp_b_.d_count, sizeof(int), hipMemcpyDeviceToHost, stream_), "d2h");
    // Host synchronization 1: the homography is needed on the host for the canvas.
    check(hipStreamSynchronize(stream_), "sync");
    rec->num_kp_a = pinned_counts_[0];
    rec->num_kp_b = pinned_counts_[1];
    const HomographyResult& h = *pinned_h_;
    rec->num_matches = h.num_matches;
    rec->num_inliers = h.num_inliers;
    rec->inlier_ratio = h.inlier_ratio;
    rec->inlier_coverage = h.inlier_coverage;
    rec->reproj_residual = h.reproj_residual;
    std::memcpy(rec->H, h.H, sizeof rec->H);
    auto stage_times = [&](bool with_s6) {
        float ms = 0;
        hipEventElapsedTime(&ms, ev_[0], ev_[1]); rec->t_s1 = ms;
        hipEventElapsedTime(&ms, ev_[1], ev_[2]); rec->t_s2 = ms;
        hipEventElapsedTime(&ms, ev_[2], ev_[3]); rec->t_s3 = ms;
        hipEventElapsedTime(&ms, ev_[3], ev_[4]); rec->t_s4 = ms;
        hipEventElapsedTime(&ms, ev_[4], ev_[5]); rec->t_s5 = ms;
        rec->t_s6 = 0;
        if (with_s6) {
            hipEventElapsedTime(&ms, ev_[5], ev_[6]); rec->t_s6 = ms;
            hipEventElapsedTime(&ms, ev_[0], ev_[6]); last_gpu_span_ms_ = ms;
        }
    };
    if (h.status != Status::OK) {
        rec->t_e2e = ms_since(t0);
        stage_times(false);
        return fail(status_name(h.status), "s5");
    }
    if (collect_features_) {
        // Gate observations of this attempt (harness/pipeline_runner.cu:445-458).
        const auto begin = Clock::now();
        const s4::Top2* forward = nullptr;
        if (s4::s4_forward_neighbors(ka.capacity, kb.capacity, {cfg.ratio_threshold, cfg.cross_check}, ws_s4_,
                                     bytes_s4_, &forward) != Status::OK)
            throw std::runtime_error("Missing matching feature observations");
        const Status fs = features::extract_gate(use_a_, use_b_, ka, kb, matches_, inlier_mask_, h, sa, sb, d_features_,
                                                 pinned_features_, stream_, forward, 7);
        if (fs == Status::CUDA_ERROR) throw std::runtime_error("Gate feature device error");
        check(hipStreamSynchronize(stream_), "sync");
        if (fs == Status::OK) {
            std::copy_n(pinned_features_, features::kGateCount, rec->gate_features.begin());
            rec->features_present = true;
        }
        rec->t_features = ms_since(begin);
    }
    s6::Geometry geom{};
    Status gs = s6::prepare_geometry(wa_, ha_, wb_, hb_, h.H, s6::GeometryPolicy{}, &geom);
    if (gs == Status::OK && (geom.width > canvas_cap_w_ || geom.height > canvas_cap_h_)) gs = Status::INSUFFICIENT_WORKSPACE;
    if (gs != Status::OK) {
        rec->t_e2e = ms_since(t0);
        stage_times(false);
        return fail(gs == Status::INSUFFICIENT_WORKSPACE ? "ERR_CANVAS_OVERFLOW" : status_name(gs), "s6");
    }
    canvas_.width = canvas_mask_.width = geom.width;
    canvas_.height = canvas_mask_.height = geom.height;
    s = s6::render(use_a_, use_b_, geom, cfg.blend_mode, canvas_, canvas_mask_, ws_s6_, bytes_s6_, stream_, stats_);
    if (s != Status::OK) {
        rec->t_e2e = ms_since(t0);
        return fail(status_name(s), "s6");
    }
    check(hipEventRecord(ev_[6], stream_), "event");
    check(hipMemcpyAsync(pinned_stats_, stats_, sizeof(*stats_), hipMemcpyDeviceToHost, stream_), "d2h");
    // Host synchronization 2: the panorama is ready.
    check(hipStreamSynchronize(stream_), "sync");
    rec->t_e2e = ms_since(t0);
    stage_times(true);
    rec->canvas_pixels = static_cast<std::uint64_t>(geom.width) * geom.height;
    rec->canvas_width = geom.width;
    rec->canvas_height = geom.height;
    rec->overlap_pixels = pinned_stats_->overlap_pixels;
    rec->num_levels = s6::effective_reductions(geom, cfg.blend_mode);
    rec->peak_workspace_bytes = bytes_s6_;
    rec->status = "OK";
    rec->fail_stage.clear();
}

void PipelineRunner::execute_adaptive_run(controller::Controller& ctl, std::uint64_t seed, RunRecord* rec) {
    *rec = RunRecord{};
    rec->seed = seed;
    rec->config_id = "adaptive_unstarted";
    attempt_trace_.count = 0;
    const auto t0 = Clock::now();
    auto timed = [&](double& total, auto&& submit) {
        check(hipEventRecord(ev_[0], stream_), "event");
        const Status s = submit();
        check(hipEventRecord(ev_[1], stream_), "event");
        const auto w0 = Clock::now();
        check(hipStreamSynchronize(stream_), "sync");
        rec->t_stream_wait += ms_since(w0);
        float ms = 0;
        hipEventElapsedTime(&ms, ev_[0], ev_[1]);
        total += ms;
        return s;
    };
    auto finish = [&](const char* status, const char* stage) {
        rec->status = status;
        rec->fail_stage = stage;
        rec->t_e2e = ms_since(t0);
    };
    ctl.reset_pair();
    const double probe_s = ctl.probe_scale();
    ScaleInfo pa{}, pb{};
    if (make_scale_info(wa_, ha_, probe_s, &pa) != Status::OK || make_scale_info(wb_, hb_, probe_s, &pb) != Status::OK)
        return finish("INVALID_PROBE_SCALE", "controller");
    gray_a_.width = pa.work_w;
    gray_a_.height = pa.work_h;
    gray_b_.width = pb.work_w;
    gray_b_.height = pb.work_h;
    // Learned policies read the scene features at the probe scale; their
    // first two values are the probe gradients (pipeline_runner.cu:607-645).
    const bool learned = ctl.uses_learned_features();
    if (probe_enabled_) {
        const Status p = timed(rec->t_probe, [&] {
            const Status s = s1::s1_preprocess(use_a_, use_b_, probe_s, &gray_a_, &gray_b_, &pa, &pb, stream_);
            if (s != Status::OK) return s;
            if (learned)
                return features::extract_scene(gray_a_, gray_b_, d_features_, pinned_features_, stream_,
                                               ctl.uses_pair_relations());
            return s1::s1_compute_gradient(gray_a_, gray_b_, ws_probe_, pinned_grad_, pinned_grad_ + 1, stream_);
        });
        rec->t_s1 = rec->t_probe;
        if (p != Status::OK) return finish(status_name(p), "s1");
        if (learned) {
            pinned_grad_[0] = pinned_features_[0];
            pinned_grad_[1] = pinned_features_[1];
            std::copy_n(pinned_features_, features::kSceneCount, rec->scene_features.begin());
        }
        rec->probe_gradient_a = pinned_grad_[0];
        rec->probe_gradient_b = pinned_grad_[1];
    } else {
        if (learned) return finish("INVALID_ARGUMENT", "controller");
        pinned_grad_[0] = pinned_grad_[1] = 1.0;
    }
    const controller::SceneSignals scene{{pa, pinned_grad_[0]}, {pb, pinned_grad_[1]}, rec->scene_features, learned};
    controller::Knobs knobs{};
    {
        const auto c0 = Clock::now();
        const Status b = ctl.begin_pair(scene, &knobs);
        rec->t_controller += ms_since(c0);
        if (b != Status::OK) return finish(status_name(b), "controller");
    }
    controller::Knobs previous{};
    std::uint64_t generation = 1;
    int completed_r = 0;
    ScaleInfo sa{}, sb{};
    HomographyResult accepted{};
    for (int attempt = 0;; ++attempt) {
        if (attempt >= 16) return finish("ATTEMPT_CAPACITY_EXCEEDED", "controller");
        rec->attempt_count = attempt + 1;
        AttemptRecord& tr = attempt_trace_.records[attempt];
        tr = AttemptRecord{};
        tr.knobs = knobs;
        attempt_trace_.count = attempt + 1;
        const auto a0 = Clock::now();
        const double s1_0 = rec->t_s1, s2_0 = rec->t_s2, s3_0 = rec->t_s3, s4_0 = rec->t_s4, s5_0 = rec->t_s5;
        const auto close_attempt = [&] {
            tr.t_s1 = rec->t_s1 - s1_0;
            tr.t_s2 = rec->t_s2 - s2_0;
            tr.t_s3 = rec->t_s3 - s3_0;
            tr.t_s4 = rec->t_s4 - s4_0;
            tr.t_s5 = rec->t_s5 - s5_0;
            tr.t_attempt = ms_since(a0);
        };
        rec->S = knobs.S;
        rec->K = knobs.K;
        rec->R = knobs.R;
        bool rerun_s1 = knobs.S != probe_s || !probe_enabled_;
        bool rerun_features = true, rerun_s4 = true;
        int previous_r = 0;
        if (attempt > 0) {
            controller::RetryPlan plan{};
            if (controller::plan_retry(previous, knobs, &plan) != Status::OK)
                return finish("INVALID_RETRY_PLAN", "controller");
            rerun_s1 = plan.restart_from == controller::RestartFrom::S1;
            tr.restart_from = plan.restart_from == controller::RestartFrom::S1   ? "S1"
                              : plan.restart_from == controller::RestartFrom::S2 ? "S2"
                                                                                 : "S5";
            rerun_features = rerun_s1 || plan.restart_from == controller::RestartFrom::S2;
            if (plan.change_generation) {
                ++generation;
                completed_r = 0;
            }
            rerun_s4 = rerun_features;
            if (plan.extend_s5) {
                if (reuse_enabled_) previous_r = completed_r;
                else rerun_s4 = true;
            }
        }
        if (make_scale_info(wa_, ha_, knobs.S, &sa) != Status::OK || make_scale_info(wb_, hb_, knobs.S, &sb) != Status::OK ||
            knobs.K > max_k_ || knobs.R > max_r_)
            return finish("INVALID_ARGUMENT", "controller");
        if (rerun_s1) {
            gray_a_.width = sa.work_w;
            gray_a_.height = sa.work_h;
            gray_b_.width = sb.work_w;
            gray_b_.height = sb.work_h;
            const Status s = timed(rec->t_s1, [&] {
                return s1::s1_preprocess(use_a_, use_b_, knobs.S, &gray_a_, &gray_b_, &sa, &sb, stream_);
            });
            if (s != Status::OK) {
                tr.execution_status = s;
                tr.fail_stage = "s1";
                close_attempt();
                return finish(status_name(s), "s1");
            }
        }
        Keypoints ka = kp_a_, kb = kp_b_;
        ka.capacity = kb.capacity = knobs.K;
        if (rerun_features) {
            Status s = timed(rec->t_s2, [&] {
                const s2::Config c{20, knobs.K, 0.04f};
                const Status x = s2::s2_detect(gray_a_, c, ka, ws_s2_a_, bytes_s2_, stream_);
                return x == Status::OK ? s2::s2_detect(gray_b_, c, kb, ws_s2_b_, bytes_s2_, stream_) : x;
            });
            if (s != Status::OK) {
                tr.execution_status = s;
                tr.fail_stage = "s2";
                close_attempt();
                return finish(status_name(s), "s2");
            }
            s = timed(rec->t_s3, [&] {
                const Status x = s3::s3_describe(gray_a_, ka, desc_a_, ws_s3_, bytes_s3_, stream_);
                return x == Status::OK ? s3::s3_describe(gray_b_, kb, desc_b_, ws_s3_, bytes_s3_, stream_) : x;
            });
            if (s != Status::OK) {
                tr.execution_status = s;
                tr.fail_stage = "s3";
                close_attempt();
                return finish(status_name(s), "s3");
            }
        }
        if (rerun_s4) {
            const Status s = timed(rec->t_s4, [&] {
                return s4::s4_match(ka, desc_a_, kb, desc_b_, {0.8, false}, matches_, ws_s4_, bytes_s4_, stream_);
            });
            if (s != Status::OK) {
                tr.execution_status = s;
                tr.fail_stage = "s4";
                close_attempt();
                return finish(status_name(s), "s4");
            }
        }
        *pinned_h_ = HomographyResult{};
        pinned_h_->status = Status::INVALID_ARGUMENT;
        if (previous_r > 0) rec->cache_hit = 1;
        const Status s5s = timed(rec->t_s5, [&] {
            return s5::estimate(ka, kb, matches_, sa, sb, {knobs.R, previous_r, 64, 1, 3.0, seed, generation},
                                inlier_mask_, pinned_h_, ws_s5_, bytes_s5_, stream_);
        });
        tr.previous_R = previous_r;
        tr.generation = generation;
        tr.execution_status = s5s;
        if (s5s != Status::OK) tr.fail_stage = "s5";
        tr.has_feedback = s5s == Status::OK;
        if (tr.has_feedback) tr.homography = *pinned_h_;
        tr.has_homography = tr.has_feedback && tr.homography.status == Status::OK;
        if (s5s == Status::OK) completed_r = knobs.R;
        controller::Feedback fb{knobs, s5s, *pinned_h_};
        // Gate observations need a usable homography: without one there is
        // nothing to observe and the reference skips the extraction.
        if (ctl.needs_gate_features() && tr.has_homography) {
            // Gate observations of this attempt (pipeline_runner.cu:780-806).
            const auto f0 = Clock::now();
            const unsigned relations = ctl.relation_mask();
            const s4::Top2* forward = nullptr;
            if ((relations & features::kMatchingRelations) &&
                s4::s4_forward_neighbors(ka.capacity, kb.capacity, {0.8, false}, ws_s4_, bytes_s4_, &forward) !=
                    Status::OK)
                throw std::runtime_error("Missing adaptive matching observations");
            const Status fs = features::extract_gate(use_a_, use_b_, ka, kb, matches_, inlier_mask_, *pinned_h_, sa, sb,
                                                     d_features_, pinned_features_, stream_, forward, relations);
            if (fs == Status::CUDA_ERROR) throw std::runtime_error("Gate feature device error");
            const auto w0 = Clock::now();
            check(hipStreamSynchronize(stream_), "sync");
            rec->t_stream_wait += ms_since(w0);
            if (fs == Status::OK) {
                std::copy_n(pinned_features_, features::kGateCount, fb.features.begin());
                fb.features_present = true;
            }
            tr.t_features = ms_since(f0);
            rec->t_features += tr.t_features;
        }
        tr.features = fb.features;
        rec->gate_features = fb.features;
        rec->features_present = fb.features_present;
        controller::Decision d{};
        const auto c0 = Clock::now();
        const Status observed = ctl.observe(fb, &d);
        rec->t_controller += ms_since(c0);
        close_attempt();
        if (observed == Status::OK) {
            tr.decision = d;
            tr.has_decision = true;
        }
        if (observed != Status::OK) return finish(status_name(observed), s5s == Status::OK ? "controller" : "s5");
        if (d.action == controller::Action::RETRY) {
            previous = knobs;
            knobs = d.next;
            continue;
        }
        if (d.action == controller::Action::REJECT) {
            if (s5s != Status::OK) return finish(status_name(s5s), "s5");
            return finish(d.reason == controller::Reason::BUDGET_EXHAUSTED  ? "BUDGET_EXHAUSTED"
                          : d.reason == controller::Reason::EXECUTION_ERROR ? "EXECUTION_ERROR"
                                                                              : "REJECTED",
                          "controller");
        }
        accepted = *pinned_h_;
        break;
    }
    s6::Geometry geom{};
    const Status gs = s6::prepare_geometry(wa_, ha_, wb_, hb_, accepted.H, {}, &geom);
    if (gs != Status::OK)
        return finish(gs == Status::INSUFFICIENT_WORKSPACE ? "ERR_CANVAS_OVERFLOW" : status_name(gs), "s6");
    if (geom.width > canvas_cap_w_ || geom.height > canvas_cap_h_) return finish("ERR_CANVAS_OVERFLOW", "s6");
    canvas_.width = canvas_mask_.width = geom.width;
    canvas_.height = canvas_mask_.height = geom.height;
    const Status rs = timed(rec->t_s6, [&] {
        const Status x =
            s6::render(use_a_, use_b_, geom, knobs.B, canvas_, canvas_mask_, ws_s6_, bytes_s6_, stream_, stats_);
        if (x == Status::OK) {
            hipMemcpyAsync(pinned_stats_, stats_, sizeof(*stats_), hipMemcpyDeviceToHost, stream_);
            hipMemcpyAsync(pinned_counts_, kp_a_.d_count, sizeof(int), hipMemcpyDeviceToHost, stream_);
            hipMemcpyAsync(pinned_counts_ + 1, kp_b_.d_count, sizeof(int), hipMemcpyDeviceToHost, stream_);
        }
        return x;
    });
    rec->t_e2e = ms_since(t0);
    if (rs != Status::OK) return finish(status_name(rs), "s6");
    rec->num_kp_a = pinned_counts_[0];
    rec->num_kp_b = pinned_counts_[1];
    rec->num_matches = accepted.num_matches;
    rec->num_inliers = accepted.num_inliers;
    rec->inlier_ratio = accepted.inlier_ratio;
    rec->inlier_coverage = accepted.inlier_coverage;
    rec->reproj_residual = accepted.reproj_residual;
    std::memcpy(rec->H, accepted.H, sizeof rec->H);
    rec->blend_mode = static_cast<int>(knobs.B);
    rec->canvas_pixels = static_cast<std::uint64_t>(geom.width) * geom.height;
    rec->canvas_width = geom.width;
    rec->canvas_height = geom.height;
    rec->overlap_pixels = pinned_stats_->overlap_pixels;
    rec->num_levels = s6::effective_reductions(geom, knobs.B);
    rec->peak_workspace_bytes = bytes_s6_;
    rec->status = "OK";
    rec->fail_stage.clear();
}

}  // namespace stitch::harness
