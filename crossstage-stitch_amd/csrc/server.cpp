// Served form (see stitch/server.hpp). The per-request flow is the one of
// the reference's PipelineRunner::execute_adaptive_run
// (harness/pipeline_runner.cu:531-885), split at every point where the host
// needs a device result so that many requests share each launch.
#include "stitch/server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace stitch::serve {
namespace {

using Clock = std::chrono::steady_clock;

enum class Phase { ProbeReady, ProbeRunning, AttemptReady, AttemptRunning, RenderReady, RenderRunning, Done };

constexpr int kMaxAttempts = 16;

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double pos = q * (v.size() - 1);
    const std::size_t lo = static_cast<std::size_t>(std::floor(pos));
    const std::size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - lo);
}

// Reservation for the canvas of a request, made at admission: the render
// pyramid of a canvas canvas_factor times the larger input.
std::size_t canvas_reservation(const ImageView& a, const ImageView& b, double factor, s6::BlendMode mode) {
    const double px = std::max(static_cast<double>(a.width) * a.height, static_cast<double>(b.width) * b.height);
    const double side = std::sqrt(factor);
    s6::Geometry g{};
    g.width = static_cast<int>(std::ceil(side * std::max(a.width, b.width)));
    g.height = static_cast<int>(std::ceil(px * factor / g.width));
    g.width_a = a.width;
    g.height_a = a.height;
    g.width_b = b.width;
    g.height_b = b.height;
    g.inverse[0] = g.inverse[4] = g.inverse[8] = 1;
    g.denominator_tolerance = 1e-12;
    g.snap_tolerance = 1e-9;
    return batch::render_bytes(g, mode);
}

s6::BlendMode deepest_mode(const std::vector<controller::Knobs>& ladder) {
    s6::BlendMode best = s6::BlendMode::Feather;
    for (const auto& k : ladder)
        if (s6::requested_reductions(k.B) > s6::requested_reductions(best)) best = k.B;
    return best;
}

}  // namespace

struct Server::Live {
    Request req{};
    Phase phase{Phase::ProbeReady};
    batch::PairBuffers buf{};
    batch::RenderBuffers rbuf{};
    controller::Controller ctl{};
    controller::Knobs knobs{}, previous{};
    int attempt{0};
    std::uint64_t generation{1};
    int completed_r{0}, previous_r{0};
    bool rerun_s1{true}, rerun_features{true}, rerun_s4{true};
    std::size_t reserved{0};
    Completion done{};
};

Server::Server() = default;

Server::~Server() {
    features_.wait();
    render_.wait();
    if (obs_stream_) hipStreamSynchronize(obs_stream_);
    if (obs_device_) hipFree(obs_device_);
    if (obs_host_) hipHostFree(obs_host_);
    if (obs_stream_) hipStreamDestroy(obs_stream_);
}

void Server::initialize(const ServerConfig& config) {
    config_ = config;
    policy_ = config.policy_json.empty() ? controller::legacy_controller_policy()
                                         : controller::parse_controller_policy(config.policy_json);
    {
        controller::Controller probe;
        if (policy_.configure(probe) != Status::OK) throw admission::Error("INVALID_POLICY");
        learned_ = probe.uses_learned_features();
    }
    if (learned_) {
        if (hipStreamCreateWithFlags(&obs_stream_, hipStreamNonBlocking) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&obs_device_), features::kDeviceDoubles * sizeof(double)) !=
                hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&obs_host_), features::kGateCount * sizeof(double),
                          hipHostMallocDefault) != hipSuccess)
            throw std::runtime_error("observation buffers");
    }
    for (const auto& k : policy_.ladder) {
        max_k_ = std::max(max_k_, k.K);
        max_r_ = std::max(max_r_, k.R);
    }
    pool_.initialize(config.hbm_budget);
    features_.initialize(config.max_batch, config.use_graphs);
    render_.initialize(config.max_batch, config.use_graphs);
}

void Server::release(std::uint64_t id) {
    for (auto it = kept_.begin(); it != kept_.end(); ++it) {
        if ((*it)->req.id != id) continue;
        batch::release_render(pool_, &(*it)->rbuf);
        batch::release_pair(pool_, &(*it)->buf);
        kept_.erase(it);
        return;
    }
}

void Server::run_trace(const Request* requests, int n, std::vector<Completion>* out, ServeStats* stats) {
    out->clear();
    ServeStats st{};
    const auto t0 = Clock::now();
    const auto now_ms = [&] { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); };
    std::vector<std::unique_ptr<Live>> live;
    std::vector<Live*> in_features, in_render;
    std::size_t reserved = 0;
    for (const auto& k : kept_) reserved += k->reserved;
    const s6::BlendMode reserve_mode = deepest_mode(policy_.ladder);
    const double probe_s = policy_.probe_scale;
    int next = 0;
    long long feature_tasks = 0, render_tasks = 0;

    auto finish = [&](Live& L, const std::string& status, const std::string& stage) {
        L.done.status = status;
        L.done.fail_stage = stage;
        L.done.done_ms = now_ms();
        L.done.attempts = L.attempt + 1;
        L.done.knobs = L.knobs;
        L.phase = Phase::Done;
    };
    // Stage flags of the next attempt after a RETRY (pipeline_runner.cu:641-660).
    auto plan_next = [&](Live& L) {
        controller::RetryPlan plan{};
        if (controller::plan_retry(L.previous, L.knobs, &plan) != Status::OK) {
            finish(L, "INVALID_RETRY_PLAN", "controller");
            return;
        }
        L.rerun_s1 = plan.restart_from == controller::RestartFrom::S1;
        L.rerun_features = L.rerun_s1 || plan.restart_from == controller::RestartFrom::S2;
        if (plan.change_generation) {
            ++L.generation;
            L.completed_r = 0;
        }
        L.rerun_s4 = L.rerun_features;
        L.previous_r = plan.extend_s5 ? L.completed_r : 0;
    };
    auto check_knobs = [&](Live& L) {
        ScaleInfo sa{}, sb{};
        if (make_scale_info(L.req.a.width, L.req.a.height, L.knobs.S, &sa) != Status::OK ||
            make_scale_info(L.req.b.width, L.req.b.height, L.knobs.S, &sb) != Status::OK || L.knobs.K > L.buf.cap.K ||
            L.knobs.R > L.buf.cap.R) {
            finish(L, "INVALID_ARGUMENT", "controller");
            return false;
        }
        return true;
    };
    auto to_render = [&](Live& L) {
        const Status gs = s6::prepare_geometry(L.req.a.width, L.req.a.height, L.req.b.width, L.req.b.height,
                                               L.done.h.H, s6::GeometryPolicy{}, &L.done.geometry);
        if (gs != Status::OK) {
            finish(L, gs == Status::INSUFFICIENT_WORKSPACE ? "ERR_CANVAS_OVERFLOW" : status_name(gs), "s6");
            return;
        }
        L.phase = Phase::RenderReady;
    };
    auto on_features = [&](Live& L, const batch::FeatureResult& r) {
        L.done.num_kp_a = r.num_kp_a;
        L.done.num_kp_b = r.num_kp_b;
        if (L.phase == Phase::ProbeRunning) {
            if (r.status != Status::OK) return finish(L, status_name(r.status), "s1");
            controller::SceneSignals scene{{r.sa, r.grad_a}, {r.sb, r.grad_b}};
            if (learned_) {
                // Scene features of the probe's gray pair; their first two
                // values are the probe gradients (pipeline_runner.cu:607-645).
                ImageView ga = L.buf.gray_a, gb = L.buf.gray_b;
                ga.width = r.sa.work_w;
                ga.height = r.sa.work_h;
                gb.width = r.sb.work_w;
                gb.height = r.sb.work_h;
                const Status fs = features::extract_scene(ga, gb, obs_device_, obs_host_, obs_stream_,
                                                          L.ctl.uses_pair_relations());
                if (fs != Status::OK || hipStreamSynchronize(obs_stream_) != hipSuccess)
                    return finish(L, status_name(fs != Status::OK ? fs : Status::CUDA_ERROR), "s1");
                std::copy_n(obs_host_, features::kSceneCount, scene.features.begin());
                scene.features_present = true;
                scene.a.mean_abs_gradient = scene.features[0];
                scene.b.mean_abs_gradient = scene.features[1];
            }
            if (const Status b = L.ctl.begin_pair(scene, &L.knobs); b != Status::OK)
                return finish(L, status_name(b), "controller");
            L.attempt = 0;
            L.rerun_s1 = L.knobs.S != probe_s;
            L.rerun_features = L.rerun_s4 = true;
            L.previous_r = 0;
            if (check_knobs(L)) L.phase = Phase::AttemptReady;
            return;
        }
        const std::string stage = r.fail_stage;
        if (r.status != Status::OK && stage != "s5") return finish(L, status_name(r.status), stage);
        if (!config_.adaptive) {
            if (r.status != Status::OK) return finish(L, status_name(r.status), "s5");
            L.done.h = r.h;
            if (r.h.status != Status::OK) return finish(L, status_name(r.h.status), "s5");
            return to_render(L);
        }
        const Status s5s = r.status;
        HomographyResult h = r.h;
        if (s5s != Status::OK) {
            h = HomographyResult{};
            h.status = Status::INVALID_ARGUMENT;
        }
        if (s5s == Status::OK) L.completed_r = L.knobs.R;
        L.done.h = h;
        controller::Feedback fb{L.knobs, s5s, h};
        if (learned_ && L.ctl.needs_gate_features() && s5s == Status::OK && h.status == Status::OK) {
            // Gate observations of this attempt (pipeline_runner.cu:780-806).
            const unsigned relations = L.ctl.relation_mask();
            const s4::Top2* forward = nullptr;
            if ((relations & features::kMatchingRelations) &&
                s4::s4_forward_neighbors(L.knobs.K, L.knobs.K, {0.8, false}, L.buf.ws_s4, L.buf.bytes_s4, &forward) !=
                    Status::OK)
                return finish(L, "CUDA_ERROR", "controller");
            Keypoints ka = L.buf.kp_a, kb = L.buf.kp_b;
            ka.capacity = kb.capacity = L.knobs.K;
            const Status fs = features::extract_gate(L.req.a, L.req.b, ka, kb, L.buf.matches, L.buf.inlier_mask, h,
                                                     r.sa, r.sb, obs_device_, obs_host_, obs_stream_, forward,
                                                     relations);
            if (fs == Status::CUDA_ERROR || hipStreamSynchronize(obs_stream_) != hipSuccess)
                return finish(L, "CUDA_ERROR", "controller");
            if (fs == Status::OK) {
                std::copy_n(obs_host_, features::kGateCount, fb.features.begin());
                fb.features_present = true;
            }
        }
        controller::Decision d{};
        const Status observed = L.ctl.observe(fb, &d);
        if (observed != Status::OK) return finish(L, status_name(observed), s5s == Status::OK ? "controller" : "s5");
        if (d.action == controller::Action::RETRY) {
            L.previous = L.knobs;
            L.knobs = d.next;
            if (++L.attempt >= kMaxAttempts) {
                L.attempt = kMaxAttempts - 1;
                return finish(L, "ATTEMPT_CAPACITY_EXCEEDED", "controller");
            }
            plan_next(L);
            if (L.phase != Phase::Done && check_knobs(L)) L.phase = Phase::AttemptReady;
            return;
        }
        if (d.action == controller::Action::REJECT) {
            if (s5s != Status::OK) return finish(L, status_name(s5s), "s5");
            return finish(L,
                          d.reason == controller::Reason::BUDGET_EXHAUSTED  ? "BUDGET_EXHAUSTED"
                          : d.reason == controller::Reason::EXECUTION_ERROR ? "EXECUTION_ERROR"
                                                                              : "REJECTED",
                          "controller");
        }
        L.done.h = h;
        to_render(L);
    };
    auto task_of = [&](Live& L) {
        batch::FeatureTask t{};
        t.a = L.req.a;
        t.b = L.req.b;
        t.buffers = &L.buf;
        if (L.phase == Phase::ProbeReady) {
            t.probe = true;
            t.S = probe_s;
            t.run_s2s3 = t.run_s4 = t.run_s5 = false;
            return t;
        }
        if (!config_.adaptive) {
            t = batch::fixed_task(L.req.fixed, &L.buf);
            t.a = L.req.a;
            t.b = L.req.b;
            t.s5.seed = L.req.seed;
            return t;
        }
        t.S = L.knobs.S;
        t.run_s1 = L.rerun_s1;
        t.run_s2s3 = L.rerun_features;
        t.run_s4 = L.rerun_s4;
        t.s2 = {20, L.knobs.K, 0.04f};
        t.s4 = {0.8, false};
        t.s5 = {L.knobs.R, L.previous_r, 64, 1, 3.0, L.req.seed, L.generation};
        return t;
    };

    while (static_cast<int>(out->size()) < n) {
        bool progress = false;
        // Admission by HBM budget and slots.
        while (next < n && requests[next].arrival_ms <= now_ms() && static_cast<int>(live.size()) < config_.max_live) {
            const Request& q = requests[next];
            auto L = std::make_unique<Live>();
            L->req = q;
            L->done.id = q.id;
            L->done.arrival_ms = q.arrival_ms;
            const batch::PairCapacity cap{std::max(q.a.width, q.b.width), std::max(q.a.height, q.b.height),
                                          config_.adaptive ? max_k_ : q.fixed.K, config_.adaptive ? max_r_ : q.fixed.R};
            std::size_t need = 0;
            bool valid = true;
            try {
                need = batch::pair_bytes(cap) +
                       canvas_reservation(q.a, q.b, config_.canvas_factor,
                                          config_.adaptive ? reserve_mode : q.fixed.blend);
            } catch (const admission::Error&) {
                valid = false;
            }
            const bool idle = live.empty() && in_features.empty() && in_render.empty();
            if (valid && reserved + need > pool_.capacity()) {
                if (!idle) break;  // wait for memory to come back
                valid = false;
            }
            if (valid && !batch::allocate_pair(pool_, cap, &L->buf)) {
                if (!idle) break;
                valid = false;
            }
            ++next;
            progress = true;
            L->done.admitted_ms = now_ms();
            if (!valid) {
                finish(*L, "ERR_DEVICE_OOM", "admission");
                out->push_back(L->done);
                continue;
            }
            L->reserved = need;
            L->done.reserved_bytes = need;
            reserved += need;
            if (config_.adaptive) {
                L->ctl.reset_pair();
                if (policy_.configure(L->ctl) != Status::OK) {
                    finish(*L, "INVALID_POLICY", "controller");
                } else {
                    L->ctl.reset_pair();
                    L->phase = Phase::ProbeReady;
                }
            } else {
                L->knobs = {q.fixed.S, q.fixed.K, q.fixed.R, q.fixed.blend};
                L->phase = Phase::AttemptReady;
            }
            live.push_back(std::move(L));
        }
        // Feature engine: collect the finished launch, start the next one.
        if (features_.ready()) {
            if (!in_features.empty()) {
                const auto& res = features_.results();
                for (std::size_t k = 0; k < in_features.size(); ++k) on_features(*in_features[k], res[k]);
                st.feature_busy_ms += features_.last_ms();
                in_features.clear();
                progress = true;
            }
            std::vector<batch::FeatureTask> tasks;
            for (auto& L : live) {
                if (static_cast<int>(tasks.size()) >= config_.max_batch) break;
                if (L->phase != Phase::ProbeReady && L->phase != Phase::AttemptReady) continue;
                tasks.push_back(task_of(*L));
                L->phase = L->phase == Phase::ProbeReady ? Phase::ProbeRunning : Phase::AttemptRunning;
                in_features.push_back(L.get());
            }
            if (!tasks.empty()) {
                features_.submit(tasks.data(), static_cast<int>(tasks.size()));
                ++st.feature_launches;
                feature_tasks += static_cast<long long>(tasks.size());
                progress = true;
            }
        }
        // Render engine.
        if (render_.ready()) {
            if (!in_render.empty()) {
                const auto& res = render_.results();
                for (std::size_t k = 0; k < in_render.size(); ++k) {
                    Live& L = *in_render[k];
                    if (res[k].status != Status::OK) {
                        finish(L, status_name(res[k].status), "s6");
                    } else {
                        L.done.stats = res[k].stats;
                        L.done.canvas = L.rbuf.canvas;
                        L.done.mask = L.rbuf.mask;
                        finish(L, "OK", "");
                    }
                }
                st.render_busy_ms += render_.last_ms();
                in_render.clear();
                progress = true;
            }
            std::vector<batch::RenderTask> tasks;
            for (auto& L : live) {
                if (static_cast<int>(tasks.size()) >= config_.max_batch) break;
                if (L->phase != Phase::RenderReady) continue;
                const s6::BlendMode mode = config_.adaptive ? L->knobs.B : L->req.fixed.blend;
                bool ok = false;
                try {
                    ok = batch::allocate_render(pool_, L->done.geometry, mode, &L->rbuf);
                } catch (const admission::Error&) {
                    finish(*L, "ERR_CANVAS_OVERFLOW", "s6");
                    continue;
                }
                if (!ok) {
                    // Nothing else holds render memory: this canvas can never fit.
                    if (in_render.empty() && tasks.empty()) {
                        bool others = false;
                        for (auto& M : live)
                            others |= M.get() != L.get() && M->phase != Phase::Done && M->rbuf.block;
                        if (!others) finish(*L, "ERR_DEVICE_OOM", "s6");
                    }
                    continue;
                }
                tasks.push_back({L->req.a, L->req.b, L->done.geometry, mode, &L->rbuf});
                L->phase = Phase::RenderRunning;
                in_render.push_back(L.get());
            }
            if (!tasks.empty()) {
                render_.submit(tasks.data(), static_cast<int>(tasks.size()));
                ++st.render_launches;
                render_tasks += static_cast<long long>(tasks.size());
                progress = true;
            }
        }
        // Retire finished requests.
        for (auto it = live.begin(); it != live.end();) {
            Live& L = **it;
            if (L.phase != Phase::Done) {
                ++it;
                continue;
            }
            out->push_back(L.done);
            progress = true;
            if (L.req.keep_output && L.done.status == "OK") {
                kept_.push_back(std::move(*it));
            } else {
                batch::release_render(pool_, &L.rbuf);
                batch::release_pair(pool_, &L.buf);
                reserved -= L.reserved;
            }
            it = live.erase(it);
        }
        if (!progress) {
            if (live.empty() && next < n) {
                const double wait = requests[next].arrival_ms - now_ms();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(std::min(wait, 1.0)));
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    }
    st.makespan_ms = now_ms();
    std::vector<double> lat;
    for (const auto& c : *out) {
        lat.push_back(c.latency_ms());
        if (c.status == "OK") ++st.completed;
        else ++st.failed;
    }
    st.p50_ms = percentile(lat, 0.50);
    st.p99_ms = percentile(lat, 0.99);
    double sum = 0;
    for (double v : lat) sum += v;
    st.mean_ms = lat.empty() ? 0 : sum / lat.size();
    st.mean_feature_batch = st.feature_launches ? static_cast<double>(feature_tasks) / st.feature_launches : 0;
    st.mean_render_batch = st.render_launches ? static_cast<double>(render_tasks) / st.render_launches : 0;
    st.peak_pool_bytes = pool_.peak();
    if (stats) *stats = st;
}

}  // namespace stitch::serve
