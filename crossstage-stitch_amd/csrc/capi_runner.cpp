// extern "C" layer for the runner, the controller and the synthetic inputs.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>

#include "cs_stitch.h"
#include "stitch/controller.hpp"
#include "stitch/runner.hpp"
#include "stitch/strip.hpp"

using namespace stitch;

struct cs_runner {
    harness::PipelineRunner impl;
    harness::RunRecord last;  // the last run's record (its feature observations)
};

struct cs_controller {
    controller::ControllerPolicy policy;
    controller::Controller impl;
};

struct cs_strip {
    strip::StripStitcher impl;
    cs_strip(int f, int w, int h, int k, int r) : impl(f, w, h, k, r) {}
};

static_assert(sizeof(cs_geometry) == sizeof(s6::Geometry), "Geometry layout");
static_assert(sizeof(cs_homography_result) == sizeof(HomographyResult), "HomographyResult layout");

namespace {

int map_exception() {
    try {
        throw;
    } catch (const admission::Error& e) {
        return e.code == "ERR_DEVICE_OOM" ? CS_INSUFFICIENT_WORKSPACE : CS_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        return CS_INSUFFICIENT_WORKSPACE;
    } catch (...) {
        return CS_DEVICE_ERROR;
    }
}

harness::HarnessConfig config(const cs_harness_config* c) {
    harness::HarnessConfig h;
    h.S = c->S;
    h.K = c->K;
    h.R = c->R;
    h.blend_mode = static_cast<s6::BlendMode>(c->blend_mode);
    h.fast_threshold = c->fast_threshold;
    h.harris_k = c->harris_k;
    h.ratio_threshold = c->ratio_threshold;
    h.cross_check = c->cross_check != 0;
    h.msac_batch_size = c->msac_batch_size;
    h.msac_refinements = c->msac_refinements;
    h.msac_threshold = c->msac_threshold;
    h.seed = c->seed;
    h.pyramid_levels = c->pyramid_levels > 1 ? c->pyramid_levels : 1;
    h.pyramid_scale_factor = c->pyramid_levels > 1 ? c->pyramid_scale_factor : 1.2;
    return h;
}

void put_record(const harness::RunRecord& r, double gpu_span, cs_run_record* o) {
    std::memset(o, 0, sizeof *o);
    o->S = r.S;
    o->K = r.K;
    o->R = r.R;
    o->blend_mode = r.blend_mode;
    o->t_s1 = r.t_s1;
    o->t_s2 = r.t_s2;
    o->t_s3 = r.t_s3;
    o->t_s4 = r.t_s4;
    o->t_s5 = r.t_s5;
    o->t_s6 = r.t_s6;
    o->t_controller = r.t_controller;
    o->t_e2e = r.t_e2e;
    o->t_probe = r.t_probe;
    o->t_stream_wait = r.t_stream_wait;
    o->t_gpu_span = gpu_span;
    o->probe_gradient_a = r.probe_gradient_a;
    o->probe_gradient_b = r.probe_gradient_b;
    o->attempt_count = r.attempt_count;
    o->num_kp_a = r.num_kp_a;
    o->num_kp_b = r.num_kp_b;
    o->num_matches = r.num_matches;
    o->num_inliers = r.num_inliers;
    o->inlier_ratio = r.inlier_ratio;
    o->inlier_coverage = r.inlier_coverage;
    o->reproj_residual = r.reproj_residual;
    std::memcpy(o->H, r.H, sizeof o->H);
    o->canvas_pixels = r.canvas_pixels;
    o->overlap_pixels = r.overlap_pixels;
    o->canvas_width = r.canvas_width;
    o->canvas_height = r.canvas_height;
    o->num_levels = r.num_levels;
    o->cache_hit = r.cache_hit;
    std::strncpy(o->status, r.status.c_str(), sizeof o->status - 1);
    std::strncpy(o->fail_stage, r.fail_stage.c_str(), sizeof o->fail_stage - 1);
}

ImageView view(const cs_image* v) {
    return {v->data, v->width, v->height, v->pitch_bytes, static_cast<PixelFormat>(v->format)};
}

void put_image(const ImageView& v, cs_image* o) {
    o->data = v.data;
    o->width = v.width;
    o->height = v.height;
    o->pitch_bytes = v.pitch_bytes;
    o->format = static_cast<int>(v.format);
}

}  // namespace

extern "C" {

cs_runner* cs_runner_create(void) {
    try {
        return new cs_runner{};
    } catch (...) {
        return nullptr;
    }
}

void cs_runner_destroy(cs_runner* r) { delete r; }

int cs_runner_initialize(cs_runner* r, int max_w, int max_h, int max_k, int max_r, int canvas_w, int canvas_h, int mode,
                         uint64_t reserve) {
    if (!r) return CS_INVALID_ARGUMENT;
    try {
        r->impl.initialize(max_w, max_h, max_k, max_r, canvas_w, canvas_h, static_cast<s6::BlendMode>(mode),
                           static_cast<std::size_t>(reserve));
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_runner_query_memory(int w, int h, int k, int r, int cw, int ch, int mode, uint64_t* device_bytes,
                           uint64_t* workspace_bytes, uint64_t* s6_bytes) {
    try {
        const auto p = harness::PipelineRunner::query_memory(w, h, k, r, cw, ch, static_cast<s6::BlendMode>(mode));
        if (device_bytes) *device_bytes = p.device_bytes;
        if (workspace_bytes) *workspace_bytes = p.workspace_bytes;
        if (s6_bytes) *s6_bytes = p.s6_workspace_bytes;
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_runner_upload_pair(cs_runner* r, const uint8_t* a, int wa, int ha, const uint8_t* b, int wb, int hb) {
    if (!r || !a || !b) return CS_INVALID_ARGUMENT;
    try {
        r->impl.upload_pair(a, wa, ha, b, wb, hb);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_runner_use_device_pair(cs_runner* r, const cs_image* a, const cs_image* b) {
    if (!r || !a || !b) return CS_INVALID_ARGUMENT;
    try {
        r->impl.use_device_pair(view(a), view(b));
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_runner_execute_run(cs_runner* r, const cs_harness_config* c, cs_run_record* out) {
    if (!r || !c || !out) return CS_INVALID_ARGUMENT;
    try {
        harness::RunRecord rec;
        r->impl.execute_run(config(c), &rec);
        put_record(rec, r->impl.last_gpu_span_ms(), out);
        r->last = std::move(rec);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_runner_execute_adaptive_run(cs_runner* r, cs_controller* ctl, uint64_t seed, cs_run_record* out) {
    if (!r || !ctl || !out) return CS_INVALID_ARGUMENT;
    try {
        harness::RunRecord rec;
        r->impl.execute_adaptive_run(ctl->impl, seed, &rec);
        put_record(rec, -1, out);
        r->last = std::move(rec);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_runner_set_collect_features(cs_runner* r, int enabled) {
    if (!r) return CS_INVALID_ARGUMENT;
    r->impl.set_collect_features(enabled != 0);
    return CS_OK;
}

int cs_runner_features(cs_runner* r, double* scene, double* gate, int* present, double* t_features) {
    if (!r) return CS_INVALID_ARGUMENT;
    const harness::RunRecord& rec = r->last;
    if (scene) std::copy(rec.scene_features.begin(), rec.scene_features.end(), scene);
    if (gate) std::copy(rec.gate_features.begin(), rec.gate_features.end(), gate);
    if (present) *present = rec.features_present;
    if (t_features) *t_features = rec.t_features;
    return CS_OK;
}

int cs_runner_attempt_trace(cs_runner* r, cs_attempt_record* out, int capacity, int* count) {
    if (!r || capacity < 0 || (capacity > 0 && !out)) return CS_INVALID_ARGUMENT;
    const harness::AttemptTrace& t = r->impl.attempt_trace();
    if (count) *count = t.count;
    for (int i = 0; i < capacity && i < t.count; ++i) {
        const harness::AttemptRecord& a = t.records[i];
        cs_attempt_record& o = out[i];
        std::memset(&o, 0, sizeof o);
        o.S = a.knobs.S;
        o.K = a.knobs.K;
        o.R = a.knobs.R;
        o.blend_mode = static_cast<int>(a.knobs.B);
        o.action = static_cast<int>(a.decision.action);
        o.reason = static_cast<int>(a.decision.reason);
        o.next_S = a.decision.next.S;
        o.next_K = a.decision.next.K;
        o.next_R = a.decision.next.R;
        o.next_blend_mode = static_cast<int>(a.decision.next.B);
        o.previous_R = a.previous_R;
        o.generation = a.generation;
        std::memcpy(&o.homography, &a.homography, sizeof o.homography);
        o.t_s1 = a.t_s1;
        o.t_s2 = a.t_s2;
        o.t_s3 = a.t_s3;
        o.t_s4 = a.t_s4;
        o.t_s5 = a.t_s5;
        o.t_attempt = a.t_attempt;
        std::strncpy(o.restart_from, a.restart_from, sizeof o.restart_from - 1);
        std::strncpy(o.fail_stage, a.fail_stage, sizeof o.fail_stage - 1);
        o.execution_status = static_cast<int>(a.execution_status);
        o.has_decision = a.has_decision ? 1 : 0;
        o.has_homography = a.has_homography ? 1 : 0;
        o.has_feedback = a.has_feedback ? 1 : 0;
    }
    return CS_OK;
}

int cs_runner_set_flags(cs_runner* r, int reuse_enabled, int probe_enabled) {
    if (!r) return CS_INVALID_ARGUMENT;
    r->impl.set_reuse_enabled(reuse_enabled != 0);
    r->impl.set_probe_enabled(probe_enabled != 0);
    return CS_OK;
}

int cs_runner_outputs(cs_runner* r, cs_image* canvas, cs_image* mask, cs_keypoints* ka, cs_keypoints* kb,
                      cs_descriptors* da, cs_descriptors* db, cs_matches* m, uint8_t** inlier_mask,
                      cs_homography_result* h) {
    if (!r) return CS_INVALID_ARGUMENT;
    const auto& R = r->impl;
    if (canvas) put_image(R.canvas(), canvas);
    if (mask) put_image(R.canvas_mask(), mask);
    auto put_kp = [](const Keypoints& k, cs_keypoints* o) {
        o->x = k.x;
        o->y = k.y;
        o->response = k.response;
        o->capacity = k.capacity;
        o->d_count = k.d_count;
    };
    auto put_desc = [](const Descriptors& d, cs_descriptors* o) {
        o->type = static_cast<int>(d.type);
        o->dim = d.dim;
        o->row_stride_bytes = d.row_stride_bytes;
        o->data = d.data;
    };
    if (ka) put_kp(R.keypoints_a(), ka);
    if (kb) put_kp(R.keypoints_b(), kb);
    if (da) put_desc(R.descriptors_a(), da);
    if (db) put_desc(R.descriptors_b(), db);
    if (m) {
        m->idx_a = R.matches().idx_a;
        m->idx_b = R.matches().idx_b;
        m->distance = R.matches().distance;
        m->capacity = R.matches().capacity;
        m->d_count = R.matches().d_count;
    }
    if (inlier_mask) *inlier_mask = const_cast<uint8_t*>(R.inlier_mask());
    if (h) std::memcpy(h, &R.homography(), sizeof *h);
    return CS_OK;
}

void* cs_runner_stream(cs_runner* r) { return r ? static_cast<void*>(r->impl.stream()) : nullptr; }

cs_controller* cs_controller_create_legacy(int blind) {
    try {
        auto* c = new cs_controller{};
        c->policy = blind ? controller::legacy_blind_s05_policy() : controller::legacy_controller_policy();
        if (c->policy.configure(c->impl) != Status::OK) {
            delete c;
            return nullptr;
        }
        return c;
    } catch (...) {
        return nullptr;
    }
}

void cs_controller_destroy(cs_controller* c) { delete c; }

int cs_plan_retry(double s0, int k0, int r0, int b0, double s1, int k1, int r1, int b1, int* out4) {
    if (!out4) return CS_INVALID_ARGUMENT;
    controller::RetryPlan p{};
    const Status st = controller::plan_retry({s0, k0, r0, static_cast<s6::BlendMode>(b0)},
                                             {s1, k1, r1, static_cast<s6::BlendMode>(b1)}, &p);
    out4[0] = static_cast<int>(p.restart_from);
    out4[1] = p.extend_s5;
    out4[2] = p.previous_hypotheses;
    out4[3] = p.change_generation;
    return static_cast<int>(st);
}

int cs_controller_walk(cs_controller* c, double grad_a, double grad_b, int wa, int ha, int wb, int hb,
                       const cs_homography_result* feedback, const int* exec_status, int n, double* trace) {
    return cs_controller_walk_features(c, grad_a, grad_b, wa, ha, wb, hb, nullptr, feedback, exec_status, nullptr,
                                       nullptr, n, trace);
}

cs_controller* cs_controller_create_policy(const char* text, char* error, size_t error_len) {
    auto say = [&](const char* what) {
        if (error && error_len) {
            std::strncpy(error, what, error_len - 1);
            error[error_len - 1] = 0;
        }
    };
    if (!text) {
        say("no policy text");
        return nullptr;
    }
    try {
        auto* c = new cs_controller{};
        c->policy = controller::parse_controller_policy(text);
        if (c->policy.configure(c->impl) != Status::OK) {
            delete c;
            say("Invalid Controller policy or ladder");
            return nullptr;
        }
        say("");
        return c;
    } catch (const std::exception& e) {
        say(e.what());
        return nullptr;
    } catch (...) {
        say("policy error");
        return nullptr;
    }
}

int cs_controller_observations(cs_controller* c, int* learned, int* pair_relations, unsigned* relation_mask,
                               int* needs_gate) {
    if (!c) return CS_INVALID_ARGUMENT;
    if (learned) *learned = c->impl.uses_learned_features();
    if (pair_relations) *pair_relations = c->impl.uses_pair_relations();
    if (relation_mask) *relation_mask = c->impl.relation_mask();
    if (needs_gate) *needs_gate = c->impl.needs_gate_features();
    return CS_OK;
}

int cs_controller_walk_features(cs_controller* c, double grad_a, double grad_b, int wa, int ha, int wb, int hb,
                                const double* scene, const cs_homography_result* feedback, const int* exec_status,
                                const double* gate, const int* gate_present, int n, double* trace) {
    if (!c || !trace || (n > 0 && (!feedback || !exec_status))) return -1;
    auto& ctl = c->impl;
    ctl.reset_pair();
    ScaleInfo pa{}, pb{};
    make_scale_info(wa, ha, ctl.probe_scale(), &pa);
    make_scale_info(wb, hb, ctl.probe_scale(), &pb);
    controller::SceneSignals signals{{pa, grad_a}, {pb, grad_b}};
    if (scene) {
        std::copy(scene, scene + features::kSceneCount, signals.features.begin());
        signals.features_present = true;
    }
    controller::Knobs knobs{};
    if (ctl.begin_pair(signals, &knobs) != Status::OK) return -2;
    int rows = 0;
    auto put = [&](int action, int reason, const controller::Knobs& k) {
        double* t = trace + 5 * rows++;
        t[0] = action;
        t[1] = reason;
        t[2] = k.S;
        t[3] = k.K;
        t[4] = k.R;
    };
    put(-1, -1, knobs);
    for (int i = 0; i < n; ++i) {
        controller::Feedback fb{knobs, static_cast<Status>(exec_status[i]),
                                *reinterpret_cast<const HomographyResult*>(&feedback[i])};
        if (gate && gate_present && gate_present[i]) {
            std::copy(gate + static_cast<std::size_t>(i) * features::kGateCount,
                      gate + static_cast<std::size_t>(i + 1) * features::kGateCount, fb.features.begin());
            fb.features_present = true;
        }
        controller::Decision d{};
        if (ctl.observe(fb, &d) != Status::OK) {
            put(-3, -3, knobs);
            break;
        }
        put(static_cast<int>(d.action), static_cast<int>(d.reason), d.next);
        if (d.action != controller::Action::RETRY) break;
        knobs = d.next;
    }
    return rows;
}

int cs_strip_create(int max_frames, int max_w, int max_h, int max_k, int max_r, cs_strip** out) {
    if (!out) return CS_INVALID_ARGUMENT;
    *out = nullptr;
    if (max_frames < 2 || max_frames > s6::kMaxChainFrames || max_w < 8 || max_h < 8 || max_k < 1 || max_r < 1)
        return CS_INVALID_ARGUMENT;
    try {
        *out = new cs_strip(max_frames, max_w, max_h, max_k, max_r);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

void cs_strip_destroy(cs_strip* s) { delete s; }

int cs_strip_run(cs_strip* s, int n, const cs_image* frames, const cs_harness_config* c, cs_strip_record* rec,
                 cs_homography_result* pairs, int* keypoints) {
    if (!s || !frames || !c || !rec || !pairs || n < 2 || n > s6::kMaxChainFrames) return CS_INVALID_ARGUMENT;
    std::memset(rec, 0, sizeof *rec);
    try {
        ImageView views[s6::kMaxChainFrames]{};
        for (int k = 0; k < n; ++k)
            views[k] = {frames[k].data, frames[k].width, frames[k].height, frames[k].pitch_bytes,
                        static_cast<PixelFormat>(frames[k].format)};
        strip::StripRecord r{};
        s->impl.run(n, views, config(c), &r, reinterpret_cast<HomographyResult*>(pairs), keypoints);
        rec->status = static_cast<int>(r.status);
        std::snprintf(rec->fail_stage, sizeof rec->fail_stage, "%s", r.fail_stage);
        rec->fail_pair = r.fail_pair;
        rec->canvas_width = r.canvas_width;
        rec->canvas_height = r.canvas_height;
        rec->origin_x = r.origin_x;
        rec->origin_y = r.origin_y;
        rec->union_pixels = r.union_pixels;
        rec->overlap_pixels = r.overlap_pixels;
        rec->t_pairs = r.t_pairs;
        rec->t_geometry = r.t_geometry;
        rec->t_render = r.t_render;
        rec->t_total = r.t_total;
        return rec->status;
    } catch (...) {
        return map_exception();
    }
}

int cs_strip_outputs(const cs_strip* s, cs_image* canvas, cs_image* mask, cs_geometry* geometry, int n) {
    if (!s) return CS_INVALID_ARGUMENT;
    const auto put = [](const ImageView& v, cs_image* o) {
        if (o) *o = {v.data, v.width, v.height, v.pitch_bytes, static_cast<int>(v.format)};
    };
    put(s->impl.canvas(), canvas);
    put(s->impl.mask(), mask);
    const auto& g = s->impl.geometry();
    if (geometry) {
        if (n != static_cast<int>(g.size())) return CS_INVALID_ARGUMENT;
        std::memcpy(geometry, g.data(), g.size() * sizeof(cs_geometry));
    }
    return CS_OK;
}

void* cs_strip_stream(cs_strip* s) { return s ? s->impl.stream() : nullptr; }

}  // extern "C"
