// extern "C" layer for the batched and served forms.
#include <cstring>
#include <new>
#include <stdexcept>

#include "cs_stitch.h"
#include "stitch/batch.hpp"
#include "stitch/server.hpp"

using namespace stitch;

struct cs_batch {
    batch::BatchRunner impl;
    std::vector<batch::PairOutcome> last;
    std::vector<cs_pair_spec> specs, previous;  // of the last two submitted batches
};

struct cs_server {
    serve::Server impl;
};

static_assert(sizeof(cs_homography_result) == sizeof(HomographyResult), "HomographyResult layout");
static_assert(sizeof(cs_render_statistics) == sizeof(s6::RenderStatistics), "RenderStatistics layout");
static_assert(sizeof(cs_scale_info) == sizeof(ScaleInfo), "ScaleInfo layout");

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

ImageView view(const cs_image& v) {
    return {v.data, v.width, v.height, v.pitch_bytes, static_cast<PixelFormat>(v.format)};
}

cs_image put_image(const ImageView& v) {
    return {v.data, v.width, v.height, v.pitch_bytes, static_cast<int>(v.format)};
}

cs_geometry put_geo(const s6::Geometry& g) {
    cs_geometry o{};
    o.width = g.width;
    o.height = g.height;
    o.origin_x = g.origin_x;
    o.origin_y = g.origin_y;
    o.width_a = g.width_a;
    o.height_a = g.height_a;
    o.width_b = g.width_b;
    o.height_b = g.height_b;
    o.denominator_tolerance = g.denominator_tolerance;
    o.snap_tolerance = g.snap_tolerance;
    std::memcpy(o.inverse, g.inverse, sizeof o.inverse);
    return o;
}

batch::PairSpec spec(const cs_pair_spec& p) {
    batch::PairSpec s{};
    s.a = view(p.a);
    s.b = view(p.b);
    s.S = p.S;
    s.K = p.K;
    s.R = p.R;
    s.blend = static_cast<s6::BlendMode>(p.blend_mode);
    s.fast_threshold = p.fast_threshold;
    s.harris_k = p.harris_k;
    s.ratio = p.ratio_threshold;
    s.cross_check = p.cross_check != 0;
    s.msac_batch_size = p.msac_batch_size;
    s.msac_refinements = p.msac_refinements;
    s.msac_threshold = p.msac_threshold;
    s.seed = p.seed;
    return s;
}

void copy_text(char* dst, std::size_t n, const std::string& s) {
    std::memset(dst, 0, n);
    std::strncpy(dst, s.c_str(), n - 1);
}

cs_keypoints put_kp(const Keypoints& k) { return {k.x, k.y, k.response, k.capacity, k.d_count}; }
cs_descriptors put_desc(const Descriptors& d) {
    return {static_cast<int>(d.type), d.dim, d.row_stride_bytes, d.data};
}

}  // namespace

extern "C" {

int cs_batch_create(int max_pairs, uint64_t pool_bytes, int use_graphs, cs_batch** out) {
    if (!out || max_pairs <= 0) return CS_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        auto* b = new cs_batch{};
        try {
            b->impl.initialize(max_pairs, static_cast<std::size_t>(pool_bytes), use_graphs != 0);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

void cs_batch_destroy(cs_batch* b) { delete b; }

namespace {
void put_outcomes(const std::vector<batch::PairOutcome>& last, const cs_pair_spec* pairs, int n,
                  cs_pair_outcome* outcomes) {
        for (int i = 0; i < n; ++i) {
            const batch::PairOutcome& o = last[i];
            cs_pair_outcome& r = outcomes[i];
            std::memset(&r, 0, sizeof r);
            r.status = static_cast<int>(o.status);
            copy_text(r.fail_stage, sizeof r.fail_stage, o.fail_stage);
            r.num_kp_a = o.features.num_kp_a;
            r.num_kp_b = o.features.num_kp_b;
            std::memcpy(&r.scale_a, &o.features.sa, sizeof r.scale_a);
            std::memcpy(&r.scale_b, &o.features.sb, sizeof r.scale_b);
            std::memcpy(&r.homography, &o.features.h, sizeof r.homography);
            r.geometry = put_geo(o.geometry);
            std::memcpy(&r.statistics, &o.render.stats, sizeof r.statistics);
            if (o.render_buffers) {
                r.canvas = put_image(o.render_buffers->canvas);
                r.mask = put_image(o.render_buffers->mask);
            }
            if (o.buffers && o.buffers->block) {
                const batch::PairBuffers& B = *o.buffers;
                r.keypoints_a = put_kp(B.kp_a);
                r.keypoints_b = put_kp(B.kp_b);
                r.keypoints_a.capacity = r.keypoints_b.capacity = pairs[i].K;
                r.descriptors_a = put_desc(B.desc_a);
                r.descriptors_b = put_desc(B.desc_b);
                r.matches = {B.matches.idx_a, B.matches.idx_b, B.matches.distance, B.matches.capacity,
                             B.matches.d_count};
                r.inlier_mask = B.inlier_mask;
            }
        }
}
}  // namespace

int cs_batch_run(cs_batch* b, const cs_pair_spec* pairs, int n, cs_pair_outcome* outcomes, cs_batch_timings* t) {
    if (!b || (n > 0 && (!pairs || !outcomes)) || n < 0) return CS_INVALID_ARGUMENT;
    try {
        std::vector<batch::PairSpec> specs(n);
        for (int i = 0; i < n; ++i) specs[i] = spec(pairs[i]);
        batch::BatchTimings tm{};
        b->impl.run(specs.data(), n, &b->last, &tm);
        b->specs.assign(pairs, pairs + n);
        b->previous.clear();
        put_outcomes(b->last, pairs, n, outcomes);
        if (t) *t = {tm.feature_ms, tm.render_ms, tm.host_ms, tm.total_ms};
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_submit(cs_batch* b, const cs_pair_spec* pairs, int n, cs_pair_outcome* previous, int previous_n,
                    cs_batch_timings* t) {
    if (!b || (n > 0 && !pairs) || n < 0 || previous_n < 0 || (previous_n > 0 && !previous)) return CS_INVALID_ARGUMENT;
    try {
        const int have = b->impl.pending() ? static_cast<int>(b->specs.size()) : 0;
        if (previous && previous_n != have) return CS_INVALID_ARGUMENT;
        std::vector<batch::PairSpec> specs(n);
        for (int i = 0; i < n; ++i) specs[i] = spec(pairs[i]);
        batch::BatchTimings tm{};
        b->impl.submit(specs.data(), n, &b->last, &tm);
        b->previous.swap(b->specs);
        b->specs.assign(pairs, pairs + n);
        if (previous && have > 0) put_outcomes(b->last, b->previous.data(), have, previous);
        if (t) *t = {tm.feature_ms, tm.render_ms, tm.host_ms, tm.total_ms};
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_finish(cs_batch* b, cs_pair_outcome* outcomes, int n) {
    if (!b || n < 0 || (n > 0 && !outcomes)) return CS_INVALID_ARGUMENT;
    if (!b->impl.pending()) return n == 0 ? CS_OK : CS_INVALID_ARGUMENT;  // nothing submitted
    try {
        b->impl.finish(&b->last, nullptr);
        if (n != static_cast<int>(b->last.size()) || n != static_cast<int>(b->specs.size())) return CS_INVALID_ARGUMENT;
        put_outcomes(b->last, b->specs.data(), n, outcomes);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_autotune(cs_batch* b, const cs_pair_spec* pairs, int n, const char* cache_path, cs_tuning* chosen) {
    if (!b || !pairs || n <= 0) return CS_INVALID_ARGUMENT;
    try {
        std::vector<batch::PairSpec> specs(n);
        for (int i = 0; i < n; ++i) specs[i] = spec(pairs[i]);
        const batch::Tuning t = b->impl.autotune(specs.data(), n, cache_path ? cache_path : "");
        if (chosen) {
            *chosen = cs_tuning{};
            chosen->s4_segment = t.s4_segment;
            chosen->workgroups_per_cu = t.workgroups_per_cu;
            chosen->graphs = t.graphs ? 1 : 0;
            for (int s = 0; s < 6; ++s) chosen->stage_workgroups_per_cu[s] = t.stage_workgroups[s];
            chosen->s1_tile_rows = t.s1_tile_rows;
        }
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_apply(cs_batch* b, const cs_tuning* t) {
    if (!b || !t) return CS_INVALID_ARGUMENT;
    try {
        batch::Tuning tuning{};
        tuning.s4_segment = t->s4_segment;
        tuning.workgroups_per_cu = t->workgroups_per_cu;
        tuning.graphs = t->graphs != 0;
        for (int s = 0; s < 6; ++s) tuning.stage_workgroups[s] = t->stage_workgroups_per_cu[s];
        tuning.s1_tile_rows = t->s1_tile_rows ? t->s1_tile_rows : 64;
        b->impl.apply(tuning);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_pool(cs_batch* b, uint64_t* capacity, uint64_t* in_use, uint64_t* peak) {
    if (!b) return CS_INVALID_ARGUMENT;
    std::size_t c = 0, u = 0, p = 0;
    b->impl.pool_use(&c, &u, &p);
    if (capacity) *capacity = c;
    if (in_use) *in_use = u;
    if (peak) *peak = p;
    return CS_OK;
}

int cs_batch_pair_bytes(int width, int height, int K, int R, uint64_t* bytes) {
    if (!bytes || width <= 0 || height <= 0 || K <= 0 || R <= 0) return CS_INVALID_ARGUMENT;
    try {
        *bytes = batch::pair_bytes({width, height, K, R});
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_set_probe(cs_batch* b, int on) {
    if (!b) return CS_INVALID_ARGUMENT;
    try {
        b->impl.set_probe(on != 0);
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_batch_probe_read(cs_batch* b, cs_kernel_probe* out, int capacity, int* count) {
    if (!b || capacity < 0 || (capacity > 0 && !out)) return CS_INVALID_ARGUMENT;
    const auto& st = b->impl.probe_stats();
    if (count) *count = static_cast<int>(st.size());
    for (int i = 0; i < capacity && i < static_cast<int>(st.size()); ++i) {
        cs_kernel_probe& o = out[i];
        std::memset(&o, 0, sizeof o);
        copy_text(o.kernel, sizeof o.kernel, st[i].name);
        o.level = st[i].level;
        o.launches = st[i].launches;
        o.total_ms = st[i].ms;
        o.algorithmic_bytes = st[i].bytes;
    }
    return CS_OK;
}

int cs_batch_render_bytes(int canvas_width, int canvas_height, int mode, uint64_t* bytes) {
    if (!bytes || canvas_width <= 0 || canvas_height <= 0) return CS_INVALID_ARGUMENT;
    try {
        s6::Geometry g{};
        g.width = canvas_width;
        g.height = canvas_height;
        g.width_a = g.height_a = g.width_b = g.height_b = 1;  // sizes depend on the canvas only
        g.inverse[0] = g.inverse[4] = g.inverse[8] = 1.0;
        *bytes = batch::render_bytes(g, static_cast<s6::BlendMode>(mode));
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_server_create(const cs_server_config* c, cs_server** out) { return cs_server_create_policy(c, nullptr, out); }

int cs_server_create_policy(const cs_server_config* c, const char* policy_json, cs_server** out) {
    if (!c || !out || c->max_batch <= 0 || c->max_live <= 0 || !(c->canvas_factor > 0)) return CS_INVALID_ARGUMENT;
    if (policy_json) {
        try {
            (void)controller::parse_controller_policy(policy_json);
        } catch (...) {
            return CS_INVALID_ARGUMENT;
        }
    }
    *out = nullptr;
    try {
        auto* s = new cs_server{};
        serve::ServerConfig cfg{};
        cfg.hbm_budget = static_cast<std::size_t>(c->hbm_budget);
        cfg.max_batch = c->max_batch;
        cfg.max_live = c->max_live;
        cfg.adaptive = c->adaptive != 0;
        cfg.use_graphs = c->use_graphs != 0;
        cfg.canvas_factor = c->canvas_factor;
        if (policy_json) cfg.policy_json = policy_json;
        try {
            s->impl.initialize(cfg);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

void cs_server_destroy(cs_server* s) { delete s; }

int cs_server_run_trace(cs_server* s, const cs_request* requests, int n, cs_completion* completions,
                        cs_serve_stats* stats) {
    if (!s || n < 0 || (n > 0 && (!requests || !completions))) return CS_INVALID_ARGUMENT;
    try {
        std::vector<serve::Request> reqs(n);
        for (int i = 0; i < n; ++i) {
            const cs_request& q = requests[i];
            if (i > 0 && q.arrival_ms < requests[i - 1].arrival_ms) return CS_INVALID_ARGUMENT;
            reqs[i].id = q.id;
            reqs[i].a = view(q.a);
            reqs[i].b = view(q.b);
            reqs[i].arrival_ms = q.arrival_ms;
            reqs[i].seed = q.seed;
            reqs[i].fixed = spec(q.fixed);
            reqs[i].keep_output = q.keep_output != 0;
        }
        std::vector<serve::Completion> done;
        serve::ServeStats st{};
        s->impl.run_trace(reqs.data(), n, &done, &st);
        for (std::size_t i = 0; i < done.size() && static_cast<int>(i) < n; ++i) {
            const serve::Completion& c = done[i];
            cs_completion& o = completions[i];
            std::memset(&o, 0, sizeof o);
            o.id = c.id;
            copy_text(o.status, sizeof o.status, c.status);
            copy_text(o.fail_stage, sizeof o.fail_stage, c.fail_stage);
            o.arrival_ms = c.arrival_ms;
            o.admitted_ms = c.admitted_ms;
            o.done_ms = c.done_ms;
            o.attempts = c.attempts;
            o.S = c.knobs.S;
            o.K = c.knobs.K;
            o.R = c.knobs.R;
            o.blend_mode = static_cast<int>(c.knobs.B);
            o.num_kp_a = c.num_kp_a;
            o.num_kp_b = c.num_kp_b;
            std::memcpy(&o.homography, &c.h, sizeof o.homography);
            o.geometry = put_geo(c.geometry);
            std::memcpy(&o.statistics, &c.stats, sizeof o.statistics);
            o.canvas = put_image(c.canvas);
            o.mask = put_image(c.mask);
            o.reserved_bytes = c.reserved_bytes;
        }
        if (stats) {
            cs_serve_stats& o = *stats;
            o.completed = st.completed;
            o.failed = st.failed;
            o.makespan_ms = st.makespan_ms;
            o.p50_ms = st.p50_ms;
            o.p99_ms = st.p99_ms;
            o.mean_ms = st.mean_ms;
            o.feature_launches = st.feature_launches;
            o.render_launches = st.render_launches;
            o.mean_feature_batch = st.mean_feature_batch;
            o.mean_render_batch = st.mean_render_batch;
            o.peak_pool_bytes = st.peak_pool_bytes;
            o.feature_busy_ms = st.feature_busy_ms;
            o.render_busy_ms = st.render_busy_ms;
        }
        return CS_OK;
    } catch (...) {
        return map_exception();
    }
}

int cs_server_release(cs_server* s, uint64_t id) {
    if (!s) return CS_INVALID_ARGUMENT;
    s->impl.release(id);
    return CS_OK;
}

}  // extern "C"
