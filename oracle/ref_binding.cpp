// Binds host-compilable reference code (compiled from its own sources, where
// they lie in the reference checkout) to a flat C ABI, so tests can check the
// oracle and the build's host code against the reference bit for bit.
// Built by `make -C oracle ref`; the output stays in oracle/_ref/.
#include <cstdint>
#include <cstring>

#include "common/geometry.hpp"
#include "s5/s5_math.hpp"
#include "s6/geometry.hpp"
#include "controller/controller.hpp"
#include "controller/reuse.hpp"
#include "harness/controller_policy.hpp"

using namespace stitch;
namespace d5 = stitch::s5::detail;

namespace {
struct Array {
    const double* p;
    d5::Point operator()(int i) const { return {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }
};
struct Keep {
    const std::uint8_t* k;
    bool operator()(int i) const { return !k || k[i]; }
};
}  // namespace

extern "C" {

int ref_make_scale_info(int w, int h, double s, ScaleInfo* out) { return static_cast<int>(make_scale_info(w, h, s, out)); }

std::uint64_t ref_s5_random_word(std::uint64_t v) { return d5::random_word(v); }

void ref_s5_sample_four(int n, std::uint64_t seed, int hyp, int* out) { d5::sample_four(n, seed, hyp, out); }

int ref_s5_nondegenerate(const double* pts) {
    d5::Point p[4];
    for (int i = 0; i < 4; ++i) p[i] = Array{pts}(i);
    return d5::nondegenerate_four(p);
}

int ref_s5_fit(int n, const double* pts, const std::uint8_t* keep, double* out) {
    return d5::fit(n, Array{pts}, Keep{keep}, out);
}

double ref_s5_residual(const double* h, const double* p) { return d5::residual_squared(h, Array{p}(0)); }

int ref_s5_normalize(double* h) { return d5::normalize(h); }

int ref_s5_valid_domain(const double* h, int w, int hh) { return d5::valid_domain(h, w, hh); }

int ref_mat3_inverse(const double* m, double* out) { return mat3_inverse(m, out); }

// geometry[0..7] = width, height, origin_x, origin_y, wa, ha, wb, hb.
int ref_s6_prepare_geometry(int wa, int ha, int wb, int hb, const double* h, int* geometry, double* inverse) {
    s6::Geometry g{};
    const Status st = s6::prepare_geometry(wa, ha, wb, hb, h, s6::GeometryPolicy{}, &g);
    const int v[8] = {g.width, g.height, g.origin_x, g.origin_y, g.width_a, g.height_a, g.width_b, g.height_b};
    std::memcpy(geometry, v, sizeof v);
    std::memcpy(inverse, g.inverse, sizeof g.inverse);
    return static_cast<int>(st);
}

// plan_retry over knobs given as (S, K, R, B); out = restart_from, extend_s5,
// previous_hypotheses, change_generation.
int ref_plan_retry(double s0, int k0, int r0, int b0, double s1, int k1, int r1, int b1, int* out) {
    controller::RetryPlan plan{};
    const controller::Knobs a{s0, k0, r0, static_cast<s6::BlendMode>(b0)};
    const controller::Knobs b{s1, k1, r1, static_cast<s6::BlendMode>(b1)};
    const Status st = controller::plan_retry(a, b, &plan);
    out[0] = static_cast<int>(plan.restart_from);
    out[1] = plan.extend_s5;
    out[2] = plan.previous_hypotheses;
    out[3] = plan.change_generation;
    return static_cast<int>(st);
}

// Runs the built-in legacy-v1 policy over one pair: begin_pair with the probe
// gradients of images (wa, ha) and (wb, hb), then observe each feedback in
// turn. trace receives per step (action, reason, ladder S*1000, K, R); the
// first row is the begin_pair knobs (action -1). Returns the number of rows.
int ref_legacy_walk(int blind, double grad_a, double grad_b, int wa, int ha, int wb, int hb,
                    const HomographyResult* feedback, const int* exec_status, int n, double* trace) {
    const auto policy = blind ? harness::legacy_blind_s05_policy() : harness::legacy_controller_policy();
    controller::Controller c;
    if (policy.configure(c) != Status::OK) return -1;
    c.reset_pair();
    ScaleInfo pa{}, pb{};
    make_scale_info(wa, ha, c.probe_scale(), &pa);
    make_scale_info(wb, hb, c.probe_scale(), &pb);
    const controller::SceneSignals scene{{pa, grad_a}, {pb, grad_b}};
    controller::Knobs knobs{};
    if (c.begin_pair(scene, &knobs) != Status::OK) return -2;
    int rows = 0;
    auto put = [&](int action, int reason, const controller::Knobs& k) {
        double* t = trace + 5 * rows++;
        t[0] = action; t[1] = reason; t[2] = k.S; t[3] = k.K; t[4] = k.R;
    };
    put(-1, -1, knobs);
    for (int i = 0; i < n; ++i) {
        controller::Feedback fb{knobs, static_cast<Status>(exec_status[i]), feedback[i]};
        controller::Decision d{};
        if (c.observe(fb, &d) != Status::OK) { put(-3, -3, knobs); break; }
        put(static_cast<int>(d.action), static_cast<int>(d.reason), d.next);
        if (d.action != controller::Action::RETRY) break;
        knobs = d.next;
    }
    return rows;
}

// Parses a policy document with the reference's loader: 0 accepted, 1
// rejected (the reason in error).
int ref_policy_check(const char* text, char* error, int error_len) {
    try {
        (void)harness::parse_controller_policy(harness::JsonValue::parse(text));
        if (error_len > 0) error[0] = 0;
        return 0;
    } catch (const std::exception& e) {
        if (error_len > 0) {
            std::strncpy(error, e.what(), static_cast<std::size_t>(error_len) - 1);
            error[error_len - 1] = 0;
        }
        return 1;
    }
}

// ref_legacy_walk for a policy document, with observations: scene (32
// doubles or NULL) for begin_pair, gate row i (38 doubles) with feedback i
// when present[i]. Returns the number of rows, -1 when the policy is
// rejected, -2 when begin_pair fails.
int ref_policy_walk(const char* text, double grad_a, double grad_b, int wa, int ha, int wb, int hb, const double* scene,
                    const HomographyResult* feedback, const int* exec_status, const double* gate, const int* present,
                    int n, double* trace) {
    harness::ControllerPolicy policy;
    try {
        policy = harness::parse_controller_policy(harness::JsonValue::parse(text));
    } catch (const std::exception&) {
        return -1;
    }
    controller::Controller c;
    if (policy.configure(c) != Status::OK) return -1;
    c.reset_pair();
    ScaleInfo pa{}, pb{};
    make_scale_info(wa, ha, c.probe_scale(), &pa);
    make_scale_info(wb, hb, c.probe_scale(), &pb);
    controller::SceneSignals signals{{pa, grad_a}, {pb, grad_b}, {}, false};
    if (scene) {
        std::memcpy(signals.features.data(), scene, sizeof(double) * signals.features.size());
        signals.features_present = true;
    }
    controller::Knobs knobs{};
    if (c.begin_pair(signals, &knobs) != Status::OK) return -2;
    int rows = 0;
    auto put = [&](int action, int reason, const controller::Knobs& k) {
        double* t = trace + 5 * rows++;
        t[0] = action; t[1] = reason; t[2] = k.S; t[3] = k.K; t[4] = k.R;
    };
    put(-1, -1, knobs);
    for (int i = 0; i < n; ++i) {
        controller::Feedback fb{knobs, static_cast<Status>(exec_status[i]), feedback[i], {}, false};
        if (gate && present && present[i]) {
            std::memcpy(fb.features.data(), gate + static_cast<std::size_t>(i) * fb.features.size(),
                        sizeof(double) * fb.features.size());
            fb.features_present = true;
        }
        controller::Decision d{};
        if (c.observe(fb, &d) != Status::OK) { put(-3, -3, knobs); break; }
        put(static_cast<int>(d.action), static_cast<int>(d.reason), d.next);
        if (d.action != controller::Action::RETRY) break;
        knobs = d.next;
    }
    return rows;
}

}  // extern "C"
