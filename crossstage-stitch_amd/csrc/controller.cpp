// Rule-based adaptive controller: see stitch/controller.hpp for the contract.
#include "stitch/controller.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>

namespace stitch::controller {
namespace detail {
namespace {
bool unit_interval(double x) noexcept { return std::isfinite(x) && x >= 0.0 && x <= 1.0; }

bool valid_image_scene(const ImageSceneSignals& image, double probe) noexcept {
    const auto& s = image.scale;
    if (!unit_interval(image.mean_abs_gradient) || s.orig_w <= 0 || s.orig_h <= 0 || s.work_w < 2 || s.work_h < 2 ||
        s.work_w > s.orig_w || s.work_h > s.orig_h)
        return false;
    ScaleInfo expected{};
    if (make_scale_info(s.orig_w, s.orig_h, probe, &expected) != Status::OK || s.work_w != expected.work_w ||
        s.work_h != expected.work_h || s.sx != expected.sx || s.sy != expected.sy)
        return false;
    for (int i = 0; i < 9; ++i)
        if (s.D[i] != expected.D[i]) return false;
    return true;
}
}  // namespace

bool valid_policy(const Policy& p) noexcept {
    if (!std::isfinite(p.probe_scale) || p.probe_scale <= 0.0 || p.probe_scale > 1.0 || !p.scene_rules ||
        p.scene_rule_count == 0 || !p.self_checks || p.self_check_count == 0)
        return false;
    for (std::size_t i = 0; i < p.scene_rule_count; ++i) {
        const auto& r = p.scene_rules[i];
        if (!unit_interval(r.min_pair_gradient) || r.first_index >= p.self_check_count ||
            (i && r.min_pair_gradient >= p.scene_rules[i - 1].min_pair_gradient))
            return false;
    }
    if (p.scene_rules[p.scene_rule_count - 1].min_pair_gradient != 0.0) return false;
    if (const PooledModels* m = p.pooled) {
        // Pooled models replace the selector, gates and retry targets, fall
        // back to the last rung, and read 39 (77 after S5) inputs.
        if (p.selector.count || p.gates || p.gate_count || p.retry_target_count ||
            m->reference_index + 1 != p.self_check_count || p.self_check_count > 64 ||
            !valid_ensemble(m->pre_quality, kPooledContext, false) ||
            !valid_ensemble(m->post_quality, kPooledGateInputs, false) ||
            !valid_ensemble(m->latency, kPooledContext, true) || !unit_interval(m->pre_threshold) ||
            !unit_interval(m->gate_threshold) || !std::isfinite(m->extra_ms) || m->extra_ms < 0 ||
            m->relation_mask > 7)
            return false;
    }
    for (std::size_t i = 0; i < p.self_check_count; ++i) {
        const auto& c = p.self_checks[i];
        if (c.min_inliers < 4 || !unit_interval(c.min_inlier_ratio) || !unit_interval(c.min_inlier_coverage) ||
            !std::isfinite(c.max_residual_s1) || c.max_residual_s1 < 0.0)
            return false;
    }
    if (p.selector.count) {
        // One gate tree per rung, each reading the first 24 gate features.
        if (!valid_tree(p.selector, kSelectorFeatures, static_cast<int>(p.self_check_count)) || !p.gates ||
            p.gate_count != p.self_check_count)
            return false;
        for (std::size_t i = 0; i < p.gate_count; ++i)
            if (!valid_tree(p.gates[i], 24, 2)) return false;
    } else if (p.gates || p.gate_count) {
        return false;
    }
    if (p.retry_target_count) {
        // Retry targets move strictly forward (the count itself ends the run).
        if (!p.selector.count || !p.retry_targets || p.retry_target_count != p.self_check_count) return false;
        for (std::size_t i = 0; i < p.retry_target_count; ++i)
            if (p.retry_targets[i] <= i || p.retry_targets[i] > p.self_check_count) return false;
    } else if (p.retry_targets) {
        return false;
    }
    return true;
}

bool valid_knobs(const Knobs& k) noexcept {
    return std::isfinite(k.S) && k.S > 0.0 && k.S <= 1.0 && k.K > 0 && k.R > 0 && s6::requested_reductions(k.B) >= 0;
}

bool same_knobs(const Knobs& a, const Knobs& b) noexcept { return a.S == b.S && a.K == b.K && a.R == b.R && a.B == b.B; }

bool valid_scene(const SceneSignals& s, double probe_scale) noexcept {
    return valid_image_scene(s.a, probe_scale) && valid_image_scene(s.b, probe_scale);
}

bool valid_feedback(const Feedback& f) noexcept {
    switch (f.execution_status) {
        case Status::INVALID_ARGUMENT:
        case Status::INSUFFICIENT_WORKSPACE:
        case Status::CUDA_ERROR: return true;
        case Status::OK: break;
        default: return false;
    }
    const auto& h = f.homography;
    switch (h.status) {
        case Status::NOT_ENOUGH_POINTS: return h.num_matches >= 0 && h.num_matches < 4;
        case Status::DEGENERATE: return h.num_matches >= 4;
        case Status::INVALID_ARGUMENT: return true;
        case Status::OK: break;
        default: return false;
    }
    if (h.num_inliers < 4 || h.num_matches < h.num_inliers || !unit_interval(h.inlier_ratio) ||
        !unit_interval(h.inlier_coverage) || !std::isfinite(h.reproj_residual) || h.reproj_residual < 0.0)
        return false;
    const double ratio = static_cast<double>(h.num_inliers) / h.num_matches;
    if (std::fabs(h.inlier_ratio - ratio) > 32 * std::numeric_limits<double>::epsilon()) return false;
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(h.H[i]) || !std::isfinite(h.H_s[i])) return false;
    return true;
}

Reason self_check_reason(const HomographyResult& h, const SelfCheck& c) noexcept {
    if (h.status == Status::NOT_ENOUGH_POINTS) return Reason::TOO_FEW_MATCHES;
    if (h.status == Status::DEGENERATE) return Reason::DEGENERATE_MODEL;
    if (h.num_inliers < c.min_inliers) return Reason::TOO_FEW_INLIERS;
    if (h.inlier_ratio < c.min_inlier_ratio) return Reason::LOW_INLIER_RATIO;
    if (h.inlier_coverage < c.min_inlier_coverage) return Reason::LOW_INLIER_COVERAGE;
    if (h.reproj_residual > c.max_residual_s1) return Reason::HIGH_RESIDUAL;
    return Reason::SELF_CHECK_PASSED;
}
}  // namespace detail

bool valid_tree(TreeView tree, int features, int outputs) noexcept {
    if (!tree.nodes || tree.count == 0 || tree.count > 127) return false;
    const int end = static_cast<int>(tree.count);
    for (int i = 0; i < end; ++i) {
        const TreeNode& n = tree.nodes[i];
        const bool leaf = n.feature == -1;
        const bool ok = leaf ? (n.value >= 0 && n.value < outputs && n.left == -1 && n.right == -1)
                             : (n.feature >= 0 && n.feature < features && std::isfinite(n.threshold) && n.left > i &&
                                n.right > i && n.left < end && n.right < end && n.value == -1);
        if (!ok) return false;
    }
    return true;
}

int predict_tree(TreeView tree, const double* values, int features) noexcept {
    for (int f = 0; f < features; ++f)
        if (!std::isfinite(values[f])) return -1;
    int at = 0;
    for (std::size_t step = 0; step < tree.count; ++step) {
        if (at < 0 || at >= static_cast<int>(tree.count)) return -1;
        const TreeNode& n = tree.nodes[at];
        if (n.feature == -1) return n.value;
        if (n.feature < 0 || n.feature >= features) return -1;
        at = values[n.feature] <= n.threshold ? n.left : n.right;
    }
    return -1;
}

bool valid_ensemble(const EnsembleView& e, int features, bool exponential) noexcept {
    if (!e.nodes || !e.roots || e.tree_count == 0 || e.tree_count > 128 || e.node_count == 0 ||
        e.node_count > e.tree_count * 127 || e.features != features || e.exponential != exponential ||
        !std::isfinite(e.bias) || !std::isfinite(e.slope) || e.slope < 0 || !std::isfinite(e.intercept) ||
        e.roots[0] != 0)
        return false;
    for (std::size_t t = 0; t < e.tree_count; ++t) {
        // Tree t owns nodes [roots[t], roots[t + 1]); children point forward inside it.
        const std::size_t begin = e.roots[t], end = t + 1 < e.tree_count ? e.roots[t + 1] : e.node_count;
        if (begin >= end || end > e.node_count || end - begin > 127) return false;
        for (std::size_t i = begin; i < end; ++i) {
            const BoostNode& n = e.nodes[i];
            const long long self = static_cast<long long>(i), stop = static_cast<long long>(end);
            const bool ok = n.feature == -1
                                ? (n.left == -1 && n.right == -1 && std::isfinite(n.leaf))
                                : (n.feature >= 0 && n.feature < features && std::isfinite(n.threshold) &&
                                   n.left > self && n.right > self && n.left < stop && n.right < stop);
            if (!ok) return false;
        }
    }
    return true;
}

double predict_ensemble(const EnsembleView& e, const double* x) noexcept {
    for (int f = 0; f < e.features; ++f)
        if (!std::isfinite(x[f]) || std::fabs(x[f]) > FLT_MAX) return std::numeric_limits<double>::quiet_NaN();
    float sum = e.bias;  // leaves accumulate in float32, tree by tree
    for (std::size_t t = 0; t < e.tree_count; ++t) {
        std::size_t at = e.roots[t];
        for (int depth = 0; depth < 127; ++depth) {
            const BoostNode& n = e.nodes[at];
            if (n.feature == -1) {
                sum += n.leaf;
                break;
            }
            at = static_cast<std::size_t>(static_cast<float>(x[n.feature]) < n.threshold ? n.left : n.right);
        }
    }
    if (e.exponential) return std::exp(std::clamp(static_cast<double>(sum), -20.0, 20.0));
    const double z = e.slope * sum + e.intercept;
    // Logistic, evaluated on the side where exp cannot overflow.
    return z >= 0 ? 1.0 / (1.0 + std::exp(-z)) : std::exp(z) / (1.0 + std::exp(z));
}

std::array<double, kSelectorFeatures> selector_features(const features::SceneFeatures& scene) noexcept {
    std::array<double, kSelectorFeatures> v{};
    std::copy(scene.begin(), scene.begin() + 26, v.begin());
    for (int q = 0; q < 4; ++q) v[26 + q] = std::fabs(scene[2 + 3 * q] - scene[14 + 3 * q]);
    return v;
}

// ------------------------------------------------------------- controller
// The per-pair decision procedure (contract: src/controller/controller.hpp;
// rules of src/controller/controller.cpp): the first rung comes from the
// scene, each attempt is judged by its rung's self-check (and, for learned
// policies, a quality model), and a failed attempt moves to a later rung
// until the ladder runs out.

void Controller::describe_rung(std::size_t rung) noexcept {
    const Knobs& k = ladder_[rung];
    context_[30] = k.S;
    context_[31] = k.K / 2000.;
    context_[32] = k.R / 1024.;
}

Status Controller::configure(const Policy& policy) noexcept {
    if (state_ == State::Running || !detail::valid_policy(policy)) return Status::INVALID_ARGUMENT;
    policy_ = policy;
    mode_ = policy.pooled ? Mode::Pooled : (policy.selector.count ? Mode::Selector : Mode::Rules);
    state_ = State::Unbound;
    ladder_ = nullptr;
    rungs_ = rung_ = 0;
    return Status::OK;
}

double Controller::probe_scale() const noexcept {
    return mode_ == Mode::Unconfigured ? std::numeric_limits<double>::quiet_NaN() : policy_.probe_scale;
}

Status Controller::initialize(const Knobs* ladder, std::size_t count) noexcept {
    if (state_ == State::Running || mode_ == Mode::Unconfigured || !ladder || count == 0 ||
        count != policy_.self_check_count)
        return Status::INVALID_ARGUMENT;
    // Every rung is a valid, distinct knob set with one blend mode.
    for (std::size_t i = 0; i < count; ++i) {
        if (!detail::valid_knobs(ladder[i]) || ladder[i].B != ladder[0].B) return Status::INVALID_ARGUMENT;
        if (std::any_of(ladder, ladder + i, [&](const Knobs& k) { return detail::same_knobs(k, ladder[i]); }))
            return Status::INVALID_ARGUMENT;
    }
    ladder_ = ladder;
    rungs_ = count;
    rung_ = 0;
    state_ = State::Bound;
    return Status::OK;
}

// The first rung of a pair, or false when the scene cannot choose one.
bool Controller::pick_first_rung(const SceneSignals& scene, std::size_t* rung) noexcept {
    if (mode_ == Mode::Rules) {
        // Rules hold decreasing gradient floors ending at zero: the first one
        // the weaker image reaches names the starting rung.
        const double g = std::fmin(scene.a.mean_abs_gradient, scene.b.mean_abs_gradient);
        const SceneRule* end = policy_.scene_rules + policy_.scene_rule_count;
        const SceneRule* hit =
            std::find_if(policy_.scene_rules, end, [g](const SceneRule& r) { return g >= r.min_pair_gradient; });
        if (hit == end) return false;
        *rung = hit->first_index;
        return true;
    }
    if (!scene.features_present) return false;
    const auto selector = selector_features(scene.features);
    if (mode_ == Mode::Selector) {
        const int r = predict_tree(policy_.selector, selector.data(), kSelectorFeatures);
        if (r < 0 || r >= static_cast<int>(rungs_)) return false;
        *rung = static_cast<std::size_t>(r);
        return true;
    }
    // Pooled: the reference rung unless another rung is predicted to pass
    // (quality at least the threshold) at a lower expected cost - its own
    // predicted latency plus, on failure, a retry at the reference rung and a
    // fixed overhead.
    const PooledModels& m = *policy_.pooled;
    std::copy(selector.begin(), selector.end(), context_.begin());
    std::copy(scene.features.begin() + 26, scene.features.end(), context_.begin() + 33);
    describe_rung(m.reference_index);
    const double reference_ms = predict_ensemble(m.latency, context_.data());
    if (!std::isfinite(reference_ms)) return false;
    std::size_t choice = m.reference_index;
    double lowest = reference_ms;
    for (std::size_t i = 0; m.pre_threshold < 1.0 && i < rungs_; ++i) {
        if (i == m.reference_index) continue;
        describe_rung(i);
        const double pass = predict_ensemble(m.pre_quality, context_.data());
        const double ms = predict_ensemble(m.latency, context_.data());
        if (!std::isfinite(pass) || !std::isfinite(ms)) return false;
        const double expected = ms + (1 - pass) * reference_ms + m.extra_ms;
        if (pass >= m.pre_threshold && expected < lowest) {
            lowest = expected;
            choice = i;
        }
    }
    describe_rung(choice);
    *rung = choice;
    return true;
}

Status Controller::begin_pair(const SceneSignals& scene, Knobs* first) noexcept {
    if (state_ != State::Bound || !first || !detail::valid_scene(scene, policy_.probe_scale))
        return Status::INVALID_ARGUMENT;
    std::size_t rung = 0;
    if (!pick_first_rung(scene, &rung)) return Status::INVALID_ARGUMENT;
    rung_ = rung;
    *first = ladder_[rung];
    state_ = State::Running;
    return Status::OK;
}

// Verdict on an attempt that ran: the rung's self-check, then the learned
// quality model where the policy has one.
Reason Controller::judge(const Feedback& f) const noexcept {
    const Reason checked = detail::self_check_reason(f.homography, policy_.self_checks[rung_]);
    if (checked != Reason::SELF_CHECK_PASSED) return checked;
    if (mode_ == Mode::Pooled && rung_ != policy_.pooled->reference_index) {
        // 77 inputs: rung context (33), gate features 0..23, pair relations
        // (6), alignment relations (gate features 24..37).
        if (!f.features_present) return Reason::QUALITY_PREDICTOR_REJECTED;
        std::array<double, kPooledGateInputs> in{};
        auto at = std::copy_n(context_.begin(), 33, in.begin());
        at = std::copy_n(f.features.begin(), 24, at);
        at = std::copy_n(context_.begin() + 33, 6, at);
        std::copy_n(f.features.begin() + 24, 14, at);
        const double q = predict_ensemble(policy_.pooled->post_quality, in.data());
        if (!(std::isfinite(q) && q >= policy_.pooled->gate_threshold)) return Reason::QUALITY_PREDICTOR_REJECTED;
    }
    if (mode_ == Mode::Selector &&
        (!f.features_present || predict_tree(policy_.gates[rung_], f.features.data(), features::kGateCount) != 1))
        return Reason::QUALITY_PREDICTOR_REJECTED;
    return Reason::SELF_CHECK_PASSED;
}

// Candidate after `rung` for a retry (rungs_ when there is none).
std::size_t Controller::successor(std::size_t rung) const noexcept {
    if (mode_ == Mode::Pooled) {
        // One fallback, to the reference rung; from there, nothing.
        const std::size_t ref = policy_.pooled->reference_index;
        return rung == ref ? rungs_ : ref;
    }
    return policy_.retry_target_count ? policy_.retry_targets[rung] : rung + 1;
}

Status Controller::observe(const Feedback& feedback, Decision* decision) noexcept {
    if (state_ != State::Running || !decision || !detail::same_knobs(feedback.executed, ladder_[rung_]) ||
        !detail::valid_feedback(feedback))
        return Status::INVALID_ARGUMENT;
    const HomographyResult& h = feedback.homography;
    if (feedback.execution_status != Status::OK || h.status == Status::INVALID_ARGUMENT) {
        *decision = {Action::REJECT, Reason::EXECUTION_ERROR, feedback.executed};
        state_ = State::Done;
        return Status::OK;
    }
    const Reason reason = judge(feedback);
    if (reason == Reason::SELF_CHECK_PASSED) {
        *decision = {Action::ACCEPT, reason, feedback.executed};
        state_ = State::Done;
        return Status::OK;
    }
    // A rung with the same S and K only draws more hypotheses over the same
    // matches, so it is passed over when it asks for more inliers than there
    // are matches.
    auto hopeless = [&](std::size_t r) {
        const Knobs& k = ladder_[r];
        return k.S == feedback.executed.S && k.K == feedback.executed.K &&
               h.num_matches < policy_.self_checks[r].min_inliers;
    };
    std::size_t next = successor(rung_);
    while (next < rungs_ && hopeless(next))
        next = mode_ == Mode::Pooled ? next + 1 : successor(next);
    if (next >= rungs_) {
        *decision = {Action::REJECT, Reason::BUDGET_EXHAUSTED, feedback.executed};
        state_ = State::Done;
        return Status::OK;
    }
    rung_ = next;
    if (mode_ == Mode::Pooled) describe_rung(next);
    *decision = {Action::RETRY, reason, ladder_[next]};
    return Status::OK;
}

void Controller::reset_pair() noexcept {
    rung_ = 0;
    state_ = ladder_ ? State::Bound : State::Unbound;
}

Status plan_retry(const Knobs& a, const Knobs& b, RetryPlan* plan) noexcept {
    if (!plan || !detail::valid_knobs(a) || !detail::valid_knobs(b) || a.B != b.B || detail::same_knobs(a, b))
        return Status::INVALID_ARGUMENT;
    RetryPlan r{RestartFrom::S5, false, 0, false};
    if (a.S != b.S) {
        r.restart_from = RestartFrom::S1;
        r.change_generation = true;
    } else if (a.K != b.K) {
        r.restart_from = RestartFrom::S2;
        r.change_generation = true;
    } else if (b.R > a.R) {
        r.extend_s5 = true;
        r.previous_hypotheses = a.R;
    }
    *plan = r;
    return Status::OK;
}

Status ControllerPolicy::configure(Controller& target) const {
    gate_views.clear();
    for (const auto& nodes : gate_nodes) gate_views.push_back({nodes.data(), nodes.size()});
    if (pooled_mode) {
        pooled_view = pooled;
        pooled_view.pre_quality = pre_quality.view();
        pooled_view.post_quality = post_quality.view();
        pooled_view.latency = latency.view();
    }
    const Policy p{probe_scale,
                   rules.data(),
                   rules.size(),
                   checks.data(),
                   checks.size(),
                   {selector.data(), selector.size()},
                   gate_views.empty() ? nullptr : gate_views.data(),
                   gate_views.size(),
                   retry_targets.empty() ? nullptr : retry_targets.data(),
                   retry_targets.size(),
                   pooled_mode ? &pooled_view : nullptr};
    const Status s = target.configure(p);
    return s == Status::OK ? target.initialize(ladder.data(), ladder.size()) : s;
}

ControllerPolicy legacy_controller_policy() {
    constexpr auto f = s6::BlendMode::Feather;
    return {"legacy-v1",
            .25,
            {{.25, 500, 128, f}, {.5, 1000, 128, f}, {.5, 1000, 256, f}, {.75, 2000, 512, f}, {1., 2000, 1024, f}},
            {{15, .15, .05, 4.}, {20, .15, .05, 3.5}, {25, .2, .05, 3.}, {30, .2, .05, 2.5}, {35, .2, .05, 2.}},
            {{.08, 0}, {.04, 1}, {0., 2}}};
}

ControllerPolicy legacy_blind_s05_policy() {
    ControllerPolicy p = legacy_controller_policy();
    p.id = "legacy-v1-blind-s05";
    p.rules = {{0., 1}};
    return p;
}

}  // namespace stitch::controller
