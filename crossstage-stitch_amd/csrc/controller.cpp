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

void Controller::set_pooled_rung(std::size_t index) noexcept {
    pooled_context_[30] = ladder_[index].S;
    pooled_context_[31] = ladder_[index].K / 2000.;
    pooled_context_[32] = ladder_[index].R / 1024.;
}

Status Controller::configure(const Policy& policy) noexcept {
    if (phase_ == Phase::AwaitingFeedback || !detail::valid_policy(policy)) return Status::INVALID_ARGUMENT;
    policy_ = policy;
    configured_ = true;
    ladder_ = nullptr;
    count_ = position_ = 0;
    phase_ = Phase::Uninitialized;
    return Status::OK;
}

double Controller::probe_scale() const noexcept {
    return configured_ ? policy_.probe_scale : std::numeric_limits<double>::quiet_NaN();
}

Status Controller::initialize(const Knobs* ladder, std::size_t count) noexcept {
    if (phase_ == Phase::AwaitingFeedback || !configured_ || !ladder || count == 0 || count != policy_.self_check_count)
        return Status::INVALID_ARGUMENT;
    for (std::size_t i = 0; i < count; ++i) {
        if (!detail::valid_knobs(ladder[i]) || ladder[i].B != ladder[0].B) return Status::INVALID_ARGUMENT;
        for (std::size_t j = 0; j < i; ++j)
            if (detail::same_knobs(ladder[i], ladder[j])) return Status::INVALID_ARGUMENT;
    }
    ladder_ = ladder;
    count_ = count;
    position_ = 0;
    phase_ = Phase::Idle;
    return Status::OK;
}

Status Controller::begin_pair(const SceneSignals& scene, Knobs* first) noexcept {
    if (phase_ != Phase::Idle || !first || !detail::valid_scene(scene, policy_.probe_scale))
        return Status::INVALID_ARGUMENT;
    if (const PooledModels* m = policy_.pooled) {
        // Start at the rung with the lowest expected cost: its predicted
        // latency plus, weighted by its predicted failure, the reference
        // rung's latency and a fixed overhead (controller.cpp:45-69).
        if (!scene.features_present) return Status::INVALID_ARGUMENT;
        const auto selector = selector_features(scene.features);
        std::copy(selector.begin(), selector.end(), pooled_context_.begin());
        std::copy(scene.features.begin() + 26, scene.features.end(), pooled_context_.begin() + 33);
        set_pooled_rung(m->reference_index);
        const double reference_ms = predict_ensemble(m->latency, pooled_context_.data());
        if (!std::isfinite(reference_ms)) return Status::INVALID_ARGUMENT;
        double best = reference_ms;
        position_ = m->reference_index;
        if (m->pre_threshold < 1.0)
            for (std::size_t i = 0; i < count_; ++i) {
                if (i == m->reference_index) continue;
                set_pooled_rung(i);
                const double quality = predict_ensemble(m->pre_quality, pooled_context_.data());
                const double ms = predict_ensemble(m->latency, pooled_context_.data());
                if (!std::isfinite(quality) || !std::isfinite(ms)) return Status::INVALID_ARGUMENT;
                const double cost = ms + (1 - quality) * reference_ms + m->extra_ms;
                if (quality >= m->pre_threshold && cost < best) {
                    best = cost;
                    position_ = i;
                }
            }
        set_pooled_rung(position_);
        *first = ladder_[position_];
        phase_ = Phase::AwaitingFeedback;
        return Status::OK;
    }
    if (policy_.selector.count) {
        if (!scene.features_present) return Status::INVALID_ARGUMENT;
        const auto selector = selector_features(scene.features);
        const int rung = predict_tree(policy_.selector, selector.data(), kSelectorFeatures);
        if (rung < 0 || rung >= static_cast<int>(count_)) return Status::INVALID_ARGUMENT;
        position_ = static_cast<std::size_t>(rung);
        *first = ladder_[position_];
        phase_ = Phase::AwaitingFeedback;
        return Status::OK;
    }
    const double gradient = std::fmin(scene.a.mean_abs_gradient, scene.b.mean_abs_gradient);
    for (std::size_t i = 0; i < policy_.scene_rule_count; ++i) {
        const auto& rule = policy_.scene_rules[i];
        if (gradient >= rule.min_pair_gradient) {
            position_ = rule.first_index;
            *first = ladder_[position_];
            phase_ = Phase::AwaitingFeedback;
            return Status::OK;
        }
    }
    return Status::INVALID_ARGUMENT;
}

Status Controller::observe(const Feedback& feedback, Decision* decision) noexcept {
    if (phase_ != Phase::AwaitingFeedback || !decision || !detail::same_knobs(feedback.executed, ladder_[position_]) ||
        !detail::valid_feedback(feedback))
        return Status::INVALID_ARGUMENT;
    Decision result{Action::REJECT, Reason::EXECUTION_ERROR, feedback.executed};
    if (feedback.execution_status != Status::OK || feedback.homography.status == Status::INVALID_ARGUMENT) {
        phase_ = Phase::Finished;
    } else {
        result.reason = detail::self_check_reason(feedback.homography, policy_.self_checks[position_]);
        const PooledModels* m = policy_.pooled;
        if (m && position_ != m->reference_index && result.reason == Reason::SELF_CHECK_PASSED) {
            // Quality gate on 77 inputs: the rung context (33), gate features
            // 0..23, the pair relations (6), the alignment relations (14).
            std::array<double, kPooledGateInputs> in{};
            std::copy_n(pooled_context_.begin(), 33, in.begin());
            std::copy_n(feedback.features.begin(), 24, in.begin() + 33);
            std::copy_n(pooled_context_.begin() + 33, 6, in.begin() + 57);
            std::copy_n(feedback.features.begin() + 24, 14, in.begin() + 63);
            const double quality = feedback.features_present ? predict_ensemble(m->post_quality, in.data()) : -1.0;
            if (!std::isfinite(quality) || quality < m->gate_threshold) result.reason = Reason::QUALITY_PREDICTOR_REJECTED;
        }
        if (policy_.selector.count && result.reason == Reason::SELF_CHECK_PASSED &&
            (!feedback.features_present ||
             predict_tree(policy_.gates[position_], feedback.features.data(), features::kGateCount) != 1))
            result.reason = Reason::QUALITY_PREDICTOR_REJECTED;
        if (result.reason == Reason::SELF_CHECK_PASSED) {
            result.action = Action::ACCEPT;
            phase_ = Phase::Finished;
        } else {
            // Pooled policies retry once, at the reference rung; selector
            // policies follow their retry targets; rule policies the next rung.
            const bool targets = policy_.retry_target_count != 0;
            auto after = [&](std::size_t i) { return targets ? policy_.retry_targets[i] : i + 1; };
            std::size_t next = m ? (position_ == m->reference_index ? count_ : m->reference_index) : after(position_);
            // Changing R cannot add matches: skip rungs whose inlier minimum
            // already exceeds the available correspondences.
            while (next < count_ && ladder_[next].S == feedback.executed.S && ladder_[next].K == feedback.executed.K &&
                   feedback.homography.num_matches < policy_.self_checks[next].min_inliers)
                next = after(next);
            if (next < count_) {
                result.action = Action::RETRY;
                result.next = ladder_[next];
                position_ = next;
                if (m) set_pooled_rung(next);
            } else {
                result.reason = Reason::BUDGET_EXHAUSTED;
                phase_ = Phase::Finished;
            }
        }
    }
    *decision = result;
    return Status::OK;
}

void Controller::reset_pair() noexcept {
    position_ = 0;
    phase_ = ladder_ ? Phase::Idle : Phase::Uninitialized;
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
