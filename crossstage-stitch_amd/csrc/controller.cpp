// Rule-based adaptive controller: see stitch/controller.hpp for the contract.
#include "stitch/controller.hpp"

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
    for (std::size_t i = 0; i < p.self_check_count; ++i) {
        const auto& c = p.self_checks[i];
        if (c.min_inliers < 4 || !unit_interval(c.min_inlier_ratio) || !unit_interval(c.min_inlier_coverage) ||
            !std::isfinite(c.max_residual_s1) || c.max_residual_s1 < 0.0)
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
        if (result.reason == Reason::SELF_CHECK_PASSED) {
            result.action = Action::ACCEPT;
            phase_ = Phase::Finished;
        } else {
            std::size_t next = position_ + 1;
            // Changing R cannot add matches: skip rungs whose inlier minimum
            // already exceeds the available correspondences.
            while (next < count_ && ladder_[next].S == feedback.executed.S && ladder_[next].K == feedback.executed.K &&
                   feedback.homography.num_matches < policy_.self_checks[next].min_inliers)
                ++next;
            if (next < count_) {
                result.action = Action::RETRY;
                result.next = ladder_[next];
                position_ = next;
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
    const Policy p{probe_scale, rules.data(), rules.size(), checks.data(), checks.size()};
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
