// Adaptive per-pair controller (rule-based policies) and retry planning.
// Contract and semantics of the reference's src/controller/controller.hpp:12-93,
// controller.cpp:8-153 (scene rule -> first rung; self-checks on S5 feedback ->
// ACCEPT / RETRY / REJECT), validation.cpp:29-88, policy.cpp:13-50,
// reuse.hpp:12-20 / reuse.cpp:5-22 (restart at S1, S2 or S5-extend) and the
// built-in ladders of harness/controller_policy.hpp:208-226. Learned
// selector/gate/pooled models are not part of this build: a policy carrying
// them is rejected by validation.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "stitch/stages.hpp"

namespace stitch::controller {

struct Knobs {
    double S;
    int K;
    int R;
    s6::BlendMode B;
};

struct ImageSceneSignals {
    ScaleInfo scale;
    double mean_abs_gradient;
};

struct SceneSignals {
    ImageSceneSignals a;
    ImageSceneSignals b;
};

struct Feedback {
    Knobs executed;
    Status execution_status;
    HomographyResult homography;
};

enum class Action : int { ACCEPT = 0, RETRY = 1, REJECT = 2 };

enum class Reason : int {
    SELF_CHECK_PASSED = 0,
    TOO_FEW_MATCHES,
    TOO_FEW_INLIERS,
    LOW_INLIER_RATIO,
    LOW_INLIER_COVERAGE,
    HIGH_RESIDUAL,
    DEGENERATE_MODEL,
    EXECUTION_ERROR,
    BUDGET_EXHAUSTED,
    QUALITY_PREDICTOR_REJECTED
};

struct Decision {
    Action action;
    Reason reason;
    Knobs next;
};

struct SelfCheck {
    int min_inliers;
    double min_inlier_ratio;
    double min_inlier_coverage;
    double max_residual_s1;
};

struct SceneRule {
    double min_pair_gradient;  // inclusive lower bound on min(G_A, G_B)
    std::size_t first_index;
};

struct Policy {
    double probe_scale;
    const SceneRule* scene_rules;  // strictly decreasing bounds ending at zero
    std::size_t scene_rule_count;
    const SelfCheck* self_checks;  // one per ladder entry
    std::size_t self_check_count;
};

class Controller {
public:
    Controller() noexcept = default;
    Status configure(const Policy& policy) noexcept;
    double probe_scale() const noexcept;
    Status initialize(const Knobs* ladder, std::size_t count) noexcept;
    Status begin_pair(const SceneSignals& scene, Knobs* first) noexcept;
    Status observe(const Feedback& feedback, Decision* decision) noexcept;
    void reset_pair() noexcept;
    std::size_t ladder_position() const noexcept { return position_; }

private:
    enum class Phase { Uninitialized, Idle, AwaitingFeedback, Finished };
    Policy policy_{};
    bool configured_ = false;
    const Knobs* ladder_ = nullptr;
    std::size_t count_ = 0;
    std::size_t position_ = 0;
    Phase phase_ = Phase::Uninitialized;
};

enum class RestartFrom : int { S1 = 1, S2 = 2, S5 = 5 };

struct RetryPlan {
    RestartFrom restart_from;
    bool extend_s5;
    int previous_hypotheses;
    bool change_generation;
};

Status plan_retry(const Knobs& executed, const Knobs& next, RetryPlan* plan) noexcept;

namespace detail {
bool valid_policy(const Policy& p) noexcept;
bool valid_knobs(const Knobs& k) noexcept;
bool same_knobs(const Knobs& a, const Knobs& b) noexcept;
bool valid_scene(const SceneSignals& s, double probe_scale) noexcept;
bool valid_feedback(const Feedback& f) noexcept;
Reason self_check_reason(const HomographyResult& h, const SelfCheck& c) noexcept;
}  // namespace detail

// Owning policy (ladder + checks + rules) that configures a Controller.
struct ControllerPolicy {
    std::string id;
    double probe_scale{};
    std::vector<Knobs> ladder;
    std::vector<SelfCheck> checks;
    std::vector<SceneRule> rules;
    Status configure(Controller& target) const;
};

ControllerPolicy legacy_controller_policy();
ControllerPolicy legacy_blind_s05_policy();

}  // namespace stitch::controller
