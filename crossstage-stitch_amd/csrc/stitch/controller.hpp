// Adaptive per-pair controller and retry planning.
// Contract and semantics of the reference's src/controller/controller.hpp:12-93,
// controller.cpp:8-153 (first rung from a scene rule, a selector tree or the
// pooled cost model; self-checks, gate trees or the pooled quality model on
// S5 feedback -> ACCEPT / RETRY / REJECT), validation.cpp:29-88,
// policy.cpp:13-50, tree.hpp:6-40 (bounded decision trees), ensemble.hpp:9-66
// (boosted ensembles with float32 splits), reuse.hpp:12-20 / reuse.cpp:5-22
// (restart at S1, S2 or S5-extend), the built-in ladders of
// harness/controller_policy.hpp:208-226 and its JSON policy schemas v2-v4
// (controller_policy.hpp:60-206).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "stitch/features.hpp"
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
    features::SceneFeatures features{};
    bool features_present{};
};

struct Feedback {
    Knobs executed;
    Status execution_status;
    HomographyResult homography;
    features::GateFeatures features{};
    bool features_present{};
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

// Bounded decision tree: node 0 is the root, children follow their parent,
// leaves carry a class value (tree.hpp:6-40).
struct TreeNode {
    int feature{-1};
    double threshold{};
    int left{-1}, right{-1}, value{-1};
};
struct TreeView {
    const TreeNode* nodes{};
    std::size_t count{};
};
bool valid_tree(TreeView tree, int features, int outputs) noexcept;
// Class of the leaf reached (value <= threshold goes left), -1 for
// non-finite inputs or a malformed walk.
int predict_tree(TreeView tree, const double* values, int features) noexcept;

// Boosted ensemble of up to 128 trees of up to 127 nodes, float32 splits
// and leaves summed in float32 (ensemble.hpp:9-66).
struct BoostNode {
    int feature{-1}, left{-1}, right{-1};
    float threshold{}, leaf{};
};
struct EnsembleView {
    const BoostNode* nodes{};
    std::size_t node_count{};
    const std::size_t* roots{};
    std::size_t tree_count{};
    int features{};
    float bias{};
    double slope{1.}, intercept{};
    bool exponential{};
};
bool valid_ensemble(const EnsembleView& e, int features, bool exponential) noexcept;
// exp(clamp(sum, -20, 20)) for latency models, the calibrated logistic
// otherwise; NaN when an input is not finite or exceeds the float range.
double predict_ensemble(const EnsembleView& e, const double* x) noexcept;

// Pooled models of controller-policy-v4: first rung by predicted cost,
// post-S5 gate by predicted quality, fallback to the reference rung.
struct PooledModels {
    EnsembleView pre_quality{}, post_quality{}, latency{};
    double pre_threshold{}, gate_threshold{}, extra_ms{};
    std::size_t reference_index{};
    bool pair_relations{};
    unsigned relation_mask{};
};

constexpr int kSelectorFeatures = 30;
constexpr int kPooledContext = 39;   // 30 selector features, S, K/2000, R/1024, 6 pair relations
constexpr int kPooledGateInputs = 77;
// Scene features 0..25 plus |mean_A - mean_B| per quadrant (features.hpp:16-21).
std::array<double, kSelectorFeatures> selector_features(const features::SceneFeatures& scene) noexcept;

struct Policy {
    double probe_scale;
    const SceneRule* scene_rules;  // strictly decreasing bounds ending at zero
    std::size_t scene_rule_count;
    const SelfCheck* self_checks;  // one per ladder entry
    std::size_t self_check_count;
    TreeView selector{};
    const TreeView* gates{};
    std::size_t gate_count{};
    const std::size_t* retry_targets{};
    std::size_t retry_target_count{};
    const PooledModels* pooled{};
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
    std::size_t ladder_position() const noexcept { return rung_; }
    // Which observations the configured policy reads (controller.hpp:72-77).
    bool uses_learned_features() const noexcept { return mode_ == Mode::Selector || mode_ == Mode::Pooled; }
    bool uses_pair_relations() const noexcept { return mode_ == Mode::Pooled && policy_.pooled->pair_relations; }
    unsigned relation_mask() const noexcept { return mode_ == Mode::Pooled ? policy_.pooled->relation_mask : 0; }
    bool needs_gate_features() const noexcept {
        return mode_ == Mode::Selector || (mode_ == Mode::Pooled && rung_ != policy_.pooled->reference_index);
    }

private:
    // How the first rung is chosen and an attempt judged: by the scene
    // gradient rules alone, by a selector tree with per-rung gate trees, or
    // by the pooled cost and quality models.
    enum class Mode { Unconfigured, Rules, Selector, Pooled };
    // Per pair: Bound -> (begin_pair) Running -> (observe: ACCEPT / REJECT)
    // Done; a RETRY stays Running on the next rung.
    enum class State { Unbound, Bound, Running, Done };
    Policy policy_{};
    Mode mode_ = Mode::Unconfigured;
    State state_ = State::Unbound;
    const Knobs* ladder_ = nullptr;
    std::size_t rungs_ = 0;
    std::size_t rung_ = 0;
    // Inputs of the pooled models: selector features (0..29), the rung's
    // S, K / 2000 and R / 1024 (30..32), the pair relations (33..38).
    std::array<double, kPooledContext> context_{};

    bool pick_first_rung(const SceneSignals& scene, std::size_t* rung) noexcept;
    Reason judge(const Feedback& feedback) const noexcept;
    std::size_t successor(std::size_t rung) const noexcept;
    void describe_rung(std::size_t rung) noexcept;
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

// Owning boosted model of a policy file.
struct BoostedModel {
    std::vector<BoostNode> nodes;
    std::vector<std::size_t> roots;
    int features{};
    float bias{};
    double slope{1.}, intercept{};
    bool exponential{};
    EnsembleView view() const {
        return {nodes.data(), nodes.size(), roots.data(), roots.size(), features, bias, slope, intercept, exponential};
    }
};

// Owning policy (ladder, checks, rules and any learned models) that
// configures a Controller; the Controller borrows its arrays, so the policy
// must outlive it (controller_policy.hpp:27-58).
struct ControllerPolicy {
    std::string id;
    double probe_scale{};
    std::vector<Knobs> ladder;
    std::vector<SelfCheck> checks;
    std::vector<SceneRule> rules;
    std::optional<std::uint64_t> seed;
    std::vector<TreeNode> selector;
    std::vector<std::vector<TreeNode>> gate_nodes;
    std::vector<std::size_t> retry_targets;
    bool pooled_mode{false};
    BoostedModel pre_quality, post_quality, latency;
    PooledModels pooled{};  // thresholds; the model views are set by configure
    // Views into the vectors above, rebuilt by every configure call.
    mutable std::vector<TreeView> gate_views{};
    mutable PooledModels pooled_view{};
    Status configure(Controller& target) const;
};

ControllerPolicy legacy_controller_policy();
ControllerPolicy legacy_blind_s05_policy();
// Parses a controller-policy-v2, -v3 or -v4 JSON document and checks it by
// configuring a Controller; throws std::runtime_error with the reason
// (controller_policy.hpp:123-206).
ControllerPolicy parse_controller_policy(const std::string& json_text);
// "legacy-v1", "legacy-v1-blind-s05" or the path of a policy file.
ControllerPolicy load_controller_policy(const std::string& name_or_path);

}  // namespace stitch::controller
