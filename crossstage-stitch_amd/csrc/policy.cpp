// Controller policy files: the JSON schemas controller-policy-v2 (rule
// ladder), -v3 (selector and gate trees, optional retry targets) and -v4
// (pooled boosted models), with the reference's field rules and limits
// (harness/controller_policy.hpp:60-206). A parsed policy is checked by
// configuring a Controller with it.
#include <cmath>
#include <initializer_list>
#include <stdexcept>
#include <string>

#include "stitch/controller.hpp"
#include "stitch/json.hpp"

namespace stitch::controller {
namespace {

using json::Json;

[[noreturn]] void reject(const std::string& why) { throw std::runtime_error(why); }

const Json& member(const Json& parent, const char* name, Json::Kind kind) {
    const Json* v = parent.kind == Json::Object ? parent.find(name) : nullptr;
    if (!v || v->kind != kind) reject(std::string("Invalid controller policy field: ") + name);
    return *v;
}

bool has(const Json& parent, const char* name) { return parent.kind == Json::Object && parent.find(name); }

double finite_number(const Json& parent, const char* name) {
    const double v = member(parent, name, Json::Number).num;
    if (!std::isfinite(v)) reject(std::string("Nonfinite policy value: ") + name);
    return v;
}

// A whole number in [0, maximum].
int bounded_index(const Json& parent, const char* name, int maximum) {
    const double v = finite_number(parent, name);
    if (v < 0 || v > maximum || v != std::floor(v)) reject(std::string("Invalid policy integer: ") + name);
    return static_cast<int>(v);
}

// Schema v2 admits only its own fields.
void only_fields(const Json& object, std::initializer_list<const char*> allowed) {
    if (object.kind != Json::Object) reject("Expected controller policy object");
    for (const auto& kv : object.obj) {
        bool ok = false;
        for (const char* name : allowed) ok = ok || kv.first == name;
        if (!ok) reject("Unknown controller policy field: " + kv.first);
    }
}

std::vector<TreeNode> bounded_tree(const Json& nodes) {
    if (nodes.kind != Json::Array || nodes.arr.empty() || nodes.arr.size() > 127) reject("Invalid bounded tree");
    std::vector<TreeNode> out;
    out.reserve(nodes.arr.size());
    for (const Json& n : nodes.arr) {
        TreeNode t{};
        if (has(n, "value")) {
            t.value = bounded_index(n, "value", 15);
        } else {
            t.feature = bounded_index(n, "feature", kSelectorFeatures - 1);
            t.threshold = finite_number(n, "threshold");
            t.left = bounded_index(n, "left", 126);
            t.right = bounded_index(n, "right", 126);
        }
        out.push_back(t);
    }
    return out;
}

// Trees are stored back to back; child indices in the file are relative to
// their tree's first node.
BoostedModel boosted(const Json& model) {
    BoostedModel out;
    out.features = bounded_index(model, "features", kPooledGateInputs);
    out.bias = static_cast<float>(finite_number(model, "bias"));
    const std::string& transform = member(model, "transform", Json::String).str;
    if (transform != "exp" && transform != "sigmoid") reject("Unknown ensemble transform");
    out.exponential = transform == "exp";
    const Json& calibration = member(model, "calibration", Json::Object);
    out.slope = finite_number(calibration, "slope");
    out.intercept = finite_number(calibration, "intercept");
    const auto& trees = member(model, "trees", Json::Array).arr;
    if (trees.empty() || trees.size() > 128) reject("Invalid ensemble tree count");
    for (const Json& tree : trees) {
        if (tree.kind != Json::Array || tree.arr.empty() || tree.arr.size() > 127) reject("Invalid ensemble tree size");
        const int first = static_cast<int>(out.nodes.size());
        out.roots.push_back(out.nodes.size());
        for (const Json& n : tree.arr) {
            BoostNode b{};
            if (has(n, "leaf")) {
                b.leaf = static_cast<float>(finite_number(n, "leaf"));
            } else {
                b.feature = bounded_index(n, "feature", kPooledGateInputs - 1);
                b.left = first + bounded_index(n, "left", 126);
                b.right = first + bounded_index(n, "right", 126);
                b.threshold = static_cast<float>(finite_number(n, "threshold"));
            }
            out.nodes.push_back(b);
        }
    }
    if (!valid_ensemble(out.view(), out.features, out.exponential)) reject("Malformed ensemble nodes");
    return out;
}

bool identifier(const std::string& id) {
    if (id.empty()) return false;
    for (const char c : id)
        if (!((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '-' || c == '_' ||
              c == '.'))
            return false;
    return true;
}

}  // namespace

ControllerPolicy parse_controller_policy(const std::string& json_text) {
    const Json root = json::JsonReader(json_text).document();
    const std::string& schema = member(root, "schema_version", Json::String).str;
    const bool v2 = schema == "controller-policy-v2", v3 = schema == "controller-policy-v3",
               v4 = schema == "controller-policy-v4";
    if (!v2 && !v3 && !v4) reject("Unsupported controller policy schema");
    if (v2)
        only_fields(root, {"schema_version", "policy_id", "probe_scale", "ladder", "self_checks", "scene_rules",
                           "provenance", "residual_unit"});
    ControllerPolicy policy;
    policy.id = member(root, "policy_id", Json::String).str;
    policy.probe_scale = finite_number(root, "probe_scale");
    const auto& ladder = member(root, "ladder", Json::Array).arr;
    if (ladder.empty() || ladder.size() > (v4 ? 64u : 16u) || !identifier(policy.id))
        reject("Policy exceeds attempt-trace capacity or has no identity");
    if (has(root, "residual_unit") && member(root, "residual_unit", Json::String).str != "s1_pixels")
        reject("max_residual_s1 must use s1_pixels");
    if (has(root, "provenance")) {
        const Json& provenance = member(root, "provenance", Json::Object);
        if (has(provenance, "seed")) {
            const double seed = finite_number(provenance, "seed");
            if (seed < 0 || seed > 9007199254740991.0 || seed != std::floor(seed))
                reject("Policy provenance.seed must be an exactly representable nonnegative integer");
            policy.seed = static_cast<std::uint64_t>(seed);
        }
    }
    for (const Json& rung : ladder) {
        if (v2) only_fields(rung, {"S", "K", "R", "blend_mode", "config_id"});
        if (has(rung, "config_id")) member(rung, "config_id", Json::String);
        if (member(rung, "blend_mode", Json::String).str != "feather")
            reject("Calibrated policy requires fixed feather blending");
        policy.ladder.push_back({finite_number(rung, "S"), bounded_index(rung, "K", 4000), bounded_index(rung, "R", 2048),
                                 s6::BlendMode::Feather});
    }
    for (const Json& check : member(root, "self_checks", Json::Array).arr) {
        if (v2) only_fields(check, {"min_inliers", "min_inlier_ratio", "min_inlier_coverage", "max_residual_s1"});
        policy.checks.push_back({bounded_index(check, "min_inliers", 4000), finite_number(check, "min_inlier_ratio"),
                                 finite_number(check, "min_inlier_coverage"), finite_number(check, "max_residual_s1")});
    }
    for (const Json& rule : member(root, "scene_rules", Json::Array).arr) {
        if (v2) only_fields(rule, {"min_pair_gradient", "first_index"});
        policy.rules.push_back({finite_number(rule, "min_pair_gradient"),
                                static_cast<std::size_t>(bounded_index(rule, "first_index", v4 ? 63 : 15))});
    }
    if (v3) {
        if (member(root, "feature_schema", Json::String).str != "controller-features-v1")
            reject("Unknown Controller feature layout");
        policy.selector = bounded_tree(member(root, "selector", Json::Array));
        for (const Json& gate : member(root, "gates", Json::Array).arr) policy.gate_nodes.push_back(bounded_tree(gate));
        if (has(root, "retry_targets"))
            for (const Json& t : member(root, "retry_targets", Json::Array).arr) {
                if (t.kind != Json::Number || !std::isfinite(t.num) || t.num < 0 ||
                    t.num > static_cast<double>(ladder.size()) || t.num != std::floor(t.num))
                    reject("Invalid bounded retry target");
                policy.retry_targets.push_back(static_cast<std::size_t>(t.num));
            }
    }
    if (v4) {
        if (member(root, "feature_schema", Json::String).str != "controller-features-v2")
            reject("Unknown pooled feature layout");
        const Json& models = member(root, "models", Json::Object);
        policy.pooled_mode = true;
        policy.pre_quality = boosted(member(models, "pre_quality", Json::Object));
        policy.post_quality = boosted(member(models, "post_quality", Json::Object));
        policy.latency = boosted(member(models, "latency", Json::Object));
        PooledModels& m = policy.pooled;
        m.pre_threshold = finite_number(root, "pre_threshold");
        m.gate_threshold = finite_number(root, "gate_threshold");
        m.extra_ms = finite_number(root, "extra_ms");
        m.reference_index = static_cast<std::size_t>(bounded_index(root, "reference_index", 63));
        m.pair_relations = member(root, "pair_relations", Json::Bool).b;
        m.relation_mask = static_cast<unsigned>(bounded_index(root, "relation_mask", 7));
        if (m.reference_index >= policy.ladder.size()) reject("Invalid pooled fallback");
        const Knobs& ref = policy.ladder[m.reference_index];
        if (ref.S != 1. || ref.K != 2000 || ref.R != 1024) reject("Pooled reference differs from QoR baseline");
    }
    Controller check;
    if (policy.configure(check) != Status::OK) reject("Invalid Controller policy or ladder");
    return policy;
}

ControllerPolicy load_controller_policy(const std::string& name) {
    if (name == "legacy-v1") return legacy_controller_policy();
    if (name == "legacy-v1-blind-s05") return legacy_blind_s05_policy();
    std::ifstream f(name, std::ios::binary);
    if (!f) reject("cannot open " + name);
    std::stringstream text;
    text << f.rdbuf();
    return parse_controller_policy(text.str());
}

}  // namespace stitch::controller
