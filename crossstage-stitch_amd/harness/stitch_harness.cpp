// stitch_harness: the command-line campaign driver over the plain per-pair
// runner, with the reference's flags and output files (harness/main.cpp:120-483):
// fixed, oracle and adaptive modes; a JSON manifest of pairs or one pair
// given by --img-a / --img-b; a JSON list of configs or one config from -S /
// -K / -R / --blend; warmup and measured repeats; the runs CSV
// (harness/main.cpp:33-72), the per-attempt trace CSV
// (harness/attempt_trace.hpp:60-98) and the observations CSV (main.cpp:283-300).
// Images are binary PPM (P6, maxval 255), which is what the reference reads
// when it is built without OpenCV (harness/image_loader.hpp:25-95).
// Adaptive mode takes the built-in legacy-v1 / legacy-v1-blind-s05 ladders
// or a controller-policy-v2/v3/v4 JSON file (harness/controller_policy.hpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "stitch/controller.hpp"
#include "stitch/json.hpp"
#include "stitch/runner.hpp"

namespace {

using stitch::s6::BlendMode;
namespace H = stitch::harness;
namespace Ctl = stitch::controller;

using stitch::json::Json;
using stitch::json::read_json;

// ----------------------------------------------------------------- inputs
struct PairEntry {
    std::string pair_id, img_a, img_b;
};

// {"<pair_id>": {"a": path, "b": path, ...}, ...} (harness/json_parser.hpp:205-230);
// pairs run in key order, as the reference's map-backed reader yields them.
std::vector<PairEntry> manifest(const std::string& path) {
    const Json root = read_json(path);
    if (root.kind != Json::Object) throw std::runtime_error("manifest root must be an object");
    std::vector<PairEntry> out;
    for (const auto& [id, entry] : root.obj) {
        if (entry.kind != Json::Object) continue;
        const Json* a = entry.find("a");
        const Json* b = entry.find("b");
        if (!a || !b) throw std::runtime_error("manifest entry without 'a' or 'b': " + id);
        out.push_back({id, a->str, b->str});
    }
    std::stable_sort(out.begin(), out.end(), [](const PairEntry& x, const PairEntry& y) { return x.pair_id < y.pair_id; });
    return out;
}

BlendMode blend_of(const std::string& name) {
    if (name == "multiband3") return BlendMode::Multiband3;
    if (name == "multiband5") return BlendMode::Multiband5;
    if (name == "multiband6") return BlendMode::Multiband6;
    if (name == "multiband7") return BlendMode::Multiband7;
    return BlendMode::Feather;
}

const char* blend_name(BlendMode m) {
    switch (m) {
        case BlendMode::Multiband3: return "multiband3";
        case BlendMode::Multiband5: return "multiband5";
        case BlendMode::Multiband6: return "multiband6";
        case BlendMode::Multiband7: return "multiband7";
        default: return "feather";
    }
}

std::string config_id_of(double S, int K, int R, const std::string& blend) {
    std::ostringstream id;
    id << 'S' << std::fixed << std::setprecision(2) << S << "_K" << K << "_R" << R << "_B" << blend;
    return id.str();
}

struct Config {
    H::HarnessConfig cfg;
    std::string blend;
};

// [{"config_id", "S", "K", "R", "blend_mode"}, ...] (harness/json_parser.hpp:232-270).
std::vector<Config> configs(const std::string& path) {
    const Json root = read_json(path);
    if (root.kind != Json::Array) throw std::runtime_error("configs root must be an array");
    std::vector<Config> out;
    for (const Json& item : root.arr) {
        if (item.kind != Json::Object) continue;
        Config c;
        c.blend = "feather";
        if (const Json* v = item.find("S")) c.cfg.S = v->num;
        if (const Json* v = item.find("K")) c.cfg.K = static_cast<int>(v->num);
        if (const Json* v = item.find("R")) c.cfg.R = static_cast<int>(v->num);
        // Extension keys: multi-level detection.
        if (const Json* v = item.find("pyramid_levels")) c.cfg.pyramid_levels = std::max(1, static_cast<int>(v->num));
        if (const Json* v = item.find("pyramid_scale_factor")) c.cfg.pyramid_scale_factor = v->num;
        if (const Json* v = item.find("blend_mode")) c.blend = v->str;
        c.cfg.blend_mode = blend_of(c.blend);
        const Json* id = item.find("config_id");
        c.cfg.config_id = id && !id->str.empty() ? id->str : config_id_of(c.cfg.S, c.cfg.K, c.cfg.R, c.blend);
        out.push_back(c);
    }
    return out;
}

struct Rgb {
    int width{0}, height{0};
    std::vector<std::uint8_t> px;
};

// Binary PPM: "P6", width, height, maxval 255 (comments allowed between
// fields), one whitespace byte, then interleaved RGB rows.
Rgb load_ppm(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    const auto field = [&]() {
        std::string tok;
        while (true) {
            f >> std::ws;
            if (f.peek() == '#') {
                std::string skip;
                std::getline(f, skip);
                continue;
            }
            f >> tok;
            return tok;
        }
    };
    if (field() != "P6") throw std::runtime_error("not a binary PPM (P6): " + path);
    Rgb img;
    img.width = std::stoi(field());
    img.height = std::stoi(field());
    const int maxval = std::stoi(field());
    f.get();
    if (img.width <= 0 || img.height <= 0 || maxval != 255) throw std::runtime_error("unsupported PPM header: " + path);
    img.px.resize(static_cast<std::size_t>(img.width) * img.height * 3);
    f.read(reinterpret_cast<char*>(img.px.data()), static_cast<std::streamsize>(img.px.size()));
    if (!f) throw std::runtime_error("short PPM pixel data: " + path);
    return img;
}

std::string join(const std::string& dir, const std::string& rel) {
    if (dir.empty() || rel.empty() || rel.front() == '/') return rel.empty() ? dir : rel;
    return dir.back() == '/' ? dir + rel : dir + "/" + rel;
}

// ----------------------------------------------------------------- output
struct Labels {
    std::string run_id, git_rev, dataset, split, pair_id, policy_id;
    int repeat_idx{0}, is_warmup{0};
    double t_h2d{0};
};

void runs_header(std::ostream& o) {
    o << "run_id,git_rev,dataset,split,pair_id,config_id,repeat_idx,is_warmup,seed,S,K,R,blend_mode,"
         "t_h2d,t_s1,t_s2,t_s3,t_s4,t_s5,t_s6,t_controller,t_d2h,t_e2e,"
         "num_kp_a,num_kp_b,num_matches,num_inliers,inlier_ratio,inlier_coverage,reproj_residual,"
         "h0,h1,h2,h3,h4,h5,h6,h7,h8,"
         "canvas_pixels,overlap_pixels,blend_mode_effective,num_levels,peak_workspace_bytes,cache_hit,"
         "status,fail_stage,policy_id,attempt_count,t_probe_gpu,t_stream_wait,probe_gradient_a,probe_gradient_b,"
         "t_probe_s1,t_probe_grad,t_probe\n";
}

std::string fixed(double v, int digits) {
    std::ostringstream s;
    s << std::fixed << std::setprecision(digits) << v;
    return s.str();
}

std::string exact(double v) {
    std::ostringstream s;
    s << std::setprecision(std::numeric_limits<double>::max_digits10) << v;
    return s.str();
}

void runs_row(std::ostream& o, const Labels& L, const H::RunRecord& r, const std::string& blend) {
    const bool adaptive = !L.policy_id.empty();
    const int tdig = adaptive ? 6 : 3;
    o << L.run_id << ',' << L.git_rev << ',' << L.dataset << ',' << L.split << ',' << L.pair_id << ','
      << r.config_id << ',' << L.repeat_idx << ',' << L.is_warmup << ',' << r.seed << ','
      << (adaptive ? exact(r.S) : fixed(r.S, 4)) << ',' << r.K << ',' << r.R << ',' << blend << ','
      << fixed(L.t_h2d, tdig) << ',' << fixed(r.t_s1, tdig) << ',' << fixed(r.t_s2, tdig) << ','
      << fixed(r.t_s3, tdig) << ',' << fixed(r.t_s4, tdig) << ',' << fixed(r.t_s5, tdig) << ','
      << fixed(r.t_s6, tdig) << ',' << fixed(r.t_controller, tdig) << ',' << fixed(0.0, tdig) << ','
      << fixed(r.t_e2e, tdig) << ',' << r.num_kp_a << ',' << r.num_kp_b << ',' << r.num_matches << ','
      << r.num_inliers << ',' << fixed(r.inlier_ratio, 6) << ',' << fixed(r.inlier_coverage, 6) << ','
      << fixed(r.reproj_residual, 6) << ',';
    for (int i = 0; i < 9; ++i) {
        if (r.status == "OK" || !adaptive) o << exact(r.H[i]);
        o << ',';
    }
    o << r.canvas_pixels << ',' << r.overlap_pixels << ',' << blend << ',' << r.num_levels << ','
      << r.peak_workspace_bytes << ',' << r.cache_hit << ',' << r.status << ',' << r.fail_stage << ',';
    if (adaptive)
        o << L.policy_id << ',' << r.attempt_count << ',' << fixed(r.t_probe, 6) << ',' << fixed(r.t_stream_wait, 6);
    else
        o << ",,,";
    o << ',' << std::setprecision(17) << r.probe_gradient_a << ',' << r.probe_gradient_b << ','
      << fixed(r.t_probe, 4) << ',' << fixed(0.0, 4) << ',' << fixed(r.t_probe, 4) << '\n';
}

const char* action_name(Ctl::Action a) {
    return a == Ctl::Action::ACCEPT ? "ACCEPT" : a == Ctl::Action::RETRY ? "RETRY" : "REJECT";
}

const char* reason_name(Ctl::Reason r) {
    switch (r) {
        case Ctl::Reason::SELF_CHECK_PASSED: return "SELF_CHECK_PASSED";
        case Ctl::Reason::TOO_FEW_MATCHES: return "TOO_FEW_MATCHES";
        case Ctl::Reason::TOO_FEW_INLIERS: return "TOO_FEW_INLIERS";
        case Ctl::Reason::LOW_INLIER_RATIO: return "LOW_INLIER_RATIO";
        case Ctl::Reason::LOW_INLIER_COVERAGE: return "LOW_INLIER_COVERAGE";
        case Ctl::Reason::HIGH_RESIDUAL: return "HIGH_RESIDUAL";
        case Ctl::Reason::DEGENERATE_MODEL: return "DEGENERATE_MODEL";
        case Ctl::Reason::EXECUTION_ERROR: return "EXECUTION_ERROR";
        default: return "BUDGET_EXHAUSTED";
    }
}

void trace_header(std::ostream& o) {
    o << "run_id,git_rev,dataset,split,pair_id,repeat_idx,is_warmup,seed,attempt_idx,"
         "S,K,R,previous_R,eligible_R_old,generation,num_matches,num_inliers,inlier_ratio,"
         "inlier_coverage,reproj_residual,homography_status,action,reason,t_s4,t_s5,t_attempt,"
         "h0,h1,h2,h3,h4,h5,h6,h7,h8,policy_id,B,restart_from,t_s1,t_s2,t_s3,"
         "execution_status,attempt_fail_stage,terminal_status,terminal_fail_stage,t_features";
    for (int i = 0; i < stitch::features::kGateCount; ++i) o << ",gate_" << i;
    o << '\n';
}

void trace_rows(std::ostream& o, const Labels& L, const H::RunRecord& r, const H::AttemptTrace& t) {
    o << std::setprecision(17);
    for (int i = 0; i < t.count; ++i) {
        const H::AttemptRecord& a = t.records[i];
        const auto& h = a.homography;
        o << L.run_id << ',' << L.git_rev << ',' << L.dataset << ',' << L.split << ',' << L.pair_id << ','
          << L.repeat_idx << ',' << L.is_warmup << ',' << r.seed << ',' << i << ',' << a.knobs.S << ','
          << a.knobs.K << ',' << a.knobs.R << ',' << a.previous_R << ',' << a.previous_R << ',' << a.generation
          << ',';
        if (a.has_feedback)
            o << h.num_matches << ',' << h.num_inliers << ',' << h.inlier_ratio << ',' << h.inlier_coverage << ','
              << h.reproj_residual << ',' << static_cast<int>(h.status);
        else
            o << ",,,,,";
        o << ',';
        if (a.has_decision) o << action_name(a.decision.action) << ',' << reason_name(a.decision.reason);
        else o << ',';
        o << ',' << a.t_s4 << ',' << a.t_s5 << ',' << a.t_attempt;
        for (double v : h.H) {
            o << ',';
            if (a.has_homography) o << v;
        }
        o << ',' << L.policy_id << ',' << blend_name(a.knobs.B) << ',' << a.restart_from << ',' << a.t_s1 << ','
          << a.t_s2 << ',' << a.t_s3 << ',' << static_cast<int>(a.execution_status) << ',' << a.fail_stage << ','
          << r.status << ',' << r.fail_stage << ',' << a.t_features;
        for (double v : a.features) o << ',' << v;
        o << '\n';
    }
}

std::string default_run_id() {
    const std::time_t now = std::time(nullptr);
    char buf[32];
    std::strftime(buf, sizeof buf, "run_%Y%m%d_%H%M%S", std::localtime(&now));
    return buf;
}

void usage() {
    std::cout << "Usage: stitch_harness [OPTIONS]\n"
                 "Modes:\n"
                 "  --mode <fixed|oracle|adaptive>  Execution mode (default: fixed)\n"
                 "  --adaptive | --oracle           Shortcuts for --mode\n"
                 "  --controller-policy <path>      Required adaptive policy JSON, legacy-v1, or legacy-v1-blind-s05\n"
                 "  --trace-out <path>              Per-attempt controller trace CSV (adaptive)\n"
                 "  --features-out <path>           Controller observations CSV\n"
                 "  --no-reuse                      Rerun S4/S5 on R-only retries\n"
                 "  --no-probe                      Skip the S1 probe; start at the first rule rung\n"
                 "Batch options:\n"
                 "  --manifest <path>               Pairs JSON manifest\n"
                 "  --configs <path>                Configs JSON (fixed / oracle)\n"
                 "  --out <path>                    Output CSV (default: runs.csv)\n"
                 "  --dataset --split --run-id      Row labels\n"
                 "  --warmup <n> --repeats <n>      Unmeasured and measured runs per pair/config (2, 10)\n"
                 "  --seed <n>                      S5 seed\n"
                 "Single options:\n"
                 "  --img-a <ppm> --img-b <ppm>     One pair\n"
                 "  --pair-id <id>                  Its id (default: single_pair)\n"
                 "  -S <s> -K <k> -R <r>            Knobs (0.5, 1000, 256)\n"
                 "  --blend <mode>                  feather|multiband3|multiband5|multiband6|multiband7\n";
}

}  // namespace

int main(int argc, char** argv) {
    std::string manifest_path, configs_path, out_path = "runs.csv", dataset = "udis_d", split = "dev";
    std::string run_id = default_run_id(), mode = "fixed", trace_path, features_path, policy_name;
    std::string img_a, img_b, pair_id = "single_pair", blend = "feather";
    double S = 0.5;
    int K = 1000, R = 256, warmup = 2, repeats = 10;
    std::uint64_t seed = 0;
    bool seed_given = false;
    bool reuse = true, probe = true;
    const char* rev = std::getenv("STITCH_GIT_REV");
    const std::string git_rev = rev ? rev : "HEAD";
    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            const auto next = [&]() -> std::string {
                if (i + 1 >= argc) throw std::invalid_argument("missing value for " + a);
                return argv[++i];
            };
            if (a == "--manifest") manifest_path = next();
            else if (a == "--configs") configs_path = next();
            else if (a == "--out") out_path = next();
            else if (a == "--dataset") dataset = next();
            else if (a == "--split") split = next();
            else if (a == "--run-id") run_id = next();
            else if (a == "--mode") mode = next();
            else if (a == "--adaptive") mode = "adaptive";
            else if (a == "--oracle") mode = "oracle";
            else if (a == "--controller-policy") policy_name = next();
            else if (a == "--trace-out") trace_path = next();
            else if (a == "--features-out") features_path = next();
            else if (a == "--no-reuse") reuse = false;
            else if (a == "--no-probe") probe = false;
            else if (a == "--warmup") warmup = std::stoi(next());
            else if (a == "--repeats") repeats = std::stoi(next());
            else if (a == "--seed") {
                const std::string v = next();
                if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
                    throw std::invalid_argument("invalid --seed value");
                seed = std::stoull(v);
                seed_given = true;
            } else if (a == "--img-a") img_a = next();
            else if (a == "--img-b") img_b = next();
            else if (a == "--pair-id") pair_id = next();
            else if (a == "-S") S = std::stod(next());
            else if (a == "-K") K = std::stoi(next());
            else if (a == "-R") R = std::stoi(next());
            else if (a == "--blend") blend = next();
            else if (a == "--help" || a == "-h") {
                usage();
                return 0;
            } else {
                throw std::invalid_argument("unknown option " + a);
            }
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    if ((mode != "fixed" && mode != "oracle" && mode != "adaptive") ||
        (mode != "adaptive" && (!policy_name.empty() || !trace_path.empty())) || warmup < 0 || repeats <= 0) {
        std::cerr << "Invalid benchmark mode, policy option or repeat count\n";
        return 1;
    }
    std::error_code ec;
    const std::string policy_file = policy_name.rfind("legacy-v1", 0) == 0 ? std::string() : policy_name;
    for (const std::string& in : {manifest_path, configs_path, policy_file, img_a, img_b, trace_path, features_path})
        if (!in.empty() && std::filesystem::weakly_canonical(in, ec) == std::filesystem::weakly_canonical(out_path, ec)) {
            std::cerr << "Output path error: --out aliases another file\n";
            return 1;
        }
    if (!trace_path.empty() && std::filesystem::exists(trace_path)) {
        std::cerr << "Output path error: the attempt trace needs a fresh path\n";
        return 1;
    }

    std::vector<PairEntry> pairs;
    std::vector<Config> cfgs;
    try {
        if (!manifest_path.empty()) pairs = manifest(manifest_path);
        else if (!img_a.empty() && !img_b.empty()) pairs.push_back({pair_id, img_a, img_b});
        else throw std::runtime_error("must specify either --manifest or both --img-a and --img-b");
        if (mode != "adaptive") {
            if (!configs_path.empty()) {
                cfgs = configs(configs_path);
            } else {
                Config c;
                c.blend = blend;
                c.cfg.S = S;
                c.cfg.K = K;
                c.cfg.R = R;
                c.cfg.blend_mode = blend_of(blend);
                c.cfg.config_id = config_id_of(S, K, R, blend);
                cfgs.push_back(c);
            }
            for (Config& c : cfgs) c.cfg.seed = seed;
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }

    Ctl::Controller controller;
    Ctl::ControllerPolicy policy;
    if (mode == "adaptive") {
        // Built-in ladders by name, otherwise a policy file (main.cpp:293-314);
        // a policy's provenance seed replaces the default and must agree with --seed.
        try {
            if (policy_name.empty())
                throw std::runtime_error("Adaptive mode requires --controller-policy <path> or explicit legacy-v1");
            policy = Ctl::load_controller_policy(policy_name);
            if (policy.seed) {
                if (seed_given && seed != *policy.seed)
                    throw std::runtime_error("CLI seed differs from policy provenance.seed");
                seed = *policy.seed;
            }
            if (policy.configure(controller) != stitch::Status::OK)
                throw std::runtime_error("failed to initialize Controller");
        } catch (const std::exception& e) {
            std::cerr << "Controller policy error: " << e.what() << "\n";
            return 1;
        }
    }

    const std::string dir = manifest_path.empty() ? "" : std::filesystem::path(manifest_path).parent_path().string();
    H::PipelineRunner runner;
    runner.set_reuse_enabled(reuse);
    runner.set_probe_enabled(probe);
    // Observations are collected by the fixed and oracle modes only (main.cpp:322).
    runner.set_collect_features(!features_path.empty() && mode != "adaptive");
    try {
        int max_w = 2048, max_h = 2048;
        for (const PairEntry& p : pairs)
            for (const std::string& path : {p.img_a, p.img_b}) {
                try {
                    const Rgb img = load_ppm(join(dir, path));
                    max_w = std::max(max_w, img.width);
                    max_h = std::max(max_h, img.height);
                } catch (const std::exception&) {
                    // recorded per run below
                }
            }
        runner.initialize(max_w, max_h, 4000, 2048);
    } catch (const std::exception& e) {
        std::cerr << "Failed to initialize pipeline runner: " << e.what() << "\n";
        return 1;
    }

    std::ofstream runs(out_path);
    if (!runs) {
        std::cerr << "Error opening output CSV: " << out_path << "\n";
        return 1;
    }
    runs_header(runs);
    std::ofstream trace, features;
    if (!trace_path.empty()) {
        trace.open(trace_path);
        if (!trace) return 1;
        trace_header(trace);
    }
    if (!features_path.empty()) {
        features.open(features_path);
        if (!features) return 1;
        features << "pair_id,config_id,seed,repeat_idx,is_warmup,features_present,t_features";
        for (int i = 0; i < stitch::features::kSceneCount; ++i) features << ",scene_" << i;
        for (int i = 0; i < stitch::features::kGateCount; ++i) features << ",gate_" << i;
        features << '\n';
    }
    std::cout << "Starting benchmark harness:\n  Mode: " << mode << "\n  Pairs: " << pairs.size()
              << "\n  Configs: " << (mode == "adaptive" ? policy.ladder.size() : cfgs.size())
              << "\n  Policy: " << (mode == "adaptive" ? policy.id : "none") << "\n  Seed: " << seed
              << "\n  Warmup: " << warmup << ", Repeats: " << repeats << "\n  Output: " << out_path << "\n";

    std::uint64_t measured = 0;
    for (const PairEntry& pair : pairs) {
        std::string input_status, input_stage = "input";
        double t_h2d = 0;
        try {
            const Rgb a = load_ppm(join(dir, pair.img_a));
            const Rgb b = load_ppm(join(dir, pair.img_b));
            input_stage = "h2d";
            const auto t0 = std::chrono::steady_clock::now();
            runner.upload_pair(a.px.data(), a.width, a.height, b.px.data(), b.width, b.height);
            t_h2d = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        } catch (const stitch::admission::Error& e) {
            input_status = e.code;
        } catch (const std::exception& e) {
            std::cerr << "Input failed for " << pair.pair_id << ": " << e.what() << "\n";
            input_status = "ERR_INPUT";
        }
        const auto emit = [&](const Config* c) {
            for (int i = 0; i < warmup + repeats; ++i) {
                Labels L{run_id, git_rev, dataset, split, pair.pair_id, "", 0, 0, t_h2d};
                L.is_warmup = i < warmup;
                L.repeat_idx = L.is_warmup ? i : i - warmup;
                if (!c) L.policy_id = policy.id;
                H::RunRecord rec;
                std::string blend_out = c ? c->blend : "feather";
                if (!input_status.empty()) {
                    rec.seed = seed;
                    rec.config_id = c ? c->cfg.config_id : "adaptive_unstarted";
                    if (c) {
                        rec.S = c->cfg.S;
                        rec.K = c->cfg.K;
                        rec.R = c->cfg.R;
                    }
                    rec.status = input_status;
                    rec.fail_stage = input_stage;
                } else if (!c) {
                    try {
                        runner.execute_adaptive_run(controller, seed, &rec);
                    } catch (const std::exception& e) {
                        std::cerr << "Device failure stopped the adaptive campaign at " << pair.pair_id << ": "
                                  << e.what() << "\n";
                        return false;
                    }
                    const H::AttemptTrace& t = runner.attempt_trace();
                    if (t.count) {
                        const Ctl::Knobs& k = t.records[t.count - 1].knobs;
                        blend_out = blend_name(k.B);
                        std::ostringstream id;
                        id << 'S';
                        if (std::round(k.S * 100) / 100 == k.S) id << std::fixed << std::setprecision(2);
                        else id << std::setprecision(std::numeric_limits<double>::max_digits10);
                        id << k.S << "_K" << k.K << "_R" << k.R << "_B" << blend_out;
                        rec.config_id = id.str();
                    }
                    if (trace) trace_rows(trace, L, rec, t);
                } else {
                    try {
                        runner.execute_run(c->cfg, &rec);
                    } catch (const std::exception& e) {
                        std::cerr << "Device failure at " << pair.pair_id << ": " << e.what() << "\n";
                        return false;
                    }
                    rec.config_id = c->cfg.config_id;
                }
                runs_row(runs, L, rec, blend_out);
                if (features) {
                    features << std::setprecision(17) << pair.pair_id << ',' << rec.config_id << ',' << rec.seed
                             << ',' << L.repeat_idx << ',' << L.is_warmup << ',' << rec.features_present << ','
                             << rec.t_features;
                    for (double v : rec.scene_features) features << ',' << v;
                    for (double v : rec.gate_features) features << ',' << v;
                    features << '\n';
                }
                if (!L.is_warmup) ++measured;
            }
            runs.flush();
            if (trace) trace.flush();
            if (features) features.flush();
            return static_cast<bool>(runs) && (!features_path.size() || static_cast<bool>(features));
        };
        if (mode == "adaptive") {
            if (!emit(nullptr)) return 2;
        } else {
            for (const Config& c : cfgs)
                if (!emit(&c)) return 2;
        }
    }
    std::cout << "Benchmark harness completed successfully. Total measurement runs: " << measured << "\n";
    return 0;
}
