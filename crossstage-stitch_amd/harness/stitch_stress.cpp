// stitch_stress: the multi-resolution stress driver.
//
// Command line, files and status codes are those of the reference's stress
// driver (harness/stress_multires.cu, artifacts of harness/stress_io.hpp):
//   --track s6_only    S6 alone on a given homography (--h-file), with guard
//                      bytes behind every image row and the workspace
//                      checked afterwards, and regions of interest (--rois,
//                      up to 32 "x y w h" lines) dumped as raw RGB / mask
//                      bytes, optionally the whole canvas (--export-full);
//   --track end_to_end the plain per-pair runner, S1-S6.
// Each run writes native_metrics.csv (one row per warmup and repeat) and
// native_receipt.json into a new output directory; every file is written
// under a temporary name created exclusively and then linked into place, so
// nothing is overwritten and no reader sees a partial file. Argument errors
// exit with 2 before anything runs; run failures are recorded in the receipt.
// The device check compares the GPU architecture name (--expected-arch,
// default gfx950); --expected-sm is accepted and ignored.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <array>
#include <cerrno>
#include <chrono>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "stitch/runner.hpp"

namespace {

namespace fs = std::filesystem;
using namespace stitch;
using admission::Error;
using Clock = std::chrono::steady_clock;

struct HipFailure : std::runtime_error {
    hipError_t code;
    explicit HipFailure(hipError_t e) : std::runtime_error(hipGetErrorString(e)), code(e) {}
};
void ok(hipError_t e) {
    if (e != hipSuccess) throw HipFailure(e);
}

// ------------------------------------------------------------ command line
struct UsageError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

class Flags {
public:
    Flags(int argc, char** argv, const std::vector<std::string>& known, const std::vector<std::string>& switches) {
        auto listed = [](const std::vector<std::string>& v, const std::string& s) {
            for (const auto& x : v)
                if (x == s) return true;
            return false;
        };
        for (int i = 1; i < argc; ++i) {
            const std::string name = argv[i];
            std::string value = "1";
            if (!listed(switches, name)) {
                if (!listed(known, name)) throw UsageError("unknown argument " + name);
                if (i + 1 == argc) throw UsageError("missing value for " + name);
                value = argv[++i];
            }
            if (!values_.emplace(name, value).second) throw UsageError("duplicate argument " + name);
        }
    }
    bool has(const std::string& name) const { return values_.count(name) != 0; }
    const std::string& text(const std::string& name) const {
        const auto it = values_.find(name);
        if (it == values_.end()) throw UsageError("required argument missing: " + name);
        return it->second;
    }
    std::size_t count(const std::string& name, std::size_t fallback) const {
        if (!has(name)) return fallback;
        const std::string& v = text(name);
        if (v.empty() || v.size() > 19 || v.find_first_not_of("0123456789") != std::string::npos)
            throw UsageError(name + " needs a nonnegative integer");
        return std::stoull(v);
    }
    int extent(const std::string& name) const {
        const std::size_t v = count(name, 0);
        if (v < 1 || v > static_cast<std::size_t>(INT_MAX)) throw UsageError(name + " needs a positive dimension");
        return static_cast<int>(v);
    }

private:
    std::map<std::string, std::string> values_;
};

struct Options {
    bool render_only{true};  // s6_only; false: end_to_end
    std::string track, mode, a_path, b_path, h_path, roi_path, out_dir, arch{"gfx950"};
    int width{0}, height{0}, canvas_w{0}, canvas_h{0};
    int warmups{2}, repeats{10};
    std::size_t host_budget{512ull << 20}, device_reserve{256ull << 20};
    bool export_full{false};
    harness::HarnessConfig config{};
};

Options read_options(int argc, char** argv) {
    const Flags f(argc, argv,
                  {"--track", "--mode", "--a", "--b", "--h-file", "--rois", "--output", "--width", "--height",
                   "--canvas-width", "--canvas-height", "--warmups", "--repeats", "--expected-sm", "--expected-arch",
                   "--host-budget-bytes", "--device-reserve-bytes", "--scale", "--keypoints", "--ransac", "--seed"},
                  {"--export-full"});
    Options o;
    o.track = f.text("--track");
    if (o.track != "s6_only" && o.track != "end_to_end") throw UsageError("unknown track " + o.track);
    o.render_only = o.track == "s6_only";
    o.mode = f.text("--mode");
    const std::map<std::string, s6::BlendMode> blends{{"feather", s6::BlendMode::Feather},
                                                      {"multiband_3", s6::BlendMode::Multiband3},
                                                      {"multiband_5", s6::BlendMode::Multiband5}};
    if (!blends.count(o.mode)) throw UsageError("unknown blender " + o.mode);
    o.config.blend_mode = blends.at(o.mode);
    o.a_path = f.text("--a");
    o.b_path = f.text("--b");
    o.out_dir = f.text("--output");
    o.width = f.extent("--width");
    o.height = f.extent("--height");
    if (o.render_only) {
        o.h_path = f.text("--h-file");
        o.roi_path = f.text("--rois");
        o.export_full = f.has("--export-full");
    } else {
        // The end-to-end track estimates H itself: no ground truth, no dumps.
        if (f.has("--h-file") || f.has("--rois") || f.has("--export-full"))
            throw UsageError("end_to_end takes no --h-file, --rois or --export-full");
        o.canvas_w = f.extent("--canvas-width");
        o.canvas_h = f.extent("--canvas-height");
    }
    const std::size_t warm = f.count("--warmups", 2), reps = f.count("--repeats", 10);
    if (warm > 1000 || reps < 1 || reps > 1000) throw UsageError("warmups 0..1000, repeats 1..1000");
    o.warmups = static_cast<int>(warm);
    o.repeats = static_cast<int>(reps);
    f.count("--expected-sm", 0);
    if (f.has("--expected-arch")) o.arch = f.text("--expected-arch");
    o.host_budget = f.count("--host-budget-bytes", o.host_budget);
    o.device_reserve = f.count("--device-reserve-bytes", o.device_reserve);
    const std::size_t k = f.count("--keypoints", 1000), r = f.count("--ransac", 256);
    if (k < 1 || k > INT_MAX || r < 1 || r > INT_MAX) throw UsageError("--keypoints and --ransac must be positive");
    o.config.K = static_cast<int>(k);
    o.config.R = static_cast<int>(r);
    o.config.seed = f.count("--seed", 0);
    o.config.config_id = "stress_fixed";
    if (f.has("--scale")) {
        const std::string& s = f.text("--scale");
        char* end = nullptr;
        o.config.S = std::strtod(s.c_str(), &end);
        if (s.empty() || *end || !std::isfinite(o.config.S) || o.config.S <= 0 || o.config.S > 1)
            throw UsageError("--scale must lie in (0, 1]");
    }
    return o;
}

// ---------------------------------------------------------------- artifacts
// A fresh directory whose files appear whole or not at all.
class Artifacts {
public:
    explicit Artifacts(const std::string& dir) : dir_(dir) {
        if (fs::exists(dir_)) throw UsageError("output already exists");
        fs::create_directories(dir_);
    }
    void write(const std::string& name, const void* data, std::size_t bytes) const {
        const fs::path final_path = dir_ / name;
        const std::string staging = final_path.string() + ".partial";
        const int fd = ::open(staging.c_str(), O_WRONLY | O_CREAT | O_EXCL | O_CLOEXEC, 0644);
        if (fd < 0) throw std::runtime_error("cannot create " + staging);
        const auto* p = static_cast<const char*>(data);
        std::size_t left = bytes;
        bool good = true;
        while (left && good) {
            const ssize_t n = ::write(fd, p, left);
            if (n < 0 && errno == EINTR) continue;
            good = n > 0;
            if (good) {
                p += n;
                left -= static_cast<std::size_t>(n);
            }
        }
        good = (::close(fd) == 0) && good;
        // link() fails rather than replace an existing file.
        if (!good || ::link(staging.c_str(), final_path.c_str()) != 0) {
            ::unlink(staging.c_str());
            throw std::runtime_error("cannot publish " + final_path.string());
        }
        ::unlink(staging.c_str());
    }
    void write(const std::string& name, const std::string& text) const { write(name, text.data(), text.size()); }

private:
    fs::path dir_;
};

// Binary PPM (P6, maxval 255) of exactly width x height pixels and nothing after them.
std::vector<std::uint8_t> load_ppm(const std::string& path, int width, int height) {
    std::ifstream in(path, std::ios::binary);
    auto token = [&]() {
        std::string t;
        for (int c; (c = in.get()) != EOF;) {
            if (c == '#' && t.empty()) {
                in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
            } else if (std::isspace(c)) {
                if (!t.empty()) break;
            } else {
                t.push_back(static_cast<char>(c));
            }
        }
        return t;
    };
    if (!in || token() != "P6" || token() != std::to_string(width) || token() != std::to_string(height) ||
        token() != "255")
        throw Error("ERR_INPUT");
    const std::size_t bytes = admission::multiply(admission::multiply(width, height), 3);
    std::vector<std::uint8_t> px(bytes);
    in.read(reinterpret_cast<char*>(px.data()), static_cast<std::streamsize>(bytes));
    if (!in || in.peek() != EOF) throw Error("ERR_INPUT");
    return px;
}

// ------------------------------------------------------- device resources
constexpr std::uint8_t kGuard = 0xa5;
constexpr int kRowPad = 64;
constexpr std::size_t kTail = 256;

struct DeviceBlock {
    void* ptr{nullptr};
    explicit DeviceBlock(std::size_t bytes) {
        if (bytes) ok(hipMalloc(&ptr, bytes));
    }
    ~DeviceBlock() {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceBlock(const DeviceBlock&) = delete;
    DeviceBlock& operator=(const DeviceBlock&) = delete;
    std::uint8_t* bytes() const { return static_cast<std::uint8_t*>(ptr); }
};

// Pitched image whose 64 bytes behind each row hold a guard pattern.
class GuardedImage {
public:
    GuardedImage(int w, int h, int channels)
        : block_(admission::image_bytes(w, h, channels)), channels_(channels) {
        view_ = {block_.bytes(), w, h, w * channels + kRowPad, channels == 3 ? PixelFormat::RGB8 : PixelFormat::GRAY8};
        ok(hipMemset(block_.ptr, kGuard, admission::image_bytes(w, h, channels)));
    }
    const ImageView& view() const { return view_; }
    void upload(const std::vector<std::uint8_t>& rgb) {
        ok(hipMemcpy2D(view_.data, view_.pitch_bytes, rgb.data(), static_cast<std::size_t>(view_.width) * channels_,
                       static_cast<std::size_t>(view_.width) * channels_, view_.height, hipMemcpyHostToDevice));
    }
    bool guards_intact() const {
        std::vector<std::uint8_t> pads(static_cast<std::size_t>(kRowPad) * view_.height);
        ok(hipMemcpy2D(pads.data(), kRowPad, view_.data + static_cast<std::size_t>(view_.width) * channels_,
                       view_.pitch_bytes, kRowPad, view_.height, hipMemcpyDeviceToHost));
        for (std::uint8_t b : pads)
            if (b != kGuard) return false;
        return true;
    }
    // Raw bytes of the region (x, y, w, h), row after row.
    std::vector<std::uint8_t> region(const std::array<int, 4>& r) const {
        const auto [x, y, w, h] = r;
        if (x < 0 || y < 0 || w < 1 || h < 1 || w > view_.width - x || h > view_.height - y)
            throw Error("ERR_CANVAS_OVERFLOW");
        const std::size_t row = static_cast<std::size_t>(w) * channels_;
        std::vector<std::uint8_t> out(admission::multiply(row, h));
        ok(hipMemcpy2D(out.data(), row,
                       view_.data + static_cast<std::size_t>(y) * view_.pitch_bytes +
                           static_cast<std::size_t>(x) * channels_,
                       view_.pitch_bytes, row, h, hipMemcpyDeviceToHost));
        return out;
    }

private:
    DeviceBlock block_;
    int channels_;
    ImageView view_{};
};

// ------------------------------------------------------------ measurements
struct Timings {
    // Stage times (ms) or negative when absent: s1..s6, GPU span, host end to end.
    std::array<double, 8> v{-1, -1, -1, -1, -1, -1, -1, -1};
};

std::string num(double v) {
    std::ostringstream s;
    s << std::setprecision(17) << v;
    return s.str();
}

class MetricsTable {
public:
    MetricsTable() {
        text_ << "repeat_idx,is_warmup,status_code,t_s1,t_s2,t_s3,t_s4,t_s5,t_s6,t_gpu_steady,t_host_e2e,"
                 "inlier_count,inlier_ratio,H00,H01,H02,H10,H11,H12,H20,H21,H22,canvas_width,canvas_height,"
                 "origin_x,origin_y,workspace_bytes\n";
    }
    void add(int index, bool warmup, const std::string& status, const Timings& t, const harness::RunRecord* run,
             const s6::Geometry* canvas, std::size_t workspace) {
        std::vector<std::string> cells{std::to_string(index), warmup ? "1" : "0", status};
        for (double x : t.v) cells.push_back(x >= 0 && std::isfinite(x) ? num(x) : std::string());
        cells.push_back(run ? std::to_string(run->num_inliers) : std::string());
        cells.push_back(run ? num(run->inlier_ratio) : std::string());
        for (double h : run ? std::vector<double>(run->H, run->H + 9) : std::vector<double>(9))
            cells.push_back(run && status == "SUCCESS" ? num(h) : std::string());
        if (canvas) {
            for (int x : {canvas->width, canvas->height, canvas->origin_x, canvas->origin_y})
                cells.push_back(std::to_string(x));
            cells.push_back(std::to_string(workspace));
        } else {
            cells.resize(cells.size() + 5);
        }
        for (std::size_t i = 0; i < cells.size(); ++i) text_ << (i ? "," : "") << cells[i];
        text_ << '\n';
    }
    std::string str() const { return text_.str(); }

private:
    std::ostringstream text_;
};

// Ordered JSON object of scalars.
class Receipt {
public:
    Receipt& put(const std::string& key, const std::string& text) { return raw(key, quoted(text)); }
    Receipt& put(const std::string& key, const char* text) { return raw(key, quoted(text)); }
    template <class N>
    Receipt& put(const std::string& key, N number) {
        std::ostringstream s;
        s << std::setprecision(17) << number;
        return raw(key, s.str());
    }
    Receipt& raw(const std::string& key, const std::string& json) {
        fields_.emplace_back(key, json);
        return *this;
    }
    std::string str() const {
        std::string out = "{";
        for (std::size_t i = 0; i < fields_.size(); ++i)
            out += (i ? "," : "") + quoted(fields_[i].first) + ":" + fields_[i].second;
        return out + "}";
    }

private:
    static std::string quoted(const std::string& v) {
        std::string o = "\"";
        for (unsigned char c : v) {
            if (c == '"' || c == '\\') {
                o += '\\';
                o += static_cast<char>(c);
            } else if (c < 0x20) {
                char u[8];
                std::snprintf(u, sizeof u, "\\u%04x", c);
                o += u;
            } else {
                o += static_cast<char>(c);
            }
        }
        return o + "\"";
    }
    std::vector<std::pair<std::string, std::string>> fields_;
};

// What a run learned, for the receipt.
struct RunFacts {
    std::string status{"ERR_RUNNER"}, reason, gpu_arch;
    int runtime_version{0}, driver_version{0};
    s6::Geometry canvas{};
    std::size_t workspace{0}, workspace_reserved{0}, pipeline_capacity{0};
    std::size_t declared_device{0}, free_device{0}, total_device{0};
    bool workspace_queried{false};
};

void require_device(const Options& o, RunFacts* facts) {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) throw Error("ERR_NO_DEVICE");
    int device = 0;
    ok(hipGetDevice(&device));
    hipDeviceProp_t props{};
    ok(hipGetDeviceProperties(&props, device));
    facts->gpu_arch = props.gcnArchName;
    ok(hipRuntimeGetVersion(&facts->runtime_version));
    ok(hipDriverGetVersion(&facts->driver_version));
    if (facts->gpu_arch.compare(0, o.arch.size(), o.arch) != 0) throw Error("ERR_DEVICE_MISMATCH");
    ok(hipMemGetInfo(&facts->free_device, &facts->total_device));
    admission::fit(facts->declared_device, facts->free_device, o.device_reserve, "ERR_DEVICE_OOM");
}

std::array<double, 9> read_homography(const std::string& path) {
    std::ifstream in(path);
    std::array<double, 9> h{};
    for (double& v : h)
        if (!(in >> v) || !std::isfinite(v)) throw Error("ERR_INPUT");
    if (!(in >> std::ws).eof()) throw Error("ERR_INPUT");
    return h;
}

std::vector<std::array<int, 4>> read_rois(const std::string& path) {
    std::ifstream in(path);
    std::vector<std::array<int, 4>> boxes;
    std::array<int, 4> b{};
    while (in >> b[0] >> b[1] >> b[2] >> b[3]) {
        boxes.push_back(b);
        if (boxes.size() > 32) throw Error("ERR_CAPACITY");
    }
    if (!in.eof() || boxes.empty()) throw Error("ERR_INPUT");
    return boxes;
}

void render_track(const Options& o, const Artifacts& out, MetricsTable* table, RunFacts* facts) {
    using namespace admission;
    const auto h = read_homography(o.h_path);
    const Status g = s6::prepare_geometry(o.width, o.height, o.width, o.height, h.data(), {}, &facts->canvas);
    if (g != Status::OK) throw Error(g == Status::INSUFFICIENT_WORKSPACE ? "ERR_CANVAS_OVERFLOW" : "ERR_DEGENERATE_GEOM");
    const s6::Geometry& c = facts->canvas;
    if (s6::s6_workspace_bytes(c, o.config.blend_mode, &facts->workspace) != Status::OK) throw Error("ERR_CAPACITY");
    facts->workspace_queried = true;
    facts->workspace_reserved = facts->workspace ? add(facts->workspace, kTail) : 0;
    facts->pipeline_capacity = facts->workspace;
    facts->declared_device = add(multiply(2, image_bytes(o.width, o.height, 3)),
                                 add(add(image_bytes(c.width, c.height, 3), image_bytes(c.width, c.height, 1)),
                                     facts->workspace_reserved));
    if (o.export_full)
        fit(add(multiply(multiply(c.width, c.height), 3), multiply(multiply(o.width, o.height), 6)), o.host_budget,
            16ull << 20, "ERR_HOST_OOM");
    require_device(o, facts);
    const auto a = load_ppm(o.a_path, o.width, o.height), b = load_ppm(o.b_path, o.width, o.height);
    GuardedImage da(o.width, o.height, 3), db(o.width, o.height, 3), pano(c.width, c.height, 3),
        mask(c.width, c.height, 1);
    DeviceBlock work(facts->workspace_reserved);
    if (facts->workspace) ok(hipMemset(work.bytes() + facts->workspace, kGuard, kTail));
    da.upload(a);
    db.upload(b);
    hipStream_t stream{};
    ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    std::unique_ptr<std::remove_pointer_t<hipStream_t>, void (*)(hipStream_t)> stream_owner(
        stream, [](hipStream_t s) { (void)hipStreamDestroy(s); });
    hipEvent_t ev[2]{};
    for (auto& e : ev) ok(hipEventCreate(&e));
    for (int i = 0; i < o.warmups + o.repeats; ++i) {
        const auto t0 = Clock::now();
        ok(hipEventRecord(ev[0], stream));
        const Status r = s6::render(da.view(), db.view(), c, o.config.blend_mode, pano.view(), mask.view(), work.ptr,
                                    facts->workspace, stream);
        if (r != Status::OK) {
            for (auto e : ev) (void)hipEventDestroy(e);
            throw Error(r == Status::CUDA_ERROR ? "ERR_CUDA" : "ERR_CAPACITY");
        }
        ok(hipEventRecord(ev[1], stream));
        ok(hipEventSynchronize(ev[1]));
        float gpu_ms = 0;
        ok(hipEventElapsedTime(&gpu_ms, ev[0], ev[1]));
        Timings t;
        t.v[5] = t.v[6] = gpu_ms;
        t.v[7] = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        const bool warm = i < o.warmups;
        table->add(warm ? i : i - o.warmups, warm, "SUCCESS", t, nullptr, &c, facts->workspace);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    // Nothing wrote past a row or past the workspace.
    if (!da.guards_intact() || !db.guards_intact() || !pano.guards_intact() || !mask.guards_intact())
        throw Error("ERR_CORRECTNESS");
    if (facts->workspace) {
        std::array<std::uint8_t, kTail> tail{};
        ok(hipMemcpy(tail.data(), work.bytes() + facts->workspace, kTail, hipMemcpyDeviceToHost));
        for (std::uint8_t v : tail)
            if (v != kGuard) throw Error("ERR_CORRECTNESS");
    }
    const auto rois = read_rois(o.roi_path);
    for (std::size_t i = 0; i < rois.size(); ++i) {
        const auto rgb = pano.region(rois[i]), m = mask.region(rois[i]);
        out.write("roi" + std::to_string(i) + ".rgb", rgb.data(), rgb.size());
        out.write("roi" + std::to_string(i) + ".mask", m.data(), m.size());
    }
    if (o.export_full) {
        const std::array<int, 4> all{0, 0, c.width, c.height};
        const auto rgb = pano.region(all), m = mask.region(all);
        out.write("full.rgb", rgb.data(), rgb.size());
        out.write("full.mask", m.data(), m.size());
    }
}

void pipeline_track(const Options& o, MetricsTable* table, RunFacts* facts) {
    const auto plan = harness::PipelineRunner::query_memory(o.width, o.height, o.config.K, o.config.R, o.canvas_w,
                                                            o.canvas_h, o.config.blend_mode);
    facts->declared_device = plan.device_bytes;
    facts->workspace_reserved = plan.s6_workspace_bytes ? plan.s6_workspace_bytes + kTail : 0;
    facts->pipeline_capacity = plan.workspace_bytes;
    require_device(o, facts);
    const auto a = load_ppm(o.a_path, o.width, o.height), b = load_ppm(o.b_path, o.width, o.height);
    harness::PipelineRunner runner;
    runner.initialize(o.width, o.height, o.config.K, o.config.R, o.canvas_w, o.canvas_h, o.config.blend_mode,
                      o.device_reserve);
    runner.upload_pair(a.data(), o.width, o.height, b.data(), o.width, o.height);
    for (int i = 0; i < o.warmups + o.repeats; ++i) {
        harness::RunRecord run{};
        const auto t0 = Clock::now();
        runner.execute_run(o.config, &run);
        runner.synchronize();
        const double host_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        const bool stitched = run.status == "OK";
        if (stitched) {
            // The canvas of the estimated H, as S6 used it.
            if (s6::prepare_geometry(o.width, o.height, o.width, o.height, run.H, {}, &facts->canvas) != Status::OK)
                throw Error("ERR_DEGENERATE_GEOM");
            if (s6::s6_workspace_bytes(facts->canvas, o.config.blend_mode, &facts->workspace) != Status::OK)
                throw Error("ERR_CAPACITY");
            facts->workspace_queried = true;
        }
        Timings t;
        if (stitched) t.v = {run.t_s1, run.t_s2, run.t_s3, run.t_s4, run.t_s5, run.t_s6, 0, 0};
        t.v[6] = runner.last_fixed_gpu_span_ms();
        t.v[7] = host_ms;
        const bool warm = i < o.warmups;
        table->add(warm ? i : i - o.warmups, warm, stitched ? "SUCCESS" : run.status, t, &run,
                   stitched ? &facts->canvas : nullptr, facts->workspace);
    }
}

std::string receipt_of(const Options& o, const RunFacts& f) {
    const s6::Geometry& c = f.canvas;
    Receipt r;
    r.put("schema_version", "multires-native-v1")
        .put("status_code", f.status)
        .put("reason", f.reason)
        .put("track", o.track)
        .put("blend_mode", o.mode)
        .put("gpu_name", f.gpu_arch)
        .put("sm", 0)
        .put("gpu_arch", f.gpu_arch)
        .put("cuda_runtime_version", f.runtime_version)
        .put("cuda_driver_version", f.driver_version)
        .put("workspace_bytes", f.workspace)
        .put("workspace_reserved_bytes", f.workspace_reserved)
        .put("workspace_query_status", f.workspace_queried ? "QUERIED_ACTUAL_GEOMETRY" : "NOT_QUERIED_ACTUAL_GEOMETRY")
        .put("workspace_query_scope", "s6_actual_geometry")
        .put("pipeline_workspace_capacity_bytes", f.pipeline_capacity)
        .raw("workspace_allocated_bytes", f.status == "COMPLETED" ? std::to_string(f.workspace_reserved) : "null")
        .put("declared_device_bytes", f.declared_device)
        .put("free_device_bytes_before", f.free_device)
        .put("total_device_bytes", f.total_device)
        .raw("peak_vram_bytes", "null")
        .put("canvas_width", c.width)
        .put("canvas_height", c.height)
        .put("origin_x", c.origin_x)
        .put("origin_y", c.origin_y)
        .put("homography_source", o.render_only ? "frozen_ground_truth" : "s5_estimated_original_rgb")
        .put("input_width", o.width)
        .put("input_height", o.height)
        .put("seed", o.config.seed)
        .put("warmups", o.warmups)
        .put("repeats", o.repeats)
        .put("S", o.config.S)
        .put("K", o.config.K)
        .put("R", o.config.R)
        .put("memory_scope", "declared_allocations_not_peak_vram");
    return r.str();
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    std::unique_ptr<Artifacts> out;
    try {
        o = read_options(argc, argv);
        out = std::make_unique<Artifacts>(o.out_dir);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    RunFacts facts;
    MetricsTable table;
    try {
        // The host copies of both inputs (and decoding headroom) must fit the host budget.
        admission::fit(admission::add(admission::multiply(admission::multiply(o.width, o.height), 6), 64ull << 20),
                       o.host_budget, 0, "ERR_HOST_OOM");
        if (o.render_only) render_track(o, *out, &table, &facts);
        else pipeline_track(o, &table, &facts);
        facts.status = "COMPLETED";
    } catch (const Error& e) {
        facts.status = e.code;
        facts.reason = e.what();
    } catch (const HipFailure& e) {
        facts.status = e.code == hipErrorOutOfMemory ? "ERR_DEVICE_OOM" : "ERR_CUDA";
        facts.reason = e.what();
    } catch (const std::bad_alloc& e) {
        facts.status = "ERR_HOST_OOM";
        facts.reason = e.what();
    } catch (const std::exception& e) {
        facts.status = "ERR_RUNNER";
        facts.reason = e.what();
    }
    try {
        out->write("native_metrics.csv", table.str());
        out->write("native_receipt.json", receipt_of(o, facts));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "artifact publication failed: %s\n", e.what());
        return 2;
    }
    return 0;
}
