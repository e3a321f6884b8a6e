// stitch_stress: the multi-resolution stress driver (the reference's CH-05
// driver, harness/stress_multires.cu:14-200, with its bounded PPM input and
// artifact publication, harness/stress_io.hpp:17-98).
//
// Two tracks: s6_only renders a given homography (--h-file) with S6 alone,
// checks the guard bytes past every image row and past the workspace, and
// dumps regions of interest (--rois) and optionally the whole canvas;
// end_to_end runs the plain per-pair runner. Both write native_metrics.csv
// (one row per warmup and repeat) and native_receipt.json into a fresh
// output directory, each file created exclusively and linked into place.
// The device check takes the GPU architecture (--expected-arch, default
// gfx950); --expected-sm is accepted for command-line compatibility.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "stitch/runner.hpp"

namespace {

namespace fs = std::filesystem;
using namespace stitch;
using Clock = std::chrono::steady_clock;

struct DeviceError : std::runtime_error {
    hipError_t code;
    explicit DeviceError(hipError_t e) : std::runtime_error(hipGetErrorString(e)), code(e) {}
};
void hip(hipError_t e) {
    if (e != hipSuccess) throw DeviceError(e);
}

std::string quote(const std::string& v) {
    std::ostringstream o;
    o << '"';
    for (unsigned char c : v) {
        if (c == '\\' || c == '"') o << '\\' << static_cast<char>(c);
        else if (c < 32) o << "\\u" << std::hex << std::setw(4) << std::setfill('0') << int(c) << std::dec;
        else o << static_cast<char>(c);
    }
    o << '"';
    return o.str();
}

// Exclusive create of <path>.partial, then a hard link to <path>: a reader
// never sees a partly written artifact, and nothing is overwritten.
void publish(const fs::path& path, const void* bytes, std::size_t n) {
    const std::string tmp = path.string() + ".partial";
    std::FILE* f = std::fopen(tmp.c_str(), "wbx");
    if (!f) throw std::runtime_error("exclusive artifact create failed");
    const std::size_t written = std::fwrite(bytes, 1, n, f);
    const int closed = std::fclose(f);
    if (written != n || closed) throw std::runtime_error("artifact write failed");
    fs::create_hard_link(tmp, path);
    fs::remove(tmp);
}
void publish(const fs::path& path, const std::string& text) { publish(path, text.data(), text.size()); }

// Binary PPM of exactly width x height (header lines "P6", "W H", "255").
std::vector<std::uint8_t> read_ppm(const std::string& path, int w, int h) {
    std::ifstream in(path, std::ios::binary);
    std::string magic, dims, depth;
    std::getline(in, magic);
    std::getline(in, dims);
    std::getline(in, depth);
    if (!in || magic != "P6" || dims != std::to_string(w) + " " + std::to_string(h) || depth != "255")
        throw admission::Error("ERR_INPUT");
    const std::size_t bytes = admission::multiply(admission::multiply(w, h), 3);
    const auto offset = static_cast<std::uintmax_t>(in.tellg());
    if (fs::file_size(path) != offset + bytes) throw admission::Error("ERR_INPUT");
    std::vector<std::uint8_t> px(bytes);
    in.read(reinterpret_cast<char*>(px.data()), static_cast<std::streamsize>(bytes));
    if (!in) throw admission::Error("ERR_INPUT");
    return px;
}

struct DeviceBuffer {
    void* data{nullptr};
    explicit DeviceBuffer(std::size_t n) {
        if (n) hip(hipMalloc(&data, n));
    }
    ~DeviceBuffer() {
        if (data) hipFree(data);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// Pitched device image (rows padded by 64 bytes of 0xa5 guard).
struct DeviceImage {
    DeviceBuffer mem;
    ImageView view{};
    int channels;
    DeviceImage(int w, int h, int c) : mem(admission::image_bytes(w, h, c)), channels(c) {
        view = {static_cast<std::uint8_t*>(mem.data), w, h, w * c + 64, c == 3 ? PixelFormat::RGB8 : PixelFormat::GRAY8};
        hip(hipMemset(mem.data, 0xa5, admission::image_bytes(w, h, c)));
    }
    void upload(const std::vector<std::uint8_t>& px) {
        hip(hipMemcpy2D(view.data, view.pitch_bytes, px.data(), view.width * 3, view.width * 3, view.height,
                        hipMemcpyHostToDevice));
    }
    void check_guard() const {
        std::vector<std::uint8_t> pad(static_cast<std::size_t>(64) * view.height);
        hip(hipMemcpy2D(pad.data(), 64, view.data + static_cast<std::size_t>(view.width) * channels, view.pitch_bytes, 64,
                        view.height, hipMemcpyDeviceToHost));
        for (std::uint8_t b : pad)
            if (b != 0xa5) throw admission::Error("ERR_CORRECTNESS");
    }
    void save(const fs::path& path, const std::array<int, 4>& box) const {
        const int x = box[0], y = box[1], w = box[2], h = box[3];
        if (x < 0 || y < 0 || w <= 0 || h <= 0 || w > view.width - x || h > view.height - y)
            throw admission::Error("ERR_CANVAS_OVERFLOW");
        std::vector<std::uint8_t> out(admission::multiply(admission::multiply(w, h), channels));
        hip(hipMemcpy2D(out.data(), static_cast<std::size_t>(w) * channels,
                        view.data + static_cast<std::size_t>(y) * view.pitch_bytes + static_cast<std::size_t>(x) * channels,
                        view.pitch_bytes, static_cast<std::size_t>(w) * channels, h, hipMemcpyDeviceToHost));
        publish(path, out.data(), out.size());
    }
};

struct Options {
    std::string track, mode, a, b, h, rois, output, arch{"gfx950"};
    int width{0}, height{0}, canvas_w{0}, canvas_h{0}, warmups{2}, repeats{10};
    std::size_t host_budget{512u * 1024u * 1024u}, reserve{256u * 1024u * 1024u};
    bool full{false};
    harness::HarnessConfig config{};
};

Options parse(int argc, char** argv) {
    std::map<std::string, std::string> args;
    for (int i = 1; i < argc; ++i) {
        const std::string key = argv[i];
        if (key == "--export-full") {
            if (!args.emplace(key, "1").second) throw std::invalid_argument("duplicate flag");
            continue;
        }
        if (i + 1 >= argc || !args.emplace(key, argv[++i]).second)
            throw std::invalid_argument("missing/duplicate argument");
    }
    static const char* known[] = {"--track", "--mode", "--a", "--b", "--h-file", "--rois", "--output", "--width",
                                  "--height", "--canvas-width", "--canvas-height", "--warmups", "--repeats",
                                  "--expected-sm", "--expected-arch", "--host-budget-bytes",
                                  "--device-reserve-bytes", "--scale", "--keypoints", "--ransac", "--seed",
                                  "--export-full"};
    for (const auto& kv : args)
        if (std::find_if(std::begin(known), std::end(known), [&](const char* k) { return kv.first == k; }) ==
            std::end(known))
            throw std::invalid_argument("unknown argument");
    auto required = [&](const char* name) {
        if (!args.count(name)) throw std::invalid_argument("required argument missing");
        return args.at(name);
    };
    auto integer = [&](const char* name, std::size_t fallback) {
        if (!args.count(name)) return fallback;
        const std::string& v = args.at(name);
        if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
            throw std::invalid_argument("nonnegative integer required");
        return static_cast<std::size_t>(std::stoull(v));
    };
    auto dimension = [&](const char* name) {
        const std::size_t v = integer(name, 0);
        if (!v || v > static_cast<std::size_t>(INT_MAX)) throw std::invalid_argument("invalid dimension");
        return static_cast<int>(v);
    };
    Options o;
    o.track = required("--track");
    o.mode = required("--mode");
    o.a = required("--a");
    o.b = required("--b");
    o.output = required("--output");
    o.width = dimension("--width");
    o.height = dimension("--height");
    if (o.track != "s6_only" && o.track != "end_to_end") throw std::invalid_argument("unknown track");
    if (o.track == "end_to_end" && (args.count("--h-file") || args.count("--rois") || args.count("--export-full")))
        throw std::invalid_argument("end-to-end forbids ground-truth and render-oracle arguments");
    if (o.track == "s6_only") {
        o.h = required("--h-file");
        o.rois = required("--rois");
    } else {
        o.canvas_w = dimension("--canvas-width");
        o.canvas_h = dimension("--canvas-height");
    }
    if (o.mode == "feather") o.config.blend_mode = s6::BlendMode::Feather;
    else if (o.mode == "multiband_3") o.config.blend_mode = s6::BlendMode::Multiband3;
    else if (o.mode == "multiband_5") o.config.blend_mode = s6::BlendMode::Multiband5;
    else throw std::invalid_argument("unknown blender");
    const std::size_t warm = integer("--warmups", 2), rep = integer("--repeats", 10);
    if (warm > 1000 || rep < 1 || rep > 1000) throw std::invalid_argument("invalid repetitions");
    o.warmups = static_cast<int>(warm);
    o.repeats = static_cast<int>(rep);
    (void)integer("--expected-sm", 0);
    if (args.count("--expected-arch")) o.arch = args.at("--expected-arch");
    o.host_budget = integer("--host-budget-bytes", o.host_budget);
    o.reserve = integer("--device-reserve-bytes", o.reserve);
    const std::size_t k = integer("--keypoints", 1000), r = integer("--ransac", 256);
    if (!k || k > INT_MAX || !r || r > INT_MAX) throw std::invalid_argument("invalid K/R");
    o.config.K = static_cast<int>(k);
    o.config.R = static_cast<int>(r);
    o.config.seed = integer("--seed", 0);
    o.config.config_id = "stress_fixed";
    if (args.count("--scale")) {
        std::size_t end = 0;
        o.config.S = std::stod(args.at("--scale"), &end);
        if (end != args.at("--scale").size() || !std::isfinite(o.config.S) || !(o.config.S > 0 && o.config.S <= 1))
            throw std::invalid_argument("invalid scale");
    }
    o.full = args.count("--export-full") != 0;
    return o;
}

std::string number(double v) {
    std::ostringstream s;
    s << std::setprecision(17) << v;
    return s.str();
}

void write_row(std::ostream& out, int repeat, bool warmup, const std::string& status, const std::array<double, 8>& t,
               const harness::RunRecord* rec, const s6::Geometry* g, std::size_t workspace) {
    std::vector<std::string> f{std::to_string(repeat), warmup ? "1" : "0", status};
    for (double v : t) f.push_back(v >= 0 && std::isfinite(v) ? number(v) : "");
    f.push_back(rec ? std::to_string(rec->num_inliers) : "");
    f.push_back(rec ? number(rec->inlier_ratio) : "");
    for (int i = 0; i < 9; ++i) f.push_back(rec && status == "SUCCESS" ? number(rec->H[i]) : "");
    if (g) {
        for (int v : {g->width, g->height, g->origin_x, g->origin_y}) f.push_back(std::to_string(v));
        f.push_back(std::to_string(workspace));
    } else {
        for (int i = 0; i < 5; ++i) f.emplace_back();
    }
    for (std::size_t i = 0; i < f.size(); ++i) out << (i ? "," : "") << f[i];
    out << '\n';
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    try {
        o = parse(argc, argv);
        if (fs::exists(o.output)) throw std::runtime_error("output already exists");
        fs::create_directories(o.output);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    std::string status = "ERR_RUNNER", reason, gpu_name;
    int cw = 0, ch = 0, ox = 0, oy = 0, runtime_version = 0, driver_version = 0;
    std::size_t workspace = 0, allocated_workspace = 0, device_bytes = 0, free_bytes = 0, total_bytes = 0;
    std::size_t pipeline_workspace_capacity = 0;
    bool workspace_queried = false;
    std::ostringstream rows;
    rows << "repeat_idx,is_warmup,status_code,t_s1,t_s2,t_s3,t_s4,t_s5,t_s6,t_gpu_steady,t_host_e2e,inlier_count,"
            "inlier_ratio,H00,H01,H02,H10,H11,H12,H20,H21,H22,canvas_width,canvas_height,origin_x,origin_y,"
            "workspace_bytes\n";
    try {
        using namespace admission;
        fit(add(multiply(multiply(o.width, o.height), 6), 64u * 1024u * 1024u), o.host_budget, 0, "ERR_HOST_OOM");
        s6::Geometry geometry{};
        if (o.track == "s6_only") {
            double h[9];
            std::ifstream in(o.h);
            for (double& v : h)
                if (!(in >> v) || !std::isfinite(v)) throw Error("ERR_INPUT");
            in >> std::ws;
            if (!in.eof()) throw Error("ERR_INPUT");
            const Status g = s6::prepare_geometry(o.width, o.height, o.width, o.height, h, {}, &geometry);
            if (g != Status::OK)
                throw Error(g == Status::INSUFFICIENT_WORKSPACE ? "ERR_CANVAS_OVERFLOW" : "ERR_DEGENERATE_GEOM");
            cw = geometry.width;
            ch = geometry.height;
            ox = geometry.origin_x;
            oy = geometry.origin_y;
            if (s6::s6_workspace_bytes(geometry, o.config.blend_mode, &workspace) != Status::OK)
                throw Error("ERR_CAPACITY");
            workspace_queried = true;
            allocated_workspace = workspace ? add(workspace, 256) : 0;
            pipeline_workspace_capacity = workspace;
            device_bytes = add(multiply(2, image_bytes(o.width, o.height, 3)),
                               add(image_bytes(cw, ch, 3), add(image_bytes(cw, ch, 1), allocated_workspace)));
            if (o.full)
                fit(add(multiply(multiply(cw, ch), 3), multiply(multiply(o.width, o.height), 6)), o.host_budget,
                    16u * 1024u * 1024u, "ERR_HOST_OOM");
        } else {
            const auto plan = harness::PipelineRunner::query_memory(o.width, o.height, o.config.K, o.config.R,
                                                                    o.canvas_w, o.canvas_h, o.config.blend_mode);
            device_bytes = plan.device_bytes;
            allocated_workspace = plan.s6_workspace_bytes ? plan.s6_workspace_bytes + 256 : 0;
            pipeline_workspace_capacity = plan.workspace_bytes;
        }
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) throw Error("ERR_NO_DEVICE");
        int device = 0;
        hip(hipGetDevice(&device));
        hipDeviceProp_t props{};
        hip(hipGetDeviceProperties(&props, device));
        gpu_name = props.gcnArchName;
        hip(hipRuntimeGetVersion(&runtime_version));
        hip(hipDriverGetVersion(&driver_version));
        if (gpu_name.rfind(o.arch, 0) != 0) throw Error("ERR_DEVICE_MISMATCH");
        hip(hipMemGetInfo(&free_bytes, &total_bytes));
        fit(device_bytes, free_bytes, o.reserve, "ERR_DEVICE_OOM");
        const auto a = read_ppm(o.a, o.width, o.height), b = read_ppm(o.b, o.width, o.height);
        if (o.track == "s6_only") {
            DeviceImage da(o.width, o.height, 3), db(o.width, o.height, 3), out(cw, ch, 3), mask(cw, ch, 1);
            DeviceBuffer work(allocated_workspace);
            hipStream_t stream = nullptr;
            hipEvent_t start = nullptr, stop = nullptr;
            hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            hip(hipEventCreate(&start));
            hip(hipEventCreate(&stop));
            if (workspace) hip(hipMemset(static_cast<std::uint8_t*>(work.data) + workspace, 0xa5, 256));
            da.upload(a);
            db.upload(b);
            for (int i = 0; i < o.warmups + o.repeats; ++i) {
                const auto before = Clock::now();
                hip(hipEventRecord(start, stream));
                const Status r = s6::render(da.view, db.view, geometry, o.config.blend_mode, out.view, mask.view,
                                            work.data, workspace, stream);
                if (r != Status::OK) throw Error(r == Status::CUDA_ERROR ? "ERR_CUDA" : "ERR_CAPACITY");
                hip(hipEventRecord(stop, stream));
                hip(hipEventSynchronize(stop));
                const double host = std::chrono::duration<double, std::milli>(Clock::now() - before).count();
                float gpu = 0;
                hip(hipEventElapsedTime(&gpu, start, stop));
                write_row(rows, i < o.warmups ? i : i - o.warmups, i < o.warmups, "SUCCESS",
                          {-1, -1, -1, -1, -1, gpu, gpu, host}, nullptr, &geometry, workspace);
            }
            hipEventDestroy(start);
            hipEventDestroy(stop);
            hipStreamDestroy(stream);
            da.check_guard();
            db.check_guard();
            out.check_guard();
            mask.check_guard();
            if (workspace) {
                std::array<std::uint8_t, 256> guard{};
                hip(hipMemcpy(guard.data(), static_cast<std::uint8_t*>(work.data) + workspace, 256,
                              hipMemcpyDeviceToHost));
                for (std::uint8_t v : guard)
                    if (v != 0xa5) throw Error("ERR_CORRECTNESS");
            }
            std::ifstream roi_file(o.rois);
            std::array<int, 4> box{};
            int index = 0;
            while (roi_file >> box[0] >> box[1] >> box[2] >> box[3]) {
                if (++index > 32) throw Error("ERR_CAPACITY");
                out.save(fs::path(o.output) / ("roi" + std::to_string(index - 1) + ".rgb"), box);
                mask.save(fs::path(o.output) / ("roi" + std::to_string(index - 1) + ".mask"), box);
            }
            if (!roi_file.eof() || index == 0) throw Error("ERR_INPUT");
            if (o.full) {
                out.save(fs::path(o.output) / "full.rgb", {0, 0, cw, ch});
                mask.save(fs::path(o.output) / "full.mask", {0, 0, cw, ch});
            }
        } else {
            harness::PipelineRunner runner;
            runner.initialize(o.width, o.height, o.config.K, o.config.R, o.canvas_w, o.canvas_h, o.config.blend_mode,
                              o.reserve);
            runner.upload_pair(a.data(), o.width, o.height, b.data(), o.width, o.height);
            for (int i = 0; i < o.warmups + o.repeats; ++i) {
                harness::RunRecord r{};
                const auto before = Clock::now();
                runner.execute_run(o.config, &r);
                runner.synchronize();
                const double host = std::chrono::duration<double, std::milli>(Clock::now() - before).count();
                const bool ok = r.status == "OK";
                if (ok) {
                    if (s6::prepare_geometry(o.width, o.height, o.width, o.height, r.H, {}, &geometry) != Status::OK)
                        throw Error("ERR_DEGENERATE_GEOM");
                    cw = geometry.width;
                    ch = geometry.height;
                    ox = geometry.origin_x;
                    oy = geometry.origin_y;
                    if (s6::s6_workspace_bytes(geometry, o.config.blend_mode, &workspace) != Status::OK)
                        throw Error("ERR_CAPACITY");
                    workspace_queried = true;
                }
                write_row(rows, i < o.warmups ? i : i - o.warmups, i < o.warmups, ok ? "SUCCESS" : r.status,
                          {ok ? r.t_s1 : -1, ok ? r.t_s2 : -1, ok ? r.t_s3 : -1, ok ? r.t_s4 : -1, ok ? r.t_s5 : -1,
                           ok ? r.t_s6 : -1, runner.last_fixed_gpu_span_ms(), host},
                          &r, ok ? &geometry : nullptr, workspace);
            }
        }
        status = "COMPLETED";
    } catch (const admission::Error& e) {
        status = e.code;
        reason = e.what();
    } catch (const DeviceError& e) {
        status = e.code == hipErrorOutOfMemory ? "ERR_DEVICE_OOM" : "ERR_CUDA";
        reason = e.what();
    } catch (const std::bad_alloc& e) {
        status = "ERR_HOST_OOM";
        reason = e.what();
    } catch (const std::exception& e) {
        status = "ERR_RUNNER";
        reason = e.what();
    }
    try {
        publish(fs::path(o.output) / "native_metrics.csv", rows.str());
        std::ostringstream m;
        m << std::setprecision(17) << "{\"schema_version\":\"multires-native-v1\",\"status_code\":" << quote(status)
          << ",\"reason\":" << quote(reason) << ",\"track\":" << quote(o.track) << ",\"blend_mode\":" << quote(o.mode)
          << ",\"gpu_name\":" << quote(gpu_name) << ",\"sm\":0,\"gpu_arch\":" << quote(gpu_name)
          << ",\"cuda_runtime_version\":" << runtime_version << ",\"cuda_driver_version\":" << driver_version
          << ",\"workspace_bytes\":" << workspace << ",\"workspace_reserved_bytes\":" << allocated_workspace
          << ",\"workspace_query_status\":"
          << quote(workspace_queried ? "QUERIED_ACTUAL_GEOMETRY" : "NOT_QUERIED_ACTUAL_GEOMETRY")
          << ",\"workspace_query_scope\":\"s6_actual_geometry\",\"pipeline_workspace_capacity_bytes\":"
          << pipeline_workspace_capacity << ",\"workspace_allocated_bytes\":"
          << (status == "COMPLETED" ? std::to_string(allocated_workspace) : "null")
          << ",\"declared_device_bytes\":" << device_bytes << ",\"free_device_bytes_before\":" << free_bytes
          << ",\"total_device_bytes\":" << total_bytes << ",\"peak_vram_bytes\":null"
          << ",\"canvas_width\":" << cw << ",\"canvas_height\":" << ch << ",\"origin_x\":" << ox
          << ",\"origin_y\":" << oy << ",\"homography_source\":"
          << quote(o.track == "s6_only" ? "frozen_ground_truth" : "s5_estimated_original_rgb")
          << ",\"input_width\":" << o.width << ",\"input_height\":" << o.height << ",\"seed\":" << o.config.seed
          << ",\"warmups\":" << o.warmups << ",\"repeats\":" << o.repeats << ",\"S\":" << o.config.S
          << ",\"K\":" << o.config.K << ",\"R\":" << o.config.R
          << ",\"memory_scope\":\"declared_allocations_not_peak_vram\"}";
        publish(fs::path(o.output) / "native_receipt.json", m.str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "artifact publication failed: %s\n", e.what());
        return 2;
    }
    return 0;
}
