// Batched form (see stitch/batch.hpp): device pool, feature and render
// engines with graph replay, the batch runner and its autotuner, and the
// gather kernel that brings a batch's small results back in one copy.
#include <hip/hip_runtime.h>

#include "stitch/batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>

#include "stitch/geometry.hpp"
#include "stitch/jobs.hpp"
#include "stitch/s6_jobs.hpp"
#include "stitch/tuning.hpp"

namespace stitch::batch {

namespace {

// One workgroup per copy; 4-byte words when both ends allow it.
__global__ void copy_list(const CopyOp* ops, const int* count_dev, int count) {
    const int n = count_dev ? *count_dev : count;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const CopyOp op = ops[i];
        const auto* s = static_cast<const std::uint8_t*>(op.src);
        auto* d = static_cast<std::uint8_t*>(op.dst);
        const bool words = (op.bytes % 4 == 0) && (reinterpret_cast<std::uintptr_t>(s) % 4 == 0) &&
                           (reinterpret_cast<std::uintptr_t>(d) % 4 == 0);
        if (words) {
            for (int k = threadIdx.x; k < op.bytes / 4; k += blockDim.x)
                reinterpret_cast<std::uint32_t*>(d)[k] = reinterpret_cast<const std::uint32_t*>(s)[k];
        } else {
            for (int k = threadIdx.x; k < op.bytes; k += blockDim.x) d[k] = s[k];
        }
    }
}

}  // namespace

hipError_t launch_copy_list(const CopyOp* ops, const int* count_dev, int capacity, hipStream_t stream) {
    if (capacity <= 0) return hipSuccess;
    copy_list<<<capacity < 1024 ? capacity : 1024, 64, 0, stream>>>(ops, count_dev, capacity);
    return hipGetLastError();
}


namespace {

constexpr std::size_t kAlign = 256;
constexpr int kPadding = 64;               // row padding of gray images and canvases
constexpr long long kReplayTiles = 1LL << 40;  // grid sizing of captured launches: full persistent grids

std::size_t align_up(std::size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Lays consecutive 256-B aligned buffers out of one block (sizing pass when base is null).
struct Carver {
    std::uint8_t* base;
    std::size_t offset = 0;
    void* take(std::size_t bytes) {
        offset = align_up(offset);
        void* p = base ? base + offset : nullptr;
        offset += std::max<std::size_t>(bytes, 1);
        return p;
    }
};

std::size_t workspace(Status s, std::size_t bytes) {
    if (s != Status::OK) throw admission::Error("ERR_CAPACITY");
    return bytes;
}

void carve_pair(Carver& c, const PairCapacity& cap, PairBuffers* b) {
    using admission::image_bytes;
    const int w = cap.width, h = cap.height, K = cap.K;
    auto gray = [&](ImageView* v) {
        v->width = w;
        v->height = h;
        v->pitch_bytes = w + kPadding;
        v->format = PixelFormat::GRAY8;
        v->data = static_cast<std::uint8_t*>(c.take(static_cast<std::size_t>(v->pitch_bytes) * h));
    };
    gray(&b->gray_a);
    gray(&b->gray_b);
    for (Keypoints* k : {&b->kp_a, &b->kp_b}) {
        k->capacity = K;
        k->x = static_cast<float*>(c.take(sizeof(float) * K));
        k->y = static_cast<float*>(c.take(sizeof(float) * K));
        k->response = static_cast<float*>(c.take(sizeof(float) * K));
        k->d_count = static_cast<int*>(c.take(sizeof(int)));
    }
    for (Descriptors* d : {&b->desc_a, &b->desc_b}) {
        d->type = DescriptorType::BINARY_U8;
        d->dim = 32;
        d->row_stride_bytes = 32;
        d->data = c.take(static_cast<std::size_t>(K) * 32);
    }
    b->matches.capacity = K;
    b->matches.idx_a = static_cast<std::int32_t*>(c.take(sizeof(std::int32_t) * K));
    b->matches.idx_b = static_cast<std::int32_t*>(c.take(sizeof(std::int32_t) * K));
    b->matches.distance = static_cast<float*>(c.take(sizeof(float) * K));
    b->matches.d_count = static_cast<int*>(c.take(sizeof(int)));
    b->inlier_mask = static_cast<std::uint8_t*>(c.take(K));
    b->grad_accum = static_cast<unsigned long long*>(c.take(2 * sizeof(unsigned long long)));
    b->grad = static_cast<double*>(c.take(2 * sizeof(double)));
    std::size_t bytes = 0;
    b->bytes_s2 = workspace(s2::s2_workspace_bytes(w, h, {20, K, 0.04f}, &bytes), bytes);
    b->bytes_s3 = workspace(s3::s3_workspace_bytes(w, h, &bytes), bytes);
    b->bytes_s4 = workspace(s4::s4_workspace_bytes(K, K, {0.8, true}, &bytes), bytes);
    b->bytes_s5 = workspace(s5::s5_workspace_bytes({cap.R, 0, 64, 1, 3.0, 0, 1}, &bytes), bytes);
    b->ws_s2_a = c.take(b->bytes_s2);
    b->ws_s2_b = c.take(b->bytes_s2);
    b->ws_s3_a = c.take(b->bytes_s3);
    b->ws_s3_b = c.take(b->bytes_s3);
    b->ws_s4 = c.take(b->bytes_s4);
    b->ws_s5 = c.take(b->bytes_s5);
}

void validate(const PairCapacity& cap) {
    if (cap.width <= 0 || cap.height <= 0 || cap.K <= 0 || cap.R <= 0) throw admission::Error("ERR_CAPACITY");
    admission::image_bytes(cap.width, cap.height, 3);
}

void carve_render(Carver& c, const s6::Geometry& g, s6::BlendMode mode, RenderBuffers* r) {
    r->canvas.width = g.width;
    r->canvas.height = g.height;
    r->canvas.pitch_bytes = g.width * 3 + kPadding;
    r->canvas.format = PixelFormat::RGB8;
    r->canvas.data = static_cast<std::uint8_t*>(c.take(static_cast<std::size_t>(r->canvas.pitch_bytes) * g.height));
    r->mask.width = g.width;
    r->mask.height = g.height;
    r->mask.pitch_bytes = g.width + kPadding;
    r->mask.format = PixelFormat::GRAY8;
    r->mask.data = static_cast<std::uint8_t*>(c.take(static_cast<std::size_t>(r->mask.pitch_bytes) * g.height));
    r->stats = static_cast<s6::RenderStatistics*>(c.take(sizeof(s6::RenderStatistics)));
    std::size_t bytes = 0;
    if (s6::s6_workspace_bytes(g, mode, &bytes) != Status::OK) throw admission::Error("ERR_CAPACITY");
    r->ws_bytes = bytes;
    r->ws = bytes ? c.take(bytes) : nullptr;
}

// Results gathered on the device and copied back in one piece.
struct FeatureOut {
    HomographyResult h;
    int kp_a, kp_b;
    double grad[2];
};

template <class T>
T* at(std::uint8_t* base, std::size_t off) {
    return reinterpret_cast<T*>(base + off);
}

}  // namespace

// ------------------------------------------------------------------ pool
DevicePool::~DevicePool() {
    if (base_) hipFree(base_);
}

void DevicePool::initialize(std::size_t bytes) {
    if (base_) return;
    bytes = align_up(std::max<std::size_t>(bytes, kAlign));
    std::size_t free_bytes = 0, total = 0;
    check(hipMemGetInfo(&free_bytes, &total), "hipMemGetInfo");
    admission::fit(bytes, free_bytes, 0, "ERR_DEVICE_OOM");
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw admission::Error("ERR_DEVICE_OOM");
    base_ = static_cast<std::uint8_t*>(p);
    size_ = bytes;
    reset();
}

void* DevicePool::allocate(std::size_t bytes) {
    const std::size_t n = align_up(std::max<std::size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        const std::size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        used_[off] = n;
        in_use_ += n;
        peak_ = std::max(peak_, in_use_);
        return base_ + off;
    }
    return nullptr;
}

void DevicePool::release(void* p) {
    if (!p || !base_) return;
    const std::size_t off = static_cast<std::size_t>(static_cast<std::uint8_t*>(p) - base_);
    const auto u = used_.find(off);
    if (u == used_.end()) return;
    std::size_t start = off, len = u->second;
    used_.erase(u);
    in_use_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && start + len == next->first) {
        len += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == start) {
            prev->second += len;
            return;
        }
    }
    free_[start] = len;
}

void DevicePool::reset() {
    free_.clear();
    used_.clear();
    in_use_ = 0;
    if (size_) free_[0] = size_;
}

// ---------------------------------------------------------- pair buffers
std::size_t pair_bytes(const PairCapacity& cap) {
    validate(cap);
    Carver c{nullptr};
    PairBuffers b{};
    carve_pair(c, cap, &b);
    return align_up(c.offset);
}

bool allocate_pair(DevicePool& pool, const PairCapacity& cap, PairBuffers* out) {
    const std::size_t bytes = pair_bytes(cap);
    void* block = pool.allocate(bytes);
    if (!block) return false;
    PairBuffers b{};
    b.cap = cap;
    Carver c{static_cast<std::uint8_t*>(block)};
    carve_pair(c, cap, &b);
    b.block = block;
    b.block_bytes = bytes;
    *out = b;
    return true;
}

void release_pair(DevicePool& pool, PairBuffers* b) {
    if (b && b->block) pool.release(b->block);
    if (b) *b = PairBuffers{};
}

std::size_t render_bytes(const s6::Geometry& g, s6::BlendMode mode) {
    admission::image_bytes(g.width, g.height, 3);
    Carver c{nullptr};
    RenderBuffers r{};
    carve_render(c, g, mode, &r);
    return align_up(c.offset);
}

bool allocate_render(DevicePool& pool, const s6::Geometry& g, s6::BlendMode mode, RenderBuffers* out) {
    const std::size_t bytes = render_bytes(g, mode);
    void* block = pool.allocate(bytes);
    if (!block) return false;
    RenderBuffers r{};
    Carver c{static_cast<std::uint8_t*>(block)};
    carve_render(c, g, mode, &r);
    r.block = block;
    r.block_bytes = bytes;
    *out = r;
    return true;
}

void release_render(DevicePool& pool, RenderBuffers* r) {
    if (r && r->block) pool.release(r->block);
    if (r) *r = RenderBuffers{};
}

// -------------------------------------------------------- feature engine
struct FeatureEngine::Header {
    int n_s1, n_grad, n_s2, n_s3, n_s4, n_s5, n_copy, pad;
    long long t_s1, t_grad;
    jobs::S2Totals t2;
    jobs::S3Totals t3;
    jobs::S4Totals t4;
    jobs::S5Totals t5;
};

FeatureEngine::~FeatureEngine() {
    if (stream_) hipStreamSynchronize(stream_);
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    if (dev_) hipFree(dev_);
    if (staging_) hipHostFree(staging_);
    if (out_) hipHostFree(out_);
    if (ev_begin_) hipEventDestroy(ev_begin_);
    if (ev_end_) hipEventDestroy(ev_end_);
    if (stream_) hipStreamDestroy(stream_);
}

void FeatureEngine::initialize(int max_tasks, bool use_graph) {
    if (dev_) return;
    if (max_tasks <= 0) throw admission::Error("ERR_CAPACITY");
    max_tasks_ = max_tasks;
    use_graph_ = use_graph;
    const std::size_t M = max_tasks, M2 = 2 * M;
    std::size_t off = align_up(sizeof(Header));
    auto place = [&](std::size_t bytes) {
        const std::size_t o = off;
        off = align_up(off + bytes);
        return o;
    };
    off_s1_ = place(M2 * sizeof(jobs::S1Job));
    off_grad_ = place(M2 * sizeof(jobs::GradJob));
    off_s2_ = place(M2 * sizeof(jobs::S2Job));
    off_s3_ = place(M2 * sizeof(jobs::S3Job));
    off_s4_ = place(M * sizeof(jobs::S4Job));
    off_s5_ = place(M * sizeof(jobs::S5Job));
    off_copy_ = place(4 * M * sizeof(CopyOp));
    table_bytes_ = off;
    off_out_ = place(M * sizeof(FeatureOut));
    out_bytes_ = M * sizeof(FeatureOut);
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
    check(hipEventCreate(&ev_begin_), "event");
    check(hipEventCreate(&ev_end_), "event");
    check(hipMalloc(reinterpret_cast<void**>(&dev_), off), "hipMalloc");
    check(hipHostMalloc(reinterpret_cast<void**>(&staging_), table_bytes_, hipHostMallocDefault), "hipHostMalloc");
    check(hipHostMalloc(reinterpret_cast<void**>(&out_), out_bytes_, hipHostMallocDefault), "hipHostMalloc");
    std::memset(staging_, 0, table_bytes_);
}

void FeatureEngine::launch(bool replay) {
    auto* d = reinterpret_cast<const Header*>(dev_);
    Header cap{};
    const Header& h = replay ? cap : *reinterpret_cast<const Header*>(staging_);
    if (replay) {
        cap.n_s1 = cap.n_grad = cap.n_s2 = cap.n_s3 = 2 * max_tasks_;
        cap.n_s4 = cap.n_s5 = max_tasks_;
        cap.n_copy = 4 * max_tasks_;
        cap.t_s1 = cap.t_grad = kReplayTiles;
        // Every totals field gets the capacity (a field left at 0 would size
        // its kernel's captured grid to a single workgroup).
        static_assert(sizeof(jobs::S2Totals) == 6 * sizeof(long long), "S2Totals fields");
        static_assert(sizeof(jobs::S3Totals) == 3 * sizeof(long long), "S3Totals fields");
        static_assert(sizeof(jobs::S4Totals) == 6 * sizeof(long long), "S4Totals fields");
        static_assert(sizeof(jobs::S5Totals) == 2 * sizeof(long long), "S5Totals fields");
        cap.t2 = {kReplayTiles, kReplayTiles, kReplayTiles, kReplayTiles, kReplayTiles, kReplayTiles};
        cap.t3 = {kReplayTiles, kReplayTiles, kReplayTiles};
        cap.t4 = {kReplayTiles, kReplayTiles, kReplayTiles, kReplayTiles, kReplayTiles, kReplayTiles};
        cap.t5 = {kReplayTiles, kReplayTiles};
    }
    const dev::Replay<long long> r1{&d->n_s1, &d->t_s1}, rg{&d->n_grad, &d->t_grad};
    const dev::Replay<jobs::S2Totals> r2{&d->n_s2, &d->t2};
    const dev::Replay<jobs::S3Totals> r3{&d->n_s3, &d->t3};
    const dev::Replay<jobs::S4Totals> r4{&d->n_s4, &d->t4};
    const dev::Replay<jobs::S5Totals> r5{&d->n_s5, &d->t5};
    hipError_t e = hipSuccess;
    auto run = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    if (h.n_s1 > 0) run(jobs::launch_s1(at<jobs::S1Job>(dev_, off_s1_), h.n_s1, h.t_s1, nullptr, stream_, replay ? &r1 : nullptr));
    if (h.n_grad > 0)
        run(jobs::launch_gradient(at<jobs::GradJob>(dev_, off_grad_), h.n_grad, h.t_grad, nullptr, stream_,
                                  replay ? &rg : nullptr));
    if (h.n_s2 > 0) run(jobs::launch_s2(at<jobs::S2Job>(dev_, off_s2_), h.n_s2, h.t2, nullptr, stream_, replay ? &r2 : nullptr));
    if (h.n_s3 > 0) run(jobs::launch_s3(at<jobs::S3Job>(dev_, off_s3_), h.n_s3, h.t3, nullptr, stream_, replay ? &r3 : nullptr));
    if (h.n_s4 > 0) run(jobs::launch_s4(at<jobs::S4Job>(dev_, off_s4_), h.n_s4, h.t4, nullptr, stream_, replay ? &r4 : nullptr));
    if (h.n_s5 > 0) run(jobs::launch_s5(at<jobs::S5Job>(dev_, off_s5_), h.n_s5, h.t5, nullptr, stream_, replay ? &r5 : nullptr));
    if (h.n_copy > 0) run(launch_copy_list(at<CopyOp>(dev_, off_copy_), replay ? &d->n_copy : nullptr, h.n_copy, stream_));
    check(e, "feature launch");
}

void FeatureEngine::capture() {
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    persistent_grid(1, jobs::Stage::S1);  // device query outside the capture
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "begin capture");
    try {
        launch(true);
    } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(stream_, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    check(hipStreamEndCapture(stream_, &graph_), "end capture");
    check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "instantiate");
}

void FeatureEngine::invalidate() {
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

void FeatureEngine::submit(const FeatureTask* tasks, int n) {
    if (!dev_) throw admission::Error("ERR_CAPACITY");
    if (pending_) wait();
    if (n < 0 || n > max_tasks_) throw admission::Error("ERR_CAPACITY");
    results_.assign(n, FeatureResult{});
    out_slot_.assign(n, -1);
    std::vector<jobs::S1Job> v1;
    std::vector<jobs::GradJob> vg;
    std::vector<jobs::S2Job> v2;
    std::vector<jobs::S3Job> v3;
    std::vector<jobs::S4Job> v4;
    std::vector<jobs::S5Job> v5;
    std::vector<CopyOp> vc;
    auto* outs = at<FeatureOut>(dev_, off_out_);
    for (int i = 0; i < n; ++i) {
        const FeatureTask& t = tasks[i];
        FeatureResult& r = results_[i];
        auto fail = [&](Status s, const char* stage) {
            r.status = s;
            r.fail_stage = stage;
            r.h = HomographyResult{};
            r.h.status = s;
        };
        if (!t.buffers) {
            fail(Status::INVALID_ARGUMENT, "s1");
            continue;
        }
        PairBuffers& B = *t.buffers;
        if (make_scale_info(t.a.width, t.a.height, t.S, &r.sa) != Status::OK ||
            make_scale_info(t.b.width, t.b.height, t.S, &r.sb) != Status::OK || r.sa.work_w > B.cap.width ||
            r.sa.work_h > B.cap.height || r.sb.work_w > B.cap.width || r.sb.work_h > B.cap.height) {
            fail(Status::INVALID_ARGUMENT, "s1");
            continue;
        }
        ImageView ga = B.gray_a, gb = B.gray_b;
        ga.width = r.sa.work_w;
        ga.height = r.sa.work_h;
        gb.width = r.sb.work_w;
        gb.height = r.sb.work_h;
        jobs::S1Job j1[2]{};
        jobs::GradJob jg[2]{};
        jobs::S2Job j2[2]{};
        jobs::S3Job j3[2]{};
        jobs::S4Job j4{};
        jobs::S5Job j5{};
        Status s = Status::OK;
        if (t.run_s1 && ((s = jobs::make_s1_job(t.a, ga, r.sa, &j1[0])) != Status::OK ||
                         (s = jobs::make_s1_job(t.b, gb, r.sb, &j1[1])) != Status::OK)) {
            fail(s, "s1");
            continue;
        }
        if (t.probe) {
            if (ga.width < 2 || ga.height < 2 || gb.width < 2 || gb.height < 2) {
                fail(Status::INVALID_ARGUMENT, "s1");
                continue;
            }
            const ImageView* g2[2] = {&ga, &gb};
            for (int k = 0; k < 2; ++k)
                jg[k] = {g2[k]->data, g2[k]->pitch_bytes, g2[k]->width, g2[k]->height, B.grad_accum + k, B.grad + k,
                         (g2[k]->width + jobs::kS1TileW - 1) / jobs::kS1TileW, 0};
        }
        Keypoints ka = B.kp_a, kb = B.kp_b;
        const int K = t.s2.max_keypoints;
        if (K < 0 || K > B.cap.K) {
            fail(Status::INVALID_ARGUMENT, "s2");
            continue;
        }
        ka.capacity = kb.capacity = K;
        if (t.run_s2s3) {
            if ((s = jobs::make_s2_job(ga, t.s2, ka, B.ws_s2_a, B.bytes_s2, &j2[0])) != Status::OK ||
                (s = jobs::make_s2_job(gb, t.s2, kb, B.ws_s2_b, B.bytes_s2, &j2[1])) != Status::OK) {
                fail(s, "s2");
                continue;
            }
            if ((s = jobs::make_s3_job(ga, ka, B.desc_a, B.ws_s3_a, B.bytes_s3, &j3[0])) != Status::OK ||
                (s = jobs::make_s3_job(gb, kb, B.desc_b, B.ws_s3_b, B.bytes_s3, &j3[1])) != Status::OK) {
                fail(s, "s3");
                continue;
            }
        }
        if (t.run_s4 &&
            (s = jobs::make_s4_job(ka, B.desc_a, kb, B.desc_b, t.s4, B.matches, B.ws_s4, B.bytes_s4, &j4)) !=
                Status::OK) {
            fail(s, "s4");
            continue;
        }
        if (t.run_s5) {
            if (t.s5.hypotheses > B.cap.R || t.s5.hypotheses - t.s5.previous_hypotheses > jobs::kS5MaxModels) {
                fail(Status::INVALID_ARGUMENT, "s5");
                continue;
            }
            if ((s = jobs::make_s5_job(ka, kb, B.matches, r.sa, r.sb, t.s5, B.inlier_mask, B.ws_s5, B.bytes_s5,
                                       &j5)) != Status::OK) {
                fail(s, "s5");
                continue;
            }
        }
        // Every job of the task was built: commit them.
        if (t.run_s1) v1.insert(v1.end(), j1, j1 + 2);
        if (t.probe) vg.insert(vg.end(), jg, jg + 2);
        if (t.run_s2s3) {
            v2.insert(v2.end(), j2, j2 + 2);
            v3.insert(v3.end(), j3, j3 + 2);
        }
        if (t.run_s4) v4.push_back(j4);
        if (t.run_s5) {
            v5.push_back(j5);
            vc.push_back({j5.result_dev, &outs[i].h, static_cast<int>(sizeof(HomographyResult))});
        }
        vc.push_back({ka.d_count, &outs[i].kp_a, static_cast<int>(sizeof(int))});
        vc.push_back({kb.d_count, &outs[i].kp_b, static_cast<int>(sizeof(int))});
        if (t.probe) vc.push_back({B.grad, &outs[i].grad[0], static_cast<int>(2 * sizeof(double))});
        out_slot_[i] = i;
    }
    // Tile prefixes of every kernel.
    Header& h = *reinterpret_cast<Header*>(staging_);
    h = Header{};
    for (auto& j : v1) {
        j.tile_begin = h.t_s1;
        h.t_s1 += jobs::s1_tiles(j);
    }
    for (auto& j : vg) {
        j.tile_begin = h.t_grad;
        h.t_grad += jobs::gradient_tiles(j);
    }
    h.t2 = jobs::s2_prefix(v2.data(), static_cast<int>(v2.size()));
    h.t3 = jobs::s3_prefix(v3.data(), static_cast<int>(v3.size()));
    h.t4 = jobs::s4_prefix(v4.data(), static_cast<int>(v4.size()));
    h.t5 = jobs::s5_prefix(v5.data(), static_cast<int>(v5.size()));
    h.n_s1 = static_cast<int>(v1.size());
    h.n_grad = static_cast<int>(vg.size());
    h.n_s2 = static_cast<int>(v2.size());
    h.n_s3 = static_cast<int>(v3.size());
    h.n_s4 = static_cast<int>(v4.size());
    h.n_s5 = static_cast<int>(v5.size());
    h.n_copy = static_cast<int>(vc.size());
    auto put = [&](std::size_t off, const auto& v) {
        if (!v.empty()) std::memcpy(staging_ + off, v.data(), v.size() * sizeof(v[0]));
    };
    put(off_s1_, v1);
    put(off_grad_, vg);
    put(off_s2_, v2);
    put(off_s3_, v3);
    put(off_s4_, v4);
    put(off_s5_, v5);
    put(off_copy_, vc);
    check(hipMemcpyAsync(dev_, staging_, table_bytes_, hipMemcpyHostToDevice, stream_), "upload tables");
    check(hipEventRecord(ev_begin_, stream_), "event");
    if (use_graph_) {
        if (!exec_) capture();
        check(hipGraphLaunch(exec_, stream_), "graph launch");
    } else {
        launch(false);
    }
    check(hipMemcpyAsync(out_, dev_ + off_out_, out_bytes_, hipMemcpyDeviceToHost, stream_), "download results");
    check(hipEventRecord(ev_end_, stream_), "event");
    pending_ = true;
}

bool FeatureEngine::ready() {
    if (!pending_) return true;
    const hipError_t q = hipEventQuery(ev_end_);
    if (q == hipErrorNotReady) return false;
    check(q, "feature phase");
    wait();
    return true;
}

void FeatureEngine::wait() {
    if (!pending_) return;
    check(hipEventSynchronize(ev_end_), "feature phase");
    pending_ = false;
    const auto* outs = reinterpret_cast<const FeatureOut*>(out_);
    for (std::size_t i = 0; i < results_.size(); ++i) {
        if (out_slot_[i] < 0) continue;
        FeatureResult& r = results_[i];
        const FeatureOut& o = outs[out_slot_[i]];
        r.num_kp_a = o.kp_a;
        r.num_kp_b = o.kp_b;
        r.grad_a = o.grad[0];
        r.grad_b = o.grad[1];
        r.h = o.h;
    }
}

float FeatureEngine::last_ms() const {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_begin_, ev_end_) != hipSuccess) return -1;
    return ms;
}

// --------------------------------------------------------- render engine
struct RenderEngine::Header {
    int n_s6, n_copy;
    jobs::S6Totals t6;
};

RenderEngine::~RenderEngine() {
    if (stream_) hipStreamSynchronize(stream_);
    invalidate();
    if (dev_) hipFree(dev_);
    if (staging_) hipHostFree(staging_);
    if (out_) hipHostFree(out_);
    if (ev_begin_) hipEventDestroy(ev_begin_);
    if (ev_end_) hipEventDestroy(ev_end_);
    if (probe_) {
        for (int i = 0; i < jobs::LaunchProbe::kMax; ++i) {
            if (probe_->begin[i]) hipEventDestroy(probe_->begin[i]);
            if (probe_->end[i]) hipEventDestroy(probe_->end[i]);
        }
        delete probe_;
    }
    if (stream_) hipStreamDestroy(stream_);
}

void RenderEngine::set_probe(bool on) {
    if (pending_) wait();
    if (on && !probe_) {
        probe_ = new jobs::LaunchProbe{};
        for (int i = 0; i < jobs::LaunchProbe::kMax; ++i) {
            check(hipEventCreate(&probe_->begin[i]), "probe event");
            check(hipEventCreate(&probe_->end[i]), "probe event");
        }
    }
    if (on) probe_stats_.clear();
    probe_on_ = on;
}

void RenderEngine::read_probe() {
    if (!probe_ || probe_->count == 0) return;
    const Header& h = *reinterpret_cast<const Header*>(staging_);
    const auto* js = reinterpret_cast<const jobs::S6Job*>(staging_ + off_s6_);
    const auto* st = reinterpret_cast<const s6::RenderStatistics*>(out_);
    for (int i = 0; i < probe_->count; ++i) {
        float ms = 0;
        check(hipEventElapsedTime(&ms, probe_->begin[i], probe_->end[i]), "probe elapsed");
        double bytes = 0;
        for (int k = 0; k < h.n_s6; ++k)
            bytes += jobs::s6_algorithmic_bytes(probe_->kernel[i], probe_->level[i], js[k], st[k]);
        ProbeStat* s = nullptr;
        for (ProbeStat& e : probe_stats_)
            if (e.kernel == probe_->kernel[i] && e.level == probe_->level[i]) s = &e;
        if (!s) {
            probe_stats_.push_back(
                {jobs::probe_kernel_name(probe_->kernel[i]), probe_->kernel[i], probe_->level[i], 0, 0.0, 0.0});
            s = &probe_stats_.back();
        }
        s->launches += 1;
        s->ms += ms;
        s->bytes += bytes;
    }
    probe_->count = 0;
}

void RenderEngine::initialize(int max_tasks, bool use_graph) {
    if (dev_) return;
    if (max_tasks <= 0) throw admission::Error("ERR_CAPACITY");
    max_tasks_ = max_tasks;
    use_graph_ = use_graph;
    const std::size_t M = max_tasks;
    std::size_t off = align_up(sizeof(Header));
    auto place = [&](std::size_t bytes) {
        const std::size_t o = off;
        off = align_up(off + bytes);
        return o;
    };
    off_s6_ = place(M * sizeof(jobs::S6Job));
    off_copy_ = place(M * sizeof(CopyOp));
    table_bytes_ = off;
    off_out_ = place(M * sizeof(s6::RenderStatistics));
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
    check(hipEventCreate(&ev_begin_), "event");
    check(hipEventCreate(&ev_end_), "event");
    check(hipMalloc(reinterpret_cast<void**>(&dev_), off), "hipMalloc");
    check(hipHostMalloc(reinterpret_cast<void**>(&staging_), table_bytes_, hipHostMallocDefault), "hipHostMalloc");
    check(hipHostMalloc(reinterpret_cast<void**>(&out_), M * sizeof(s6::RenderStatistics), hipHostMallocDefault),
          "hipHostMalloc");
    std::memset(staging_, 0, table_bytes_);
}

void RenderEngine::invalidate() {
    for (auto& kv : graphs_) {
        if (kv.second.second) hipGraphExecDestroy(kv.second.second);
        if (kv.second.first) hipGraphDestroy(kv.second.first);
    }
    graphs_.clear();
}

hipGraphExec_t RenderEngine::graph_for(int top, int blend_top) {
    const int key = top * 64 + blend_top + 1;
    const auto it = graphs_.find(key);
    if (it != graphs_.end()) return it->second.second;
    jobs::S6Totals cap{};
    cap.top = top;
    cap.blend_top = blend_top;
    cap.warp = kReplayTiles;
    for (int l = 0; l <= top; ++l) {
        if (l >= 1) cap.pull[l] = kReplayTiles;
        cap.push[l] = kReplayTiles;
        if (l <= blend_top) {
            if (l >= 1) cap.gauss[l] = kReplayTiles;
            cap.mix[l] = kReplayTiles;
        }
    }
    auto* d = reinterpret_cast<const Header*>(dev_);
    const dev::Replay<jobs::S6Totals> r6{&d->n_s6, &d->t6};
    persistent_grid(1, jobs::Stage::S6);  // device query outside the capture
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "begin capture");
    hipError_t e = jobs::launch_s6(at<jobs::S6Job>(dev_, off_s6_), max_tasks_, cap, nullptr, stream_, &r6);
    if (e == hipSuccess) e = launch_copy_list(at<CopyOp>(dev_, off_copy_), &d->n_copy, max_tasks_, stream_);
    const hipError_t end = hipStreamEndCapture(stream_, &g);
    if (e != hipSuccess || end != hipSuccess) {
        if (g) hipGraphDestroy(g);
        check(e != hipSuccess ? e : end, "render capture");
    }
    check(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "instantiate");
    graphs_[key] = {g, x};
    return x;
}

void RenderEngine::submit(const RenderTask* tasks, int n) {
    if (!dev_) throw admission::Error("ERR_CAPACITY");
    if (pending_) wait();
    if (n < 0 || n > max_tasks_) throw admission::Error("ERR_CAPACITY");
    results_.assign(n, RenderResult{});
    out_slot_.assign(n, -1);
    std::vector<jobs::S6Job> v6;
    std::vector<CopyOp> vc;
    auto* outs = at<s6::RenderStatistics>(dev_, off_out_);
    for (int i = 0; i < n; ++i) {
        const RenderTask& t = tasks[i];
        RenderResult& r = results_[i];
        if (!t.buffers) {
            r.status = Status::INVALID_ARGUMENT;
            continue;
        }
        const RenderBuffers& B = *t.buffers;
        jobs::S6Job j{};
        const Status s =
            jobs::make_s6_job(t.a, t.b, t.geometry, t.mode, B.canvas, B.mask, B.ws, B.ws_bytes, B.stats, &j);
        if (s != Status::OK) {
            r.status = s;
            continue;
        }
        out_slot_[i] = static_cast<int>(v6.size());
        vc.push_back({B.stats, &outs[v6.size()], static_cast<int>(sizeof(s6::RenderStatistics))});
        v6.push_back(j);
    }
    Header& h = *reinterpret_cast<Header*>(staging_);
    h = Header{};
    h.t6 = jobs::s6_prefix(v6.data(), static_cast<int>(v6.size()));
    h.n_s6 = static_cast<int>(v6.size());
    h.n_copy = static_cast<int>(vc.size());
    if (!v6.empty()) std::memcpy(staging_ + off_s6_, v6.data(), v6.size() * sizeof(jobs::S6Job));
    if (!vc.empty()) std::memcpy(staging_ + off_copy_, vc.data(), vc.size() * sizeof(CopyOp));
    check(hipMemcpyAsync(dev_, staging_, table_bytes_, hipMemcpyHostToDevice, stream_), "upload tables");
    check(hipEventRecord(ev_begin_, stream_), "event");
    if (h.n_s6 > 0) {
        if (use_graph_ && !probe_on_) {
            check(hipGraphLaunch(graph_for(h.t6.top, h.t6.blend_top), stream_), "graph launch");
        } else {
            jobs::LaunchProbe* probe = probe_on_ ? probe_ : nullptr;
            if (probe) probe->count = 0;
            check(jobs::launch_s6(at<jobs::S6Job>(dev_, off_s6_), h.n_s6, h.t6, nullptr, stream_, nullptr, probe),
                  "render launch");
            check(launch_copy_list(at<CopyOp>(dev_, off_copy_), nullptr, h.n_copy, stream_), "copy list");
        }
        check(hipMemcpyAsync(out_, dev_ + off_out_, v6.size() * sizeof(s6::RenderStatistics), hipMemcpyDeviceToHost,
                             stream_),
              "download results");
    }
    check(hipEventRecord(ev_end_, stream_), "event");
    pending_ = true;
}

bool RenderEngine::ready() {
    if (!pending_) return true;
    const hipError_t q = hipEventQuery(ev_end_);
    if (q == hipErrorNotReady) return false;
    check(q, "render phase");
    wait();
    return true;
}

void RenderEngine::wait() {
    if (!pending_) return;
    check(hipEventSynchronize(ev_end_), "render phase");
    pending_ = false;
    const auto* outs = reinterpret_cast<const s6::RenderStatistics*>(out_);
    for (std::size_t i = 0; i < results_.size(); ++i)
        if (out_slot_[i] >= 0) results_[i].stats = outs[out_slot_[i]];
    if (probe_on_) read_probe();
}

float RenderEngine::last_ms() const {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_begin_, ev_end_) != hipSuccess) return -1;
    return ms;
}

// ----------------------------------------------------------- batch runner
FeatureTask fixed_task(const PairSpec& p, PairBuffers* buffers) {
    FeatureTask t{};
    t.a = p.a;
    t.b = p.b;
    t.buffers = buffers;
    t.S = p.S;
    t.s2 = {p.fast_threshold, p.K, p.harris_k};
    t.s4 = {p.ratio, p.cross_check};
    t.s5 = {p.R, 0, p.msac_batch_size, p.msac_refinements, p.msac_threshold, p.seed, 1};
    return t;
}

void BatchRunner::initialize(int max_pairs, std::size_t pool_bytes, bool use_graphs, double pair_share) {
    if (!(pair_share > 0) || !(pair_share < 1)) throw admission::Error("ERR_CAPACITY");
    max_pairs_ = max_pairs;
    tuning_.graphs = use_graphs;
    const std::size_t slot_bytes = pool_bytes / 2;
    const std::size_t pair_bytes_total = align_up(static_cast<std::size_t>(slot_bytes * pair_share));
    for (Slot& s : slots_) {
        s.pair_pool.initialize(pair_bytes_total);
        s.render_pool.initialize(slot_bytes > pair_bytes_total ? slot_bytes - pair_bytes_total : kAlign);
    }
    features_.initialize(max_pairs, use_graphs);
    for (RenderEngine& r : render_) r.initialize(max_pairs, use_graphs);
}

namespace {
using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}
}  // namespace

void BatchRunner::complete(int slot) {
    Slot& s = slots_[slot];
    RenderEngine& render = render_[slot];
    if (!s.pending) return;
    render.wait();
    last_render_ms_ = s.owner.empty() ? 0.0 : render.last_ms();
    for (std::size_t k = 0; k < s.owner.size(); ++k) {
        PairOutcome& o = s.out[s.owner[k]];
        o.render = render.results()[k];
        if (o.render.status != Status::OK) {
            o.status = o.render.status;
            o.fail_stage = "s6";
        }
    }
    s.pending = false;
}

void BatchRunner::pool_use(std::size_t* capacity, std::size_t* in_use, std::size_t* peak) const {
    std::size_t c = 0, u = 0, p = 0;
    for (const Slot& s : slots_) {
        c += s.pair_pool.capacity() + s.render_pool.capacity();
        u += s.pair_pool.in_use() + s.render_pool.in_use();
        p += s.pair_pool.peak() + s.render_pool.peak();
    }
    if (capacity) *capacity = c;
    if (in_use) *in_use = u;
    if (peak) *peak = p;
}

void BatchRunner::submit(const PairSpec* pairs, int n, std::vector<PairOutcome>* previous, BatchTimings* timings) {
    if (n < 0 || n > max_pairs_) throw admission::Error("ERR_CAPACITY");
    const auto t0 = Clock::now();
    Slot& prev = slots_[cur_ ^ 1];
    Slot& s = slots_[cur_];
    complete(cur_);  // two batches back: long done
    if (previous) previous->clear();
    const bool had_prev = prev.pending;
    s.pair_pool.reset();
    s.pairs.assign(n, PairBuffers{});
    s.renders.assign(n, RenderBuffers{});
    s.out.assign(n, PairOutcome{});
    s.owner.clear();
    std::vector<FeatureTask> tasks(n);
    for (int i = 0; i < n; ++i) {
        const PairSpec& p = pairs[i];
        const PairCapacity cap{std::max(p.a.width, p.b.width), std::max(p.a.height, p.b.height), std::max(p.K, 1),
                               std::max(p.R, 1)};
        tasks[i] = fixed_task(p, nullptr);
        try {
            if (allocate_pair(s.pair_pool, cap, &s.pairs[i])) tasks[i].buffers = &s.pairs[i];
        } catch (const admission::Error&) {
        }
        s.out[i].buffers = &s.pairs[i];
    }
    // S1-S5 on the feature stream, alongside the previous batch's S6.
    features_.submit(tasks.data(), n);
    features_.wait();
    const auto t1 = Clock::now();
    s.render_pool.reset();
    std::vector<RenderTask> renders;
    for (int i = 0; i < n; ++i) {
        PairOutcome& o = s.out[i];
        o.features = features_.results()[i];
        if (!tasks[i].buffers) {
            o.status = Status::INSUFFICIENT_WORKSPACE;
            o.fail_stage = "ERR_DEVICE_OOM";
            continue;
        }
        if (o.features.status != Status::OK) {
            o.status = o.features.status;
            o.fail_stage = o.features.fail_stage;
            continue;
        }
        if (o.features.h.status != Status::OK) {
            o.status = o.features.h.status;
            o.fail_stage = "s5";
            continue;
        }
        const PairSpec& p = pairs[i];
        const Status gs = s6::prepare_geometry(p.a.width, p.a.height, p.b.width, p.b.height, o.features.h.H,
                                               s6::GeometryPolicy{}, &o.geometry);
        if (gs != Status::OK) {
            o.status = gs;
            o.fail_stage = "s6";
            continue;
        }
        bool ok = false;
        try {
            ok = allocate_render(s.render_pool, o.geometry, p.blend, &s.renders[i]);
        } catch (const admission::Error&) {
        }
        if (!ok) {
            o.status = Status::INSUFFICIENT_WORKSPACE;
            o.fail_stage = "s6";
            continue;
        }
        o.render_buffers = &s.renders[i];
        renders.push_back({p.a, p.b, o.geometry, p.blend, &s.renders[i]});
        s.owner.push_back(i);
    }
    const auto t2 = Clock::now();
    // This batch's S6 goes onto its slot's render stream before the previous
    // batch's S6 (on the other slot's stream) is waited for and read.
    render_[cur_].submit(renders.data(), static_cast<int>(renders.size()));
    s.pending = true;
    complete(cur_ ^ 1);
    if (previous && had_prev) *previous = prev.out;
    cur_ ^= 1;
    if (timings) {
        timings->feature_ms = features_.last_ms();
        timings->render_ms = last_render_ms_;  // the previous batch's S6
        timings->host_ms = ms_between(t1, t2);
        timings->total_ms = ms_between(t0, Clock::now());
    }
}

void BatchRunner::finish(std::vector<PairOutcome>* out, BatchTimings* timings) {
    Slot& last = slots_[cur_ ^ 1];
    complete(cur_ ^ 1);
    if (out) *out = last.out;
    if (timings) timings->render_ms = last_render_ms_;
}

void BatchRunner::run(const PairSpec* pairs, int n, std::vector<PairOutcome>* out, BatchTimings* timings) {
    const auto t0 = Clock::now();
    BatchTimings t{};
    if (pending()) finish(nullptr, nullptr);
    submit(pairs, n, nullptr, &t);
    finish(out, &t);
    if (timings) {
        *timings = t;
        timings->total_ms = ms_between(t0, Clock::now());
    }
}

void BatchRunner::apply(const Tuning& t) {
    jobs::s4_set_segment(t.s4_segment);
    jobs::set_workgroups_per_cu(t.workgroups_per_cu);
    for (int s = 0; s < jobs::kStages; ++s)
        if (t.stage_workgroups[s]) jobs::set_workgroups_per_cu(static_cast<jobs::Stage>(s), t.stage_workgroups[s]);
    jobs::s1_set_tile_rows(t.s1_tile_rows);
    tuning_ = t;
    tuning_.s4_segment = jobs::s4_segment();
    tuning_.workgroups_per_cu = jobs::workgroups_per_cu(jobs::Stage::S2);
    if (t.workgroups_per_cu >= 1 && t.workgroups_per_cu <= 64) tuning_.workgroups_per_cu = t.workgroups_per_cu;
    for (int s = 0; s < jobs::kStages; ++s)
        tuning_.stage_workgroups[s] = jobs::workgroups_per_cu(static_cast<jobs::Stage>(s));
    tuning_.s1_tile_rows = jobs::s1_tile_rows();
    features_.invalidate();
    features_.set_use_graph(t.graphs);
    for (RenderEngine& r : render_) {
        r.invalidate();
        r.set_use_graph(t.graphs);
    }
}

void BatchRunner::set_probe(bool on) {
    for (RenderEngine& r : render_) r.set_probe(on);
}

const std::vector<RenderEngine::ProbeStat>& BatchRunner::probe_stats() {
    probe_merged_.clear();
    for (const RenderEngine& r : render_)
        for (const RenderEngine::ProbeStat& e : r.probe_stats()) {
            auto it = std::find_if(probe_merged_.begin(), probe_merged_.end(), [&](const RenderEngine::ProbeStat& m) {
                return m.kernel == e.kernel && m.level == e.level;
            });
            if (it == probe_merged_.end()) {
                probe_merged_.push_back(e);
            } else {
                it->launches += e.launches;
                it->ms += e.ms;
                it->bytes += e.bytes;
            }
        }
    return probe_merged_;
}

namespace {

std::string device_key() {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    return std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount);
}

bool read_int(const std::string& text, const char* key, int* value) {
    const std::string k = std::string("\"") + key + "\"";
    const auto p = text.find(k);
    if (p == std::string::npos) return false;
    const auto c = text.find(':', p + k.size());
    if (c == std::string::npos) return false;
    return std::sscanf(text.c_str() + c + 1, "%d", value) == 1;
}

bool load_cache(const std::string& path, const std::string& key, const std::string& shape, Tuning* t) {
    std::ifstream f(path);
    if (!f) return false;
    const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const std::string want = "\"" + key + "|" + shape + "\"";
    const auto p = text.find(want);
    if (p == std::string::npos) return false;
    const auto end = text.find('}', p);
    const std::string entry = text.substr(p, end == std::string::npos ? std::string::npos : end - p);
    int seg = 0, wpc = 0, graphs = 1, rows = 64;
    if (!read_int(entry, "s4_segment", &seg) || !read_int(entry, "workgroups_per_cu", &wpc)) return false;
    read_int(entry, "graphs", &graphs);
    read_int(entry, "s1_tile_rows", &rows);
    int stages[jobs::kStages] = {};
    const auto list = entry.find("\"stage_workgroups_per_cu\"");
    if (list != std::string::npos) {
        const auto open = entry.find('[', list);
        if (open == std::string::npos ||
            std::sscanf(entry.c_str() + open, "[%d,%d,%d,%d,%d,%d]", &stages[0], &stages[1], &stages[2], &stages[3],
                        &stages[4], &stages[5]) != jobs::kStages)
            return false;
    }
    t->s4_segment = seg;
    t->workgroups_per_cu = wpc;
    t->graphs = graphs != 0;
    t->s1_tile_rows = rows;
    std::copy(stages, stages + jobs::kStages, t->stage_workgroups);
    return true;
}

void store_cache(const std::string& path, const std::string& key, const std::string& shape, const Tuning& t) {
    std::string text;
    {
        std::ifstream f(path);
        if (f) text.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    }
    std::ostringstream entry;
    entry << "\"" << key << "|" << shape << "\": {\"s4_segment\": " << t.s4_segment
          << ", \"workgroups_per_cu\": " << t.workgroups_per_cu << ", \"graphs\": " << (t.graphs ? 1 : 0)
          << ", \"s1_tile_rows\": " << t.s1_tile_rows << ", \"stage_workgroups_per_cu\": [";
    for (int s = 0; s < jobs::kStages; ++s) entry << (s ? "," : "") << t.stage_workgroups[s];
    entry << "]}";
    const std::string want = "\"" + key + "|" + shape + "\"";
    const auto p = text.find(want);
    if (p != std::string::npos) {
        const auto end = text.find('}', p);
        text.replace(p, end - p + 1, entry.str());
    } else {
        const auto close = text.rfind('}');
        if (close == std::string::npos) text = "{\n  " + entry.str() + "\n}\n";
        else text.insert(close, (text.find(':') != std::string::npos ? ",\n  " : "  ") + entry.str() + "\n");
    }
    std::ofstream f(path);
    f << text;
}

// Shape class of a batch for the cache: pairs and the log2 of the mean pixel count and K.
std::string shape_key(const PairSpec* pairs, int n) {
    double px = 0, k = 0;
    for (int i = 0; i < n; ++i) {
        px += static_cast<double>(pairs[i].a.width) * pairs[i].a.height;
        k += pairs[i].K;
    }
    const int lp = n ? static_cast<int>(std::log2(px / n + 1)) : 0;
    const int lk = n ? static_cast<int>(std::log2(k / n + 1)) : 0;
    return "n" + std::to_string(n) + "_px" + std::to_string(lp) + "_k" + std::to_string(lk);
}

}  // namespace

Tuning BatchRunner::autotune(const PairSpec* pairs, int n, const std::string& cache_path) {
    const std::string key = device_key(), shape = shape_key(pairs, n);
    Tuning best = tuning_;
    if (!cache_path.empty() && load_cache(cache_path, key, shape, &best)) {
        apply(best);
        return tuning_;
    }
    std::vector<PairOutcome> out;
    auto measure = [&](const Tuning& t, bool features_only) {
        apply(t);
        BatchTimings tm{};
        double bestms = 1e30;
        for (int rep = 0; rep < 3; ++rep) {
            run(pairs, n, &out, &tm);
            if (rep > 0) bestms = std::min(bestms, features_only ? tm.feature_ms : tm.feature_ms + tm.render_ms);
        }
        return bestms;
    };
    // Occupancy of every stage together first, then each stage's own
    // occupancy in turn with the others held (S1-S5 judged by the feature
    // phase, S6 by the render phase), then the S1 tile height and the S4
    // train segment. Each step keeps the current choice unless another one
    // is faster.
    double best_ms = 1e30;
    for (int s = 0; s < jobs::kStages; ++s) best.stage_workgroups[s] = 0;
    for (int wpc : {8, 16, 24, 32}) {
        Tuning t = best;
        t.workgroups_per_cu = wpc;
        const double ms = measure(t, false);
        if (ms < best_ms) {
            best_ms = ms;
            best.workgroups_per_cu = wpc;
        }
    }
    for (int s = 0; s < jobs::kStages; ++s) best.stage_workgroups[s] = best.workgroups_per_cu;
    const auto phase_ms = [&](const Tuning& t, bool render) {
        apply(t);
        BatchTimings tm{};
        double lowest = 1e30;
        for (int rep = 0; rep < 3; ++rep) {
            run(pairs, n, &out, &tm);
            if (rep > 0) lowest = std::min(lowest, render ? tm.render_ms : tm.feature_ms);
        }
        return lowest;
    };
    for (int s = 0; s < jobs::kStages; ++s) {
        const bool render = s == static_cast<int>(jobs::Stage::S6);
        double stage_best = phase_ms(best, render);
        for (int wpc : {4, 8, 16, 24, 32}) {
            if (wpc == best.stage_workgroups[s]) continue;
            Tuning t = best;
            t.stage_workgroups[s] = wpc;
            const double ms = phase_ms(t, render);
            if (ms < stage_best) {
                stage_best = ms;
                best.stage_workgroups[s] = wpc;
            }
        }
    }
    {
        double rows_best = phase_ms(best, false);
        for (int rows : {16, 32, 64}) {
            if (rows == best.s1_tile_rows) continue;
            Tuning t = best;
            t.s1_tile_rows = rows;
            const double ms = phase_ms(t, false);
            if (ms < rows_best) {
                rows_best = ms;
                best.s1_tile_rows = rows;
            }
        }
    }
    best_ms = 1e30;
    for (int seg : {64, 128, 256, 512}) {
        Tuning t = best;
        t.s4_segment = seg;
        const double ms = measure(t, true);
        if (ms < best_ms) {
            best_ms = ms;
            best.s4_segment = seg;
        }
    }
    apply(best);
    if (!cache_path.empty()) store_cache(cache_path, key, shape, tuning_);
    return tuning_;
}

}  // namespace stitch::batch
