// S1: RGB8 -> GRAY8 with OpenCV's 15-bit fixed-point luma and INTER_AREA
// resize, for any number of images in one persistent launch, plus the
// controller probe's mean absolute gradient.
// Semantics: src/s1/s1.cu:31-103 (kernel), :105-134 (mode and parameters),
// :140-203 (gradient), :209-354 (API).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "stitch/jobs.hpp"

namespace stitch {
namespace {

using jobs::S1Job;
using jobs::kS1TileH;
using jobs::kS1TileW;

__device__ __forceinline__ int rgb_to_gray(int r, int g, int b) {
    return (9798 * r + 19235 * g + 3735 * b + 16384) >> 15;
}

__device__ __forceinline__ int gray_at(const std::uint8_t* p) { return rgb_to_gray(p[0], p[1], p[2]); }

__device__ std::uint8_t s1_pixel(const S1Job& p, int xd, int yd) {
    if (p.mode == 0) return static_cast<std::uint8_t>(gray_at(p.src + static_cast<size_t>(yd) * p.src_pitch + 3 * xd));
    if (p.mode == 1) {
        const std::uint8_t* r0 = p.src + static_cast<size_t>(2 * yd) * p.src_pitch + 6 * xd;
        const std::uint8_t* r1 = r0 + p.src_pitch;
        return static_cast<std::uint8_t>((gray_at(r0) + gray_at(r0 + 3) + gray_at(r1) + gray_at(r1 + 3) + 2) >> 2);
    }
    if (p.mode == 2) {
        int sum = 0;
        for (int dy = 0; dy < 4; ++dy) {
            const std::uint8_t* r = p.src + static_cast<size_t>(4 * yd + dy) * p.src_pitch + 12 * xd;
            for (int dx = 0; dx < 4; ++dx) sum += gray_at(r + 3 * dx);
        }
        return static_cast<std::uint8_t>(static_cast<int>(rintf(static_cast<float>(sum) * 0.0625f)));
    }
    const float fsy1 = static_cast<float>(yd) * p.inv_sy, fsy2 = static_cast<float>(yd + 1) * p.inv_sy;
    const float fsx1 = static_cast<float>(xd) * p.inv_sx, fsx2 = static_cast<float>(xd + 1) * p.inv_sx;
    const int y_start = static_cast<int>(floorf(fsy1)), y_end = static_cast<int>(ceilf(fsy2));
    const int x_start = static_cast<int>(floorf(fsx1)), x_end = static_cast<int>(ceilf(fsx2));
    float total = 0.0f;
    for (int ys = y_start; ys < y_end && ys < p.orig_h; ++ys) {
        const float wy = fmaxf(0.0f, fminf(static_cast<float>(ys + 1), fsy2) - fmaxf(static_cast<float>(ys), fsy1));
        if (wy <= 0.0f) continue;
        const std::uint8_t* row = p.src + static_cast<size_t>(ys) * p.src_pitch;
        for (int xs = x_start; xs < x_end && xs < p.orig_w; ++xs) {
            const float wx =
                fmaxf(0.0f, fminf(static_cast<float>(xs + 1), fsx2) - fmaxf(static_cast<float>(xs), fsx1));
            if (wx <= 0.0f) continue;
            total += static_cast<float>(gray_at(row + 3 * xs)) * (wx * wy);
        }
    }
    return static_cast<std::uint8_t>(static_cast<int>(rintf(total * p.scale_area)));
}

// Persistent walk over 64 x kRows output tiles of every job (kRows 16, 32
// or 64, a launch-shape knob): the word path reads four pixels per thread
// and row step, the scalar path kRows / 4 rows per thread.
constexpr int kRowsPerThread = kS1TileH / (256 / kS1TileW);  // gradient tiles (64 x 64)
static_assert(kRowsPerThread * (256 / kS1TileW) == kS1TileH, "S1 tile");

// Four horizontally adjacent pixels of a full-size (mode 0) job from three
// dword loads and one dword store, for rows whose addresses are 4-aligned.
__device__ __forceinline__ unsigned gray4(unsigned w0, unsigned w1, unsigned w2) {
    const auto b = [](unsigned w, int k) { return static_cast<int>((w >> (8 * k)) & 0xffu); };
    const unsigned g0 = static_cast<unsigned>(rgb_to_gray(b(w0, 0), b(w0, 1), b(w0, 2)));
    const unsigned g1 = static_cast<unsigned>(rgb_to_gray(b(w0, 3), b(w1, 0), b(w1, 1)));
    const unsigned g2 = static_cast<unsigned>(rgb_to_gray(b(w1, 2), b(w1, 3), b(w2, 0)));
    const unsigned g3 = static_cast<unsigned>(rgb_to_gray(b(w2, 1), b(w2, 2), b(w2, 3)));
    return g0 | (g1 << 8) | (g2 << 16) | (g3 << 24);
}

template <int kTileH>
__device__ __forceinline__ void s1_tile_vec4(const S1Job& p, int tx, int ty) {
    constexpr int kQuads = kS1TileW / 4, kRows = 256 / kQuads, kIt = kTileH / kRows;
    static_assert(kIt * kRows == kTileH, "S1 word-path tile");
    const int xd = tx * kS1TileW + 4 * static_cast<int>(threadIdx.x % kQuads);
    unsigned w[kIt][3];
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
        const int yd = ty * kTileH + static_cast<int>(threadIdx.x / kQuads) + k * kRows;
        if (yd < p.work_h && xd + 3 < p.work_w) {
            const auto* s = reinterpret_cast<const unsigned*>(p.src + static_cast<size_t>(yd) * p.src_pitch + 3 * xd);
            w[k][0] = s[0];
            w[k][1] = s[1];
            w[k][2] = s[2];
        }
    }
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
        const int yd = ty * kTileH + static_cast<int>(threadIdx.x / kQuads) + k * kRows;
        if (yd >= p.work_h || xd >= p.work_w) continue;
        std::uint8_t* d = p.dst + static_cast<size_t>(yd) * p.dst_pitch + xd;
        if (xd + 3 < p.work_w) {
            *reinterpret_cast<unsigned*>(d) = gray4(w[k][0], w[k][1], w[k][2]);
        } else {
            for (int i = 0; xd + i < p.work_w; ++i)
                d[i] = static_cast<std::uint8_t>(gray_at(p.src + static_cast<size_t>(yd) * p.src_pitch + 3 * (xd + i)));
        }
    }
}

template <int kTileH>
__global__ void __launch_bounds__(256) s1_kernel(dev::JobSpan<S1Job> span, dev::Total work) {
    constexpr int kScalarRows = kTileH / (256 / kS1TileW);
    const long long total = work.get();
    int j = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk tw = dev::tile_walk(total); tw.t < tw.end; tw.t += tw.step) {
        const long long t = tw.t;
        j = dev::next_job(span, t, &S1Job::tile_begin, j, njobs);
        const S1Job& p = dev::job_at(span, j);
        const unsigned local = static_cast<unsigned>(t - p.tile_begin);
        const unsigned tiles_x = static_cast<unsigned>(p.tiles_x);
        const int tx = static_cast<int>(local % tiles_x), ty = static_cast<int>(local / tiles_x);
        if (p.mode == 0 && ((reinterpret_cast<std::uintptr_t>(p.src) | reinterpret_cast<std::uintptr_t>(p.dst) |
                             static_cast<std::uintptr_t>(p.src_pitch) | static_cast<std::uintptr_t>(p.dst_pitch)) &
                            3u) == 0) {
            s1_tile_vec4<kTileH>(p, tx, ty);
            continue;
        }
        const int xd = tx * kS1TileW + static_cast<int>(threadIdx.x % kS1TileW);
        std::uint8_t v[kScalarRows];
#pragma unroll
        for (int k = 0; k < kScalarRows; ++k) {
            const int yd = ty * kTileH + static_cast<int>(threadIdx.x / kS1TileW) + k * (256 / kS1TileW);
            v[k] = xd < p.work_w && yd < p.work_h ? s1_pixel(p, xd, yd) : 0;
        }
#pragma unroll
        for (int k = 0; k < kScalarRows; ++k) {
            const int yd = ty * kTileH + static_cast<int>(threadIdx.x / kS1TileW) + k * (256 / kS1TileW);
            if (xd < p.work_w && yd < p.work_h) p.dst[static_cast<size_t>(yd) * p.dst_pitch + xd] = v[k];
        }
    }
}

using jobs::GradJob;

// Sum of |I(x+1,y)-I(x,y)| + |I(x,y+1)-I(x,y)|: wave sums, one atomic per wave.
__global__ void __launch_bounds__(256) gradient_kernel(dev::JobSpan<GradJob> span, dev::Total work) {
    const long long total = work.get();
    int j = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk tw = dev::tile_walk(total); tw.t < tw.end; tw.t += tw.step) {
        const long long t = tw.t;
        j = dev::next_job(span, t, &GradJob::tile_begin, j, njobs);
        const GradJob& g = dev::job_at(span, j);
        const unsigned local = static_cast<unsigned>(t - g.tile_begin);
        const unsigned tiles_x = static_cast<unsigned>(g.tiles_x);
        const int x = static_cast<int>(local % tiles_x) * kS1TileW + static_cast<int>(threadIdx.x % kS1TileW);
        unsigned diff = 0;
#pragma unroll
        for (int k = 0; k < kRowsPerThread; ++k) {
            const int y = static_cast<int>(local / tiles_x) * kS1TileH + static_cast<int>(threadIdx.x / kS1TileW) +
                          k * (256 / kS1TileW);
            if (x < g.w && y < g.h) {
                const std::uint8_t* row = g.src + static_cast<size_t>(y) * g.pitch;
                const int c = row[x];
                if (x + 1 < g.w) diff += static_cast<unsigned>(abs(static_cast<int>(row[x + 1]) - c));
                if (y + 1 < g.h) diff += static_cast<unsigned>(abs(static_cast<int>(row[x + g.pitch]) - c));
            }
        }
        diff = dev::wave_sum(diff);
        if (dev::lane_id() == 0 && diff) atomicAdd(g.accum, static_cast<unsigned long long>(diff));
    }
}

// Zeroes the accumulators (a fill kernel rather than a memset node, so the
// sequence stays safe to capture into graphs).
__global__ void gradient_clear(unsigned long long* accum, int n) {
    if (threadIdx.x < n) accum[threadIdx.x] = 0;
}

// Per-job forms of the two helpers for job tables.
__global__ void gradient_clear_jobs(dev::JobSpan<GradJob> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x)
        if (threadIdx.x == 0) *dev::job_at(span, ji).accum = 0;
}

__global__ void gradient_normalize_jobs(dev::JobSpan<GradJob> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        const GradJob& g = dev::job_at(span, ji);
        if (threadIdx.x != 0) continue;
        const unsigned long long edges =
            static_cast<unsigned long long>(g.h) * (g.w - 1) + static_cast<unsigned long long>(g.w) * (g.h - 1);
        *g.out = edges == 0 ? 0.0 : static_cast<double>(*g.accum) / (255.0 * static_cast<double>(edges));
    }
}

__global__ void gradient_normalize(const unsigned long long* accum, int w0, int h0, int w1, int h1, int n,
                                   double* out) {
    const int z = threadIdx.x;
    if (z >= n) return;
    const int w = z ? w1 : w0, h = z ? h1 : h0;
    const unsigned long long edges =
        static_cast<unsigned long long>(h) * (w - 1) + static_cast<unsigned long long>(w) * (h - 1);
    out[z] = edges == 0 ? 0.0 : static_cast<double>(accum[z]) / (255.0 * static_cast<double>(edges));
}

int determine_mode(int ow, int oh, int ww, int wh) {
    if (ow == ww && oh == wh) return 0;
    if (ow == 2 * ww && oh == 2 * wh) return 1;
    if (ow == 4 * ww && oh == 4 * wh) return 2;
    return 3;
}

Status launch_status() { return hipGetLastError() == hipSuccess ? Status::OK : Status::CUDA_ERROR; }

}  // namespace

namespace jobs {

Status make_s1_job(const ImageView& in, const ImageView& out, const ScaleInfo& si, S1Job* p) {
    if (in.data == nullptr || out.data == nullptr || in.format != PixelFormat::RGB8 ||
        out.format != PixelFormat::GRAY8 || in.width <= 0 || in.height <= 0 || out.width != si.work_w ||
        out.height != si.work_h || in.pitch_bytes < in.width * 3 || out.pitch_bytes < out.width)
        return Status::INVALID_ARGUMENT;
    p->src = in.data;
    p->src_pitch = in.pitch_bytes;
    p->orig_w = in.width;
    p->orig_h = in.height;
    p->dst = out.data;
    p->dst_pitch = out.pitch_bytes;
    p->work_w = out.width;
    p->work_h = out.height;
    p->inv_sx = static_cast<float>(in.width) / static_cast<float>(out.width);
    p->inv_sy = static_cast<float>(in.height) / static_cast<float>(out.height);
    p->scale_area = 1.0f / (p->inv_sx * p->inv_sy);
    p->mode = determine_mode(in.width, in.height, out.width, out.height);
    p->tiles_x = (out.width + kS1TileW - 1) / kS1TileW;
    p->tile_h = s1_tile_rows();
    p->tile_begin = 0;
    return Status::OK;
}

hipError_t launch_s1(const S1Job* table, int count, long long total, const S1Job* single, hipStream_t stream,
                     const dev::Replay<long long>* replay) {
    if (total <= 0) return hipSuccess;
    dev::JobSpan<S1Job> span{table, count, single ? *single : S1Job{}, replay ? replay->count : nullptr};
    // Jobs take their tile height from the knob when they are made, and the
    // engines make them just before launching.
    const unsigned grid = persistent_grid(total, jobs::Stage::S1);
    const dev::Total work(total, replay ? replay->totals : nullptr);
    switch (s1_tile_rows()) {
        case 16: s1_kernel<16><<<grid, 256, 0, stream>>>(span, work); break;
        case 32: s1_kernel<32><<<grid, 256, 0, stream>>>(span, work); break;
        default: s1_kernel<64><<<grid, 256, 0, stream>>>(span, work); break;
    }
    return hipGetLastError();
}

hipError_t launch_gradient(const GradJob* table, int count, long long total, const GradJob* single,
                           hipStream_t stream, const dev::Replay<long long>* replay) {
    if (total <= 0) return hipSuccess;
    dev::JobSpan<GradJob> span{table, count, single ? *single : GradJob{}, replay ? replay->count : nullptr};
    gradient_clear_jobs<<<count, 64, 0, stream>>>(span);
    gradient_kernel<<<persistent_grid(total, jobs::Stage::S1), 256, 0, stream>>>(span, dev::Total(total, replay ? replay->totals : nullptr));
    gradient_normalize_jobs<<<count, 64, 0, stream>>>(span);
    return hipGetLastError();
}

}  // namespace jobs

namespace s1 {

Status s1_workspace_bytes(int orig_w, int orig_h, double scale, std::size_t* bytes) {
    if (bytes == nullptr || orig_w <= 0 || orig_h <= 0 || !(scale > 0.0) || scale > 1.0)
        return Status::INVALID_ARGUMENT;
    *bytes = 0;
    return Status::OK;
}

Status s1_preprocess_single(const ImageView& in, double scale, ImageView* out, ScaleInfo* si, hipStream_t stream) {
    if (out == nullptr || si == nullptr || !(scale > 0.0) || scale > 1.0) return Status::INVALID_ARGUMENT;
    Status s = make_scale_info(in.width, in.height, scale, si);
    if (s != Status::OK) return s;
    jobs::S1Job job{};
    if ((s = jobs::make_s1_job(in, *out, *si, &job)) != Status::OK) return s;
    return jobs::launch_s1(nullptr, 1, jobs::s1_tiles(job), &job, stream) == hipSuccess ? Status::OK
                                                                                       : Status::CUDA_ERROR;
}

Status s1_preprocess(const ImageView& a, const ImageView& b, double scale, ImageView* out_a, ImageView* out_b,
                     ScaleInfo* sa, ScaleInfo* sb, hipStream_t stream) {
    if (out_a == nullptr || out_b == nullptr || sa == nullptr || sb == nullptr || !(scale > 0.0) || scale > 1.0)
        return Status::INVALID_ARGUMENT;
    Status s = make_scale_info(a.width, a.height, scale, sa);
    if (s != Status::OK) return s;
    if ((s = make_scale_info(b.width, b.height, scale, sb)) != Status::OK) return s;
    jobs::S1Job ja{}, jb{};
    if ((s = jobs::make_s1_job(a, *out_a, *sa, &ja)) != Status::OK) return s;
    if ((s = jobs::make_s1_job(b, *out_b, *sb, &jb)) != Status::OK) return s;
    // Two single-job launches keep the call allocation- and upload-free; the
    // batched form covers many images per launch through a device table.
    if (jobs::launch_s1(nullptr, 1, jobs::s1_tiles(ja), &ja, stream) != hipSuccess) return Status::CUDA_ERROR;
    if (jobs::launch_s1(nullptr, 1, jobs::s1_tiles(jb), &jb, stream) != hipSuccess) return Status::CUDA_ERROR;
    return Status::OK;
}

Status s1_gradient_workspace_bytes(std::size_t* bytes) {
    if (bytes == nullptr) return Status::INVALID_ARGUMENT;
    *bytes = 32;
    return Status::OK;
}

namespace {
Status gradient(const ImageView* views, int n, void* ws, double* const* pinned, hipStream_t stream) {
    auto* accum = static_cast<unsigned long long*>(ws);
    auto* out = reinterpret_cast<double*>(static_cast<std::uint8_t*>(ws) + 16);
    gradient_clear<<<1, 64, 0, stream>>>(accum, n);
    for (int i = 0; i < n; ++i) {
        GradJob g{views[i].data, views[i].pitch_bytes, views[i].width, views[i].height, accum + i, out + i,
                  (views[i].width + kS1TileW - 1) / kS1TileW, 0};
        const long long tiles = static_cast<long long>(g.tiles_x) * ((g.h + kS1TileH - 1) / kS1TileH);
        gradient_kernel<<<persistent_grid(tiles, jobs::Stage::S1), 256, 0, stream>>>(dev::JobSpan<GradJob>{nullptr, 1, g}, tiles);
    }
    gradient_normalize<<<1, 64, 0, stream>>>(accum, views[0].width, views[0].height, views[n - 1].width,
                                             views[n - 1].height, n, out);
    for (int i = 0; i < n; ++i)
        if (hipMemcpyAsync(pinned[i], out + i, sizeof(double), hipMemcpyDeviceToHost, stream) != hipSuccess)
            return Status::CUDA_ERROR;
    return launch_status();
}
bool valid_gray(const ImageView& g) {
    return g.data && g.format == PixelFormat::GRAY8 && g.width >= 2 && g.height >= 2 && g.pitch_bytes >= g.width;
}
}  // namespace

Status s1_compute_gradient(const ImageView& a, const ImageView& b, void* ws, double* pa, double* pb,
                           hipStream_t stream) {
    if (!valid_gray(a) || !valid_gray(b) || ws == nullptr || pa == nullptr || pb == nullptr)
        return Status::INVALID_ARGUMENT;
    const ImageView views[2] = {a, b};
    double* const pinned[2] = {pa, pb};
    return gradient(views, 2, ws, pinned, stream);
}

Status s1_compute_gradient_single(const ImageView& g, void* ws, double* p, hipStream_t stream) {
    if (!valid_gray(g) || ws == nullptr || p == nullptr) return Status::INVALID_ARGUMENT;
    double* const pinned[1] = {p};
    return gradient(&g, 1, ws, pinned, stream);
}

}  // namespace s1
}  // namespace stitch
