// S3: oriented rBRIEF (OpenCV ORB, WTA_K = 2, 31x31 patch) for any number of
// images per launch. FP32 7-tap sigma-2 Gaussian (reflect-101, no
// contraction) on LDS tiles; then one wave64 per keypoint: intensity-centroid
// moments on the unblurred image, cv::fastAtan2 in FP32, cos/sin in FP64, and
// 256 rotated pair tests packed with four wave ballots.
// Semantics: src/s2_s3/s3.cu:26-196 and tools/s2_s3/reference.py:152-284.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>

#include "orb_pairs.h"
#include "stitch/jobs.hpp"

namespace stitch {
namespace {

using jobs::S3Job;
using jobs::kS3KeypointsPerBlock;
using jobs::kS3MarkBlock;
using jobs::kS3TileH;
using jobs::kS3TileW;

constexpr int kEdge = s2::kEdge;
constexpr int kR = 3;  // blur radius
constexpr int kInW = kS3TileW + 2 * kR, kInH = kS3TileH + 2 * kR;

// float32 taps of getGaussianKernel(7, 2, CV_32F).
__constant__ float kTaps[7] = {0x1.1f5f62p-4f, 0x1.0c70fcp-3f, 0x1.869472p-3f, 0x1.ba95c0p-3f,
                               0x1.869472p-3f, 0x1.0c70fcp-3f, 0x1.1f5f62p-4f};
// Row half-widths of the radius-15 orientation patch.
__constant__ int kUMax[16] = {15, 15, 15, 15, 14, 14, 14, 13, 13, 12, 11, 10, 9, 8, 6, 3};
__constant__ signed char kPairs[CS_ORB_PAIR_COUNT * 4] = CS_ORB_PAIRS_INIT;

__device__ int reflect101(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i : 2 * n - 2 - i;
    return i;
}

// Rounded keypoint position; false when the keypoint lies within kEdge of a
// border, where its descriptor row is zero and nothing is sampled.
__device__ __forceinline__ bool keypoint_centre(const S3Job& j, int index, int& x, int& y) {
    x = static_cast<int>(rintf(j.kx[index]));
    y = static_cast<int>(rintf(j.ky[index]));
    return x >= kEdge && x < j.w - kEdge && y >= kEdge && y < j.h - kEdge;
}

// Farthest blurred pixel a descriptor reads from its centre, per axis: the
// pattern's points lie within radius 13*sqrt(2) = 18.38 (the corner
// (-13, -13)), so a rotated offset rint(p*cos - q*sin) is at most 18.
constexpr int kSampleReach = 18;
static_assert(kEdge > kSampleReach, "samples stay inside the image");

// Marks every blur tile that some descriptor of the job samples: the tiles
// meeting the (2*18+1)^2 square around each keypoint kept by describe.
__global__ void __launch_bounds__(kS3MarkBlock) s3_mark(dev::JobSpan<S3Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    for (dev::TileWalk tw = dev::tile_walk(total); tw.t < tw.end; tw.t += tw.step) {
        const long long t = tw.t;
        ji = dev::next_job(span, t, &S3Job::mark_begin, ji, njobs);
        const S3Job& j = dev::job_at(span, ji);
        const int index = static_cast<int>(t - j.mark_begin) * kS3MarkBlock + static_cast<int>(threadIdx.x);
        int x, y;
        if (index >= j.capacity || index >= *j.d_count || !keypoint_centre(j, index, x, y)) continue;
        const int tx0 = (x - kSampleReach) / kS3TileW, tx1 = (x + kSampleReach) / kS3TileW;
        const int ty0 = (y - kSampleReach) / kS3TileH, ty1 = (y + kSampleReach) / kS3TileH;
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) j.marks[ty * j.tiles_x + tx] = j.epoch;
    }
}

// Row pass then column pass, taps accumulated 0..6, each product and sum
// rounded separately; round half to even, clamp to [0, 255]. Only the tiles
// marked for this job are computed: the rest of the blurred image is never
// read.
__global__ void __launch_bounds__(256) s3_blur(dev::JobSpan<S3Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ float in[kInH * kInW];
    __shared__ float rows[kInH * kS3TileW];
    const int tid = threadIdx.x;
    int ji = 0;
    const int njobs = span.jobs();
    const long long stride = gridDim.x;
    // The workgroup's tiles are blockIdx.x + k * gridDim.x. Most are not
    // marked, so their marks are read 64 at a time: lane k of every wave
    // looks up tile base + k * gridDim.x (its job by a binary search from
    // the current one) and the ballot lists the marked ones, the same list
    // in every wave. (Reading one mark per tile in turn would leave a
    // memory round trip per skipped tile.)
    for (long long base = blockIdx.x; base < total; base += 64 * stride) {
        const long long tk = base + dev::lane_id() * stride;
        bool marked = false;
        if (tk < total) {
            int jk = ji;
            if (span.table) {
                int hi = njobs - 1;
                while (jk < hi) {
                    const int mid = (jk + hi + 1) >> 1;
                    if (span.table[mid].blur_begin <= tk) jk = mid;
                    else hi = mid - 1;
                }
            }
            const S3Job& jj = span.table ? span.table[jk] : span.single;
            using GlobalU32 = const __attribute__((address_space(1))) unsigned;
            marked = ((GlobalU32*)jj.marks)[tk - jj.blur_begin] == jj.epoch;
        }
        for (unsigned long long todo = __ballot(marked); todo; todo &= todo - 1) {
        const long long t = base + static_cast<long long>(__ffsll(static_cast<long long>(todo)) - 1) * stride;
        ji = dev::next_job(span, t, &S3Job::blur_begin, ji, njobs);
        const S3Job& j = dev::job_at(span, ji);
        const unsigned local = static_cast<unsigned>(t - j.blur_begin), tiles_x = static_cast<unsigned>(j.tiles_x);
        const int x0 = static_cast<int>(local % tiles_x) * kS3TileW;
        const int y0 = static_cast<int>(local / tiles_x) * kS3TileH;
        if (((reinterpret_cast<std::uintptr_t>(j.img) | static_cast<std::uintptr_t>(j.pitch)) & 3u) == 0) {
            // Aligned words over columns x0 - 4 .. x0 + kS3TileW + 3; words
            // that would need a column reflected go byte by byte.
            constexpr int kWords = (kS3TileW + 8) / 4;
            static_assert(kS3TileW % 4 == 0 && kR == 3, "blur tile words");
            for (int i = tid; i < kInH * kWords; i += 256) {
                const int r = i / kWords, k = i % kWords;
                const int gx = x0 - 4 + 4 * k;
                const int gy = reflect101(y0 - kR + r, j.h);
                const std::uint8_t* row = j.img + static_cast<size_t>(gy) * j.pitch;
                unsigned word = 0;
                if (gx >= 0 && gx + 3 < j.w) {
                    word = *reinterpret_cast<const unsigned*>(row + gx);
                } else {
                    for (int b = 0; b < 4; ++b) word |= static_cast<unsigned>(row[reflect101(gx + b, j.w)]) << (8 * b);
                }
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int col = 4 * k - 1 + b;  // column in the tile (x0 - kR = column 0)
                    if (col >= 0 && col < kInW) in[r * kInW + col] = static_cast<float>((word >> (8 * b)) & 0xffu);
                }
            }
        } else {
            for (int i = tid; i < kInH * kInW; i += 256) {
                const int gx = reflect101(x0 - kR + i % kInW, j.w);
                const int gy = reflect101(y0 - kR + i / kInW, j.h);
                in[i] = static_cast<float>(j.img[static_cast<size_t>(gy) * j.pitch + gx]);
            }
        }
        __syncthreads();
        for (int i = tid; i < kInH * kS3TileW; i += 256) {
            const int rx = i % kS3TileW, ry = i / kS3TileW;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 7; ++k) acc = acc + kTaps[k] * in[ry * kInW + rx + k];
            rows[i] = acc;
        }
        __syncthreads();
        for (int i = tid; i < kS3TileW * kS3TileH; i += 256) {
            const int lx = i % kS3TileW, ly = i / kS3TileW;
            const int gx = x0 + lx, gy = y0 + ly;
            if (gx >= j.w || gy >= j.h) continue;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 7; ++k) acc = acc + kTaps[k] * rows[(ly + k) * kS3TileW + lx];
            const int v = static_cast<int>(rintf(acc));
            j.blurred[static_cast<size_t>(gy) * j.w + gx] = static_cast<std::uint8_t>(min(max(v, 0), 255));
        }
        __syncthreads();
        }
    }
}

// cv::fastAtan2 in degrees [0, 360), float32 step by step.
__device__ float fast_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const bool use_x = ax >= ay;
    const float c = use_x ? ay / (ax + 0x1p-52f) : ax / (ay + 0x1p-52f);
    const float c2 = c * c;
    float p = -0x1.4515b2p+1f * c2 + 0x1.1d3f7ep+3f;
    p = p * c2 + -0x1.2aaddcp+4f;
    p = p * c2 + 0x1.ca44dep+5f;
    p = p * c;
    float a = use_x ? p : 90.0f - p;
    if (x < 0) a = 180.0f - a;
    if (y < 0) a = 360.0f - a;
    return a;
}

__device__ __forceinline__ int sample(const std::uint8_t* centre, int width, int px, int py, float a, float b) {
    const float fx = static_cast<float>(px), fy = static_cast<float>(py);
    const int ix = static_cast<int>(rintf(fx * a - fy * b));
    const int iy = static_cast<int>(rintf(fx * b + fy * a));
    return centre[iy * width + ix];
}

// One wave per keypoint. Lanes 0..61 split the 31 patch rows in halves for the
// moments (integer sums, order-free); lane l then evaluates pairs l, 64+l,
// 128+l, 192+l and each ballot yields descriptor bytes 8r..8r+7.
__global__ void __launch_bounds__(64 * kS3KeypointsPerBlock) s3_describe(dev::JobSpan<S3Job> span, dev::Total work) {
    const long long total = work.get();
    const int lane = dev::lane_id();
    const int wave = threadIdx.x / 64;
    int ji = 0;
    const int njobs = span.jobs();
    // Tiles are sized by the keypoint capacity; only those below the
    // detected count hold keypoints.
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S3Job& jj = dev::lane_job(span, tk, &S3Job::kp_begin, ji, njobs);
            const long long index0 = (tk - jj.kp_begin) * kS3KeypointsPerBlock;
            return index0 < jj.capacity && index0 < *jj.d_count;
        },
        [&](long long t) {
        ji = dev::next_job(span, t, &S3Job::kp_begin, ji, njobs);
        const S3Job& j = dev::job_at(span, ji);
        const int index = static_cast<int>(t - j.kp_begin) * kS3KeypointsPerBlock + wave;
        if (index >= j.capacity || index >= *j.d_count) return;
        std::uint8_t* row = j.desc + static_cast<size_t>(index) * j.desc_stride;
        int x, y;
        if (!keypoint_centre(j, index, x, y)) {
            if (lane < 32) row[lane] = 0;
            return;
        }
        const std::uint8_t* centre = j.img + static_cast<size_t>(y) * j.pitch + x;
        int m10 = 0, m01 = 0;
        if (lane < 62) {
            const int v = (lane >> 1) - 15;
            const int av = v < 0 ? -v : v;
            const int d = kUMax[av];
            const int u0 = (lane & 1) ? 0 : -d, u1 = (lane & 1) ? d : -1;
            const std::uint8_t* r = centre + v * j.pitch;
            int sum = 0, msum = 0;
            for (int u = u0; u <= u1; ++u) {
                const int p = r[u];
                sum += p;
                msum += u * p;
            }
            m10 = msum;
            m01 = v * sum;
        }
        m10 = dev::wave_sum(m10);
        m01 = dev::wave_sum(m01);
        const float angle = fast_atan2(static_cast<float>(m01), static_cast<float>(m10)) * 0x1.1df46ap-6f;
        const float a = static_cast<float>(cos(static_cast<double>(angle)));
        const float b = static_cast<float>(sin(static_cast<double>(angle)));
        const std::uint8_t* smooth = j.blurred + static_cast<size_t>(y) * j.w + x;
        unsigned long long bits[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const signed char* p = &kPairs[(64 * r + lane) * 4];
            const int t0 = sample(smooth, j.w, p[0], p[1], a, b);
            const int t1 = sample(smooth, j.w, p[2], p[3], a, b);
            bits[r] = __ballot(t0 < t1);
        }
        if (lane < 32) row[lane] = static_cast<std::uint8_t>(bits[lane >> 3] >> (8 * (lane & 7)));
        });
}

std::size_t align256(std::size_t v) { return (v + 255) / 256 * 256; }

// Workspace: the blurred image, then one mark word per blur tile.
std::size_t blurred_bytes(int w, int h) { return align256(static_cast<std::size_t>(w) * h); }
std::size_t mark_bytes(int w, int h) {
    return sizeof(unsigned) * static_cast<std::size_t>((w + kS3TileW - 1) / kS3TileW) *
           static_cast<std::size_t>((h + kS3TileH - 1) / kS3TileH);
}

}  // namespace

namespace jobs {

Status make_s3_job(const ImageView& image, const Keypoints& kp, const Descriptors& out, void* ws,
                   std::size_t ws_bytes, S3Job* job) {
    if (!image.data || image.format != PixelFormat::GRAY8 || image.width <= 0 || image.height <= 0 ||
        image.width > INT_MAX / image.height || image.pitch_bytes < image.width || kp.capacity < 0 || !kp.d_count ||
        reinterpret_cast<std::uintptr_t>(kp.d_count) % alignof(int) ||
        (kp.capacity > 0 && (!kp.x || !kp.y || !out.data)) || out.type != DescriptorType::BINARY_U8 ||
        out.dim != s3::kDescriptorBytes || out.row_stride_bytes < s3::kDescriptorBytes)
        return Status::INVALID_ARGUMENT;
    std::size_t need = 0;
    s3::s3_workspace_bytes(image.width, image.height, &need);
    if (!ws || reinterpret_cast<std::uintptr_t>(ws) % 256) return Status::INVALID_ARGUMENT;
    if (ws_bytes < need) return Status::INSUFFICIENT_WORKSPACE;
    S3Job j{};
    j.img = image.data;
    j.w = image.width;
    j.h = image.height;
    j.pitch = image.pitch_bytes;
    j.blurred = static_cast<std::uint8_t*>(ws);
    j.kx = kp.x;
    j.ky = kp.y;
    j.d_count = kp.d_count;
    j.capacity = kp.capacity;
    j.desc = static_cast<std::uint8_t*>(out.data);
    j.desc_stride = out.row_stride_bytes;
    j.tiles_x = (j.w + kS3TileW - 1) / kS3TileW;
    j.tiles = kp.capacity > 0 ? j.tiles_x * ((j.h + kS3TileH - 1) / kS3TileH) : 0;
    j.marks = reinterpret_cast<unsigned*>(static_cast<std::uint8_t*>(ws) + blurred_bytes(j.w, j.h));
    static std::atomic<unsigned> next_epoch{1};
    j.epoch = next_epoch.fetch_add(1, std::memory_order_relaxed);
    *job = j;
    return Status::OK;
}

S3Totals s3_prefix(S3Job* js, int n) {
    S3Totals t{0, 0, 0};
    for (int i = 0; i < n; ++i) {
        js[i].mark_begin = t.mark;
        js[i].blur_begin = t.blur;
        js[i].kp_begin = t.kp;
        t.mark += (js[i].capacity + kS3MarkBlock - 1) / kS3MarkBlock;
        t.blur += js[i].tiles;
        t.kp += (js[i].capacity + kS3KeypointsPerBlock - 1) / kS3KeypointsPerBlock;
    }
    return t;
}

hipError_t launch_s3(const S3Job* table, int count, const S3Totals& t, const S3Job* single, hipStream_t stream,
                     const dev::Replay<S3Totals>* replay) {
    dev::JobSpan<S3Job> span{table, count, single ? *single : S3Job{}, replay ? replay->count : nullptr};
    const auto total = [&](long long S3Totals::*m) {
        return dev::Total(t.*m, replay ? &(replay->totals->*m) : nullptr);
    };
    if (t.mark > 0)
        s3_mark<<<persistent_grid(t.mark, jobs::Stage::S3), kS3MarkBlock, 0, stream>>>(span, total(&S3Totals::mark));
    if (t.blur > 0) s3_blur<<<persistent_grid(t.blur, jobs::Stage::S3), 256, 0, stream>>>(span, total(&S3Totals::blur));
    if (t.kp > 0)
        s3_describe<<<persistent_grid(t.kp, jobs::Stage::S3, 16), 64 * kS3KeypointsPerBlock, 0, stream>>>(span, total(&S3Totals::kp));
    return hipGetLastError();
}

}  // namespace jobs

namespace s3 {

Status s3_workspace_bytes(int width, int height, std::size_t* bytes) {
    if (!bytes) return Status::INVALID_ARGUMENT;
    *bytes = 0;
    if (width <= 0 || height <= 0 || width > INT_MAX / height) return Status::INVALID_ARGUMENT;
    *bytes = blurred_bytes(width, height) + align256(mark_bytes(width, height));
    return Status::OK;
}

Status s3_describe(const ImageView& image, const Keypoints& keypoints, Descriptors output, void* workspace,
                   std::size_t workspace_bytes, hipStream_t stream) {
    jobs::S3Job job{};
    const Status s = jobs::make_s3_job(image, keypoints, output, workspace, workspace_bytes, &job);
    if (s != Status::OK) return s;
    const jobs::S3Totals t = jobs::s3_prefix(&job, 1);
    return jobs::launch_s3(nullptr, 1, t, &job, stream) == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

}  // namespace s3
}  // namespace stitch
