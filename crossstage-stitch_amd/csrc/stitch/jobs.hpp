// Internal job descriptors for ragged multi-image / multi-pair launches.
// One launch per stage kernel covers every job in a span; each job carries the
// first tile it owns in the launch's flattened tile range. The public stage
// calls build one job by value; the batched and served forms build device
// tables so a whole batch runs in one launch per kernel.
#pragma once

#include <cstddef>
#include <cstdint>

#include "stitch/device.hpp"
#include "stitch/stages.hpp"
#include "stitch/tuning.hpp"

namespace stitch::jobs {

// ------------------------------------------------------------------ S1
struct S1Job {
    const std::uint8_t* src;
    std::uint8_t* dst;
    int src_pitch, dst_pitch;
    int orig_w, orig_h, work_w, work_h;
    float inv_sx, inv_sy, scale_area;
    int mode;      // 0: 1x1, 1: 2x2, 2: 4x4, 3: general area filter
    int tiles_x;   // tiles of kS1TileW x tile_h output pixels
    int tile_h;    // s1_tile_rows() when the job was made; a launch's jobs share it
    long long tile_begin;
};
constexpr int kS1TileW = 64, kS1TileH = 64;

Status make_s1_job(const ImageView& in, const ImageView& out, const ScaleInfo& si, S1Job* job);
inline long long s1_tiles(const S1Job& j) {
    return static_cast<long long>(j.tiles_x) * ((j.work_h + j.tile_h - 1) / j.tile_h);
}
hipError_t launch_s1(const S1Job* table, int count, long long total_tiles, const S1Job* single,
                     hipStream_t stream, const dev::Replay<long long>* replay = nullptr);

// Controller probe gradient of one GRAY8 image (64 x 4 tiles, like S1).
struct GradJob {
    const std::uint8_t* src;
    int pitch, w, h;
    unsigned long long* accum;  // zeroed by the launch
    double* out;                // mean absolute gradient in [0, 1]
    int tiles_x;
    long long tile_begin;
};
inline long long gradient_tiles(const GradJob& g) {
    return static_cast<long long>(g.tiles_x) * ((g.h + kS1TileH - 1) / kS1TileH);
}
hipError_t launch_gradient(const GradJob* table, int count, long long total_tiles, const GradJob* single,
                           hipStream_t stream, const dev::Replay<long long>* replay = nullptr);

// ------------------------------------------------------------------ S2
struct SelectParams {
    std::uint32_t threshold, above, ties, selected;
};

struct S2Job {
    const std::uint8_t* img;
    int w, h, pitch;
    int threshold, K, candidates;
    float harris_k;
    std::uint8_t* maxima;          // w*h, unpitched
    std::uint32_t* hist;           // [256]
    SelectParams* params;
    std::uint32_t* chunk_counts;   // [2 * chunks]
    std::uint32_t* chunk_offsets;  // [2 * chunks]
    std::uint64_t* keys;           // [candidates]
    std::uint32_t* ranks;          // [candidates]
    float* out_x;
    float* out_y;
    float* out_resp;               // may be null
    int* d_count;
    int active;                    // 0: image too small or K == 0 -> count 0 only
    int tiles_x, tiles;            // FAST/NMS tiles
    int chunks;                    // compaction chunks of kS2Chunk pixels
    long long fast_begin, chunk_begin, cand_begin, sort_begin, rank_begin;
};
constexpr int kS2TileW = 64, kS2TileH = 32, kS2Chunk = 4096;
// Top-K order: the candidate keys are sorted in segments of kS2RankSegment
// (one workgroup each, in LDS), then each key's rank is its position in its
// own segment plus a binary search in every other one.
constexpr int kS2RankKeys = 256;     // keys ranked per workgroup (one per thread)
constexpr int kS2RankSegment = 2048; // keys per sorted segment

std::size_t s2_layout_bytes(int w, int h, int K, std::size_t offsets[7]);
Status make_s2_job(const ImageView& image, const s2::Config& cfg, const Keypoints& out, void* ws,
                   std::size_t ws_bytes, S2Job* job);
struct S2Totals {
    long long jobs_tiles, fast, chunks, cands, sort, rank;
};
// Fills the per-kernel tile prefixes of jobs[0..n) and returns the totals.
S2Totals s2_prefix(S2Job* jobs, int n);
hipError_t launch_s2(const S2Job* table, int count, const S2Totals& totals, const S2Job* single,
                     hipStream_t stream, const dev::Replay<S2Totals>* replay = nullptr);

// ------------------------------------------------------------------ S3
struct S3Job {
    const std::uint8_t* img;
    int w, h, pitch;
    std::uint8_t* blurred;  // w*h, unpitched
    const float* kx;
    const float* ky;
    const int* d_count;
    int capacity;
    std::uint8_t* desc;
    int desc_stride;
    int tiles_x, tiles;
    // Blur tiles that a descriptor samples hold `epoch` in `marks` (one word
    // per tile); the blur skips the others. Each job takes a fresh epoch, so
    // marks left by earlier launches never match.
    unsigned* marks;
    unsigned epoch;
    long long mark_begin, blur_begin, kp_begin;
};
constexpr int kS3TileW = 64, kS3TileH = 16;
constexpr int kS3KeypointsPerBlock = 4;  // one wave per keypoint
constexpr int kS3MarkBlock = 256;        // keypoints per marking workgroup

Status make_s3_job(const ImageView& image, const Keypoints& kp, const Descriptors& out, void* ws,
                   std::size_t ws_bytes, S3Job* job);
struct S3Totals {
    long long mark, blur, kp;
};
S3Totals s3_prefix(S3Job* jobs, int n);
hipError_t launch_s3(const S3Job* table, int count, const S3Totals& totals, const S3Job* single,
                     hipStream_t stream, const dev::Replay<S3Totals>* replay = nullptr);

// ------------------------------------------------------------------ S4
// Binary 256-bit fast path: per (query block, train segment) tile, 4 waves x
// 64 lanes x kS4QueriesPerLane queries against a train segment in LDS.
constexpr int kS4QueriesPerLane = 2;
constexpr int kS4QueryBlock = 4 * 64 * kS4QueriesPerLane;  // 512 queries per workgroup
constexpr int kS4KeyShift = 20;                            // key = distance << 20 | index
constexpr unsigned kS4NoKey = 0xffffffffu;

struct S4Job {
    const std::uint8_t* da;
    const std::uint8_t* db;
    int stride_a, stride_b;
    const int* count_a;
    const int* count_b;
    int cap_a, cap_b;
    int dim;
    int binary;        // 1: BINARY_U8, 0: FLOAT32
    int fast;          // 1: fused 256-bit path
    int cross_check;
    double ratio;      // already squared for FLOAT32
    int segment;       // train rows per segment (fast path)
    int qblocks, segs;
    std::uint32_t* part_fwd;  // [segs][cap_a] uint2 (fast path)
    std::uint32_t* part_col;  // [qblocks][cap_b] (fast path, cross-check)
    s4::Top2* forward;        // [cap_a]
    s4::Top2* reverse;        // [cap_b] (generic path, cross-check)
    int* rev_first;           // [cap_b] reverse top-1 index (cross-check)
    std::int32_t* idx_a;
    std::int32_t* idx_b;
    float* distance;
    int* d_count;
    long long dist_begin, qmerge_begin, cmerge_begin, gfwd_begin, grev_begin;
};

std::size_t s4_layout_bytes(int cap_a, int cap_b, bool cross_check, int segment, std::size_t offsets[5]);
Status make_s4_job(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db,
                   const s4::Config& cfg, const Matches& out, void* ws, std::size_t ws_bytes, S4Job* job);
struct S4Totals {
    long long dist, qmerge, cmerge, match, generic_fwd, generic_rev;
};
S4Totals s4_prefix(S4Job* jobs, int n);
hipError_t launch_s4(const S4Job* table, int count, const S4Totals& totals, const S4Job* single,
                     hipStream_t stream, const dev::Replay<S4Totals>* replay = nullptr);

// ------------------------------------------------------------------ S5
struct S5Model {
    double h[9];
    double loss;
    int hypothesis;
    int valid;
};

struct S5State {
    S5Model best;
    HomographyResult result;
    const float* ax;
    const float* ay;
    const float* bx;
    const float* by;
    const std::int32_t* idx_a;
    const std::int32_t* idx_b;
    const int* count_a;
    const int* count_b;
    const int* count_matches;
    int na, nb, nm;
    int capacity_a, capacity_b, capacity_matches;
    int evaluated;
    int batch_size;
    double threshold;
    std::uint64_t seed, generation;
    ScaleInfo scale_a, scale_b;
    Status status;
};
constexpr std::size_t kS5ModelOffset = (sizeof(S5State) + 255) / 256 * 256;
// Hypotheses evaluated per launch; the workspace holds at most this many models.
constexpr int kS5MaxModels = 8192;

struct S5Job {
    const float* ax;
    const float* ay;
    const float* bx;
    const float* by;
    const int* count_a;
    const int* count_b;
    int cap_a, cap_b;
    const std::int32_t* idx_a;
    const std::int32_t* idx_b;
    const int* count_m;
    int cap_m;
    ScaleInfo sa, sb;
    s5::Config cfg;
    double inverse_a[9];
    double d_b[9];
    std::uint8_t* mask;
    S5State* state;
    S5Model* models;  // [hypotheses - previous]
    HomographyResult* result_dev;  // state->result
    int first, count;  // hypotheses [first, first + count)
    long long gen_begin, score_begin;
};
Status make_s5_job(const Keypoints& a, const Keypoints& b, const Matches& m, const ScaleInfo& sa,
                   const ScaleInfo& sb, const s5::Config& c, std::uint8_t* mask, void* ws, std::size_t bytes,
                   S5Job* job);
struct S5Totals {
    long long gen, score;
};
S5Totals s5_prefix(S5Job* jobs, int n);
hipError_t launch_s5(const S5Job* table, int count, const S5Totals& totals, const S5Job* single,
                     hipStream_t stream, const dev::Replay<S5Totals>* replay = nullptr);

}  // namespace stitch::jobs
