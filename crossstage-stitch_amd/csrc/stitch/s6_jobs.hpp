// S6 job descriptors: one canvas (pair) per job with its pyramid levels.
#pragma once

#include <cstddef>
#include <cstdint>

#include "stitch/device.hpp"
#include "stitch/stages.hpp"

namespace stitch::jobs {

constexpr int kS6MaxLevels = 24;  // canvases up to 2^23 pixels per side

struct S6LevelOffsets {
    int w, h;
    std::size_t a, b, va, vb, ua, ub, z, ga, gb, flags;
};

struct S6Layout {
    S6LevelOffsets lv[kS6MaxLevels];
    int top, blend_top;
    std::size_t bytes;
};

// Level 0 is never stored as A colours: A sits at an integer offset on the
// canvas, so its level-0 colour is read from the source bytes. Validity and
// ownership are bits of one byte per pixel (exactly 0 or 1 there) and B's
// warped colour is stored, packed, where B is valid.
// Coarser levels hold FP32 planes.
struct S6Level {
    int w, h;
    std::size_t n;  // w * h; colour planes are n floats apart
    float* a;   // R, G, B planes (levels >= 1): pull-push, then Gaussian at levels >= 2
    float* b;   // R, G, B planes; level 0: B's warped colour where valid, one packed 8-byte word per pixel
    float* va;  // levels >= 1
    float* vb;
    float* ua;  // levels 1..blend_top
    float* ub;
    float* z;   // R, G, B planes, levels 1..blend_top
    float* ga;  // Gaussian colour planes: own planes at level 1, a / b above
    float* gb;
    std::uint8_t* flags;  // level 0: bit 0 va, 1 vb, 2 ua, 3 ub
};

struct S6Job {
    const std::uint8_t* a;
    const std::uint8_t* b;
    int pitch_a, wa, ha;
    int pitch_b, wb, hb;
    int width, height, origin_x, origin_y;
    double denominator_tolerance, snap_tolerance;
    double inv_wa, inv_wb, inv_hb;  // 1 / wa, 1 / wb, 1 / hb, correctly rounded
    bool tol_snap;                  // snap_tolerance >= 16 eps max(1, wb, hb): B's snap threshold is the tolerance
    double inverse[9];
    std::uint8_t* out;
    std::uint8_t* mask;
    int out_pitch, mask_pitch;
    s6::RenderStatistics* stats;
    int mode;
    int top, blend_top;  // -1 for feather
    S6Level lv[kS6MaxLevels];
    int bx0, by0, bx1, by1;  // canvas box outside which B is invalid (inclusive)
    int bi[4];               // canvas box inside which B is valid (inclusive; empty when unknown)
    int ox0, oy0, ox1, oy1;  // box holding every overlap pixel, start aligned to the mix tile
    // Regions (x0, y0, x1, y1 inclusive, in the level's pixels; empty when
    // x1 < x0) that the level-0 band mix depends on: where the band mix runs
    // at levels >= 1, where the Gaussian is needed, where pushed (filled)
    // values are read, and where pulled values (raw sums and weights) are
    // read. Starts are aligned to the tiles of the kernel using them.
    int zbox[kS6MaxLevels][4];
    int gbox[kS6MaxLevels][4];
    int pbox[kS6MaxLevels][4];
    int ubox[kS6MaxLevels][4];
    // Support of A (0) and B (1) at each level: the box outside which the
    // image's raw pull sums and weights are exactly 0 (level 0: A's canvas
    // rectangle and B's box; level l: every pixel whose 5x5 footprint at
    // level l - 1 meets the support there). Pulls store nothing outside it
    // and readers take 0 there without a load.
    int sup[kS6MaxLevels][2][4];
    long long warp_begin;
    long long rb[kS6MaxLevels];  // pull (reduce into level l) tiles
    long long gb[kS6MaxLevels];  // Gaussian (reduce into level l) tiles
    long long pb[kS6MaxLevels];  // push (fill level l) tiles
    long long mb[kS6MaxLevels];  // mix/collapse (level l) tiles
};

struct S6Totals {
    long long warp;
    long long pull[kS6MaxLevels];
    long long gauss[kS6MaxLevels];
    long long push[kS6MaxLevels];
    long long mix[kS6MaxLevels];
    int top, blend_top;
};

Status s6_layout(const s6::Geometry& g, s6::BlendMode mode, S6Layout* out);
Status make_s6_job(const ImageView& a, const ImageView& b, const s6::Geometry& g, s6::BlendMode mode,
                   const ImageView& output, const ImageView& mask, void* ws, std::size_t available,
                   s6::RenderStatistics* stats, S6Job* job);
S6Totals s6_prefix(S6Job* jobs, int n);
// Events around every kernel launch of a render (direct launches only), for
// per-kernel timing; kernel[i] / level[i] name launch i.
struct LaunchProbe {
    enum Kernel { kClear, kWarp, kPull, kPush, kGauss, kMix, kKernels };
    static constexpr int kMax = 4 * kS6MaxLevels + 4;
    hipEvent_t begin[kMax]{}, end[kMax]{};
    int kernel[kMax]{}, level[kMax]{};
    int count{0};
};
const char* probe_kernel_name(int kernel);

// Compulsory HBM bytes of one job's share of a launch of `kernel` at `level`
// (stats: the job's union / overlap pixel counts after the render). Each
// input the kernel needs is counted once and each output written once:
// level-0 flags 1 B/px, A's source bytes 3 B/px, B's warped colour 12 B/px
// where valid, FP32 planes 4 B per pixel and channel.
double s6_algorithmic_bytes(int kernel, int level, const S6Job& job, const s6::RenderStatistics& stats);

hipError_t launch_s6(const S6Job* table, int count, const S6Totals& totals, const S6Job* single,
                     hipStream_t stream, const dev::Replay<S6Totals>* replay = nullptr, LaunchProbe* probe = nullptr);

}  // namespace stitch::jobs
