// S6 job descriptors: one canvas (pair) per job with its pyramid levels.
#pragma once

#include <cstddef>
#include <cstdint>

#include "stitch/device.hpp"
#include "stitch/stages.hpp"

namespace stitch::jobs {

constexpr int kS6MaxLevels = 32;

struct S6LevelOffsets {
    int w, h;
    std::size_t a, b, va, vb, ua, ub, z;
};

struct S6Layout {
    S6LevelOffsets lv[kS6MaxLevels];
    int top, blend_top;
    std::size_t bytes;
};

struct S6Level {
    int w, h;
    float* a;   // RGB
    float* b;   // RGB
    float* va;
    float* vb;
    float* ua;
    float* ub;
    float* z;   // RGB, levels 1..blend_top
};

struct S6Job {
    const std::uint8_t* a;
    const std::uint8_t* b;
    int pitch_a, wa, ha;
    int pitch_b, wb, hb;
    int width, height, origin_x, origin_y;
    double denominator_tolerance, snap_tolerance;
    double inverse[9];
    std::uint8_t* out;
    std::uint8_t* mask;
    int out_pitch, mask_pitch;
    s6::RenderStatistics* stats;
    int mode;
    int top, blend_top;  // -1 for feather
    S6Level lv[kS6MaxLevels];
    long long warp_begin;
    long long pb[kS6MaxLevels];  // pull/push tiles of level l
    long long mb[kS6MaxLevels];  // Gaussian/mix tiles of level l
};

struct S6Totals {
    long long warp;
    long long pull[kS6MaxLevels];
    long long mix[kS6MaxLevels];
    int top, blend_top;
};

Status s6_layout(const s6::Geometry& g, s6::BlendMode mode, S6Layout* out);
Status make_s6_job(const ImageView& a, const ImageView& b, const s6::Geometry& g, s6::BlendMode mode,
                   const ImageView& output, const ImageView& mask, void* ws, std::size_t available,
                   s6::RenderStatistics* stats, S6Job* job);
S6Totals s6_prefix(S6Job* jobs, int n);
hipError_t launch_s6(const S6Job* table, int count, const S6Totals& totals, const S6Job* single,
                     hipStream_t stream);

}  // namespace stitch::jobs
