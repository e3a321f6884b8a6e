// Stage APIs S1..S6. Names, structs, argument meaning, Status codes and the
// caller-allocates / workspace-query contract follow the reference headers:
//   S1 src/s1/s1.hpp:12-66          S2 src/s2_s3/s2.hpp:10-39
//   S3 src/s2_s3/s3.hpp:10-26       S4 src/s4/s4.hpp:5-60
//   S5 src/s5/s5.hpp:8-39           S6 src/s6/s6.hpp:11-47, src/s6/geometry.hpp:8-44
// All calls enqueue on the caller's HIP stream, never allocate and never
// synchronize; host argument errors enqueue nothing.
#pragma once

#include <cstddef>
#include <cstdint>

#include "stitch/geometry.hpp"
#include "stitch/types.hpp"

namespace stitch {

namespace s1 {
Status s1_workspace_bytes(int orig_w, int orig_h, double scale, std::size_t* bytes);
Status s1_preprocess_single(const ImageView& in, double scale, ImageView* out, ScaleInfo* scale_info,
                            hipStream_t stream);
Status s1_preprocess(const ImageView& img_a, const ImageView& img_b, double scale, ImageView* out_a,
                     ImageView* out_b, ScaleInfo* scale_a, ScaleInfo* scale_b, hipStream_t stream);
Status s1_gradient_workspace_bytes(std::size_t* bytes);
Status s1_compute_gradient(const ImageView& gray_a, const ImageView& gray_b, void* d_workspace,
                           double* pinned_gradient_a, double* pinned_gradient_b, hipStream_t stream);
Status s1_compute_gradient_single(const ImageView& gray, void* d_workspace, double* pinned_gradient,
                                  hipStream_t stream);
}  // namespace s1

namespace s2 {
constexpr int kEdge = 31;
constexpr int kCandidateFactor = 2;
struct Config {
    int fast_threshold;  // 1..254
    int max_keypoints;   // K >= 0
    float harris_k;      // finite, >= 0
};
Status s2_workspace_bytes(int width, int height, const Config& config, std::size_t* bytes);
Status s2_detect(const ImageView& image, const Config& config, Keypoints output, void* workspace,
                 std::size_t workspace_bytes, hipStream_t stream);
}  // namespace s2

namespace s3 {
constexpr int kDescriptorBytes = 32;
Status s3_workspace_bytes(int width, int height, std::size_t* bytes);
Status s3_describe(const ImageView& image, const Keypoints& keypoints, Descriptors output, void* workspace,
                   std::size_t workspace_bytes, hipStream_t stream);
}  // namespace s3

namespace s4 {
struct Config {
    double ratio_threshold;  // finite, in (0, 1)
    bool cross_check;
};
struct Top2 {
    double first_distance;
    double second_distance;
    int first_index;
    int second_index;
};
Status s4_workspace_bytes(int capacity_a, int capacity_b, const Config& config, std::size_t* bytes);
Status s4_knn_top2(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db,
                   Top2* output, hipStream_t stream);
Status s4_match(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db,
                const Config& config, Matches output, void* workspace, std::size_t workspace_bytes,
                hipStream_t stream);
Status s4_forward_neighbors(int capacity_a, int capacity_b, const Config& config, const void* workspace,
                            std::size_t workspace_bytes, const Top2** output);
}  // namespace s4

namespace s5 {
struct Config {
    int hypotheses;           // total R including previous_hypotheses
    int previous_hypotheses;  // 0 restarts; otherwise extends the saved raw search
    int batch_size;
    int refinement_steps;
    double threshold;         // forward reprojection distance in S1 pixels
    std::uint64_t seed;
    std::uint64_t generation;
};
Status s5_workspace_bytes(const Config& config, std::size_t* bytes);
Status estimate(const Keypoints& a, const Keypoints& b, const Matches& matches, const ScaleInfo& scale_a,
                const ScaleInfo& scale_b, const Config& config, std::uint8_t* inlier_mask, HomographyResult* output,
                void* workspace, std::size_t workspace_bytes, hipStream_t stream);
}  // namespace s5

namespace s6 {
// Values 0..2 are the reference's; 3 and 4 add the 6- and 7-band blends the
// same depth-generic algorithm supports.
enum class BlendMode : int { Feather = 0, Multiband3 = 1, Multiband5 = 2, Multiband6 = 3, Multiband7 = 4 };

inline int requested_reductions(BlendMode m) {
    switch (m) {
        case BlendMode::Feather: return 0;
        case BlendMode::Multiband3: return 3;
        case BlendMode::Multiband5: return 5;
        case BlendMode::Multiband6: return 6;
        case BlendMode::Multiband7: return 7;
    }
    return -1;
}

struct alignas(8) RenderStatistics {
    std::uint64_t union_pixels;
    std::uint64_t overlap_pixels;
};

struct GeometryPolicy {
    double denominator_tolerance = 1e-12;
    double inverse_tolerance = 1e-12;
    double snap_tolerance = 1e-10;
    std::size_t max_canvas_pixels = 16777216;
};

struct Geometry {
    int width = 0;
    int height = 0;
    int origin_x = 0;  // A-original coordinates of canvas pixel (0, 0)
    int origin_y = 0;
    int width_a = 0;
    int height_a = 0;
    int width_b = 0;
    int height_b = 0;
    double denominator_tolerance = 1e-12;
    double snap_tolerance = 1e-10;
    double inverse[9] = {};  // A-original -> B-original, normalized by its largest entry
};

Status prepare_geometry(int width_a, int height_a, int width_b, int height_b, const double* h,
                        const GeometryPolicy& policy, Geometry* out);
int effective_reductions(const Geometry& geometry, BlendMode mode);
Status s6_workspace_bytes(const Geometry& geometry, BlendMode mode, std::size_t* bytes);
Status render(const ImageView& a, const ImageView& b, const Geometry& geometry, BlendMode mode,
              const ImageView& output, const ImageView& mask, void* workspace, std::size_t workspace_bytes,
              hipStream_t stream, RenderStatistics* statistics = nullptr);

// Chains of frames onto one canvas (no counterpart in the reference, which
// renders pairs). h_adj holds n - 1 homographies, h_adj[9k..] mapping frame
// k + 1 to frame k. out[k] (n entries) share the chain canvas and frame 0's
// size; out[k].inverse maps frame-0-original to frame-k-original
// coordinates (the identity for frame 0). The default chain budget is
// kChainCanvasPixels; pass a policy to change it.
constexpr int kMaxChainFrames = 64;
constexpr std::size_t kChainCanvasPixels = std::size_t{1} << 28;
Status prepare_chain_geometry(int n, const int* widths, const int* heights, const double* h_adj,
                              const GeometryPolicy& policy, Geometry* out);
// N-frame render: frame 0 at its integer offset, frames 1.. through their
// inverse maps, feather or multiband blending generalised from the pair
// render (ownership to the highest border score, lowest frame on ties; band
// mix over every frame's pyramids). frames[k] is RGB8 of g[k]'s frame size.
Status chain_workspace_bytes(int n, const Geometry* g, BlendMode mode, std::size_t* bytes);
Status render_chain(int n, const ImageView* frames, const Geometry* g, BlendMode mode, const ImageView& output,
                    const ImageView& mask, void* workspace, std::size_t workspace_bytes, hipStream_t stream,
                    RenderStatistics* statistics = nullptr);
}  // namespace s6

}  // namespace stitch
