/* C ABI of the MI355X stitching path (libcsstitch.so).
 *
 * Plain pointers and sizes only; every function returns a status code
 * (cs_status) and never throws. Device pointers are HIP device memory;
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *
 * Each entry point replaces one interface of the reference (paths relative to
 * the reference repository root); the C++ API behind it keeps the
 * reference's names and argument meaning:
 *   cs_s1_*  src/s1/s1.hpp:12-66           cs_s2_*  src/s2_s3/s2.hpp:26-39
 *   cs_s3_*  src/s2_s3/s3.hpp:14-26        cs_s4_*  src/s4/s4.hpp:23-60
 *   cs_s5_*  src/s5/s5.hpp:21-39           cs_s6_*  src/s6/s6.hpp:21-47, src/s6/geometry.hpp:84-93
 *   cs_runner_* harness/pipeline_runner.hpp:32-81 (PipelineRunner)
 *   cs_controller_*, cs_plan_retry  src/controller/controller.hpp:60-93,
 *       src/controller/reuse.hpp:20, harness/controller_policy.hpp:208-226
 * The batched (cs_batch_*) and served (cs_server_*) forms have no
 * counterpart in the reference; they run many pairs per launch.
 */
#ifndef CS_STITCH_H
#define CS_STITCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: values of stitch::Status (src/common/status.hpp:6-13). */
enum {
    CS_OK = 0,
    CS_INVALID_ARGUMENT = 1,
    CS_INSUFFICIENT_WORKSPACE = 2,
    CS_DEVICE_ERROR = 3,
    CS_NOT_ENOUGH_POINTS = 4,
    CS_DEGENERATE = 5
};

enum { CS_RGB8 = 0, CS_GRAY8 = 1 };
enum { CS_BINARY_U8 = 0, CS_FLOAT32 = 1 };
enum { CS_FEATHER = 0, CS_MULTIBAND3 = 1, CS_MULTIBAND5 = 2, CS_MULTIBAND6 = 3, CS_MULTIBAND7 = 4 };

typedef struct { uint8_t* data; int width, height, pitch_bytes, format; } cs_image;
typedef struct { int orig_w, orig_h, work_w, work_h; double sx, sy; double D[9]; } cs_scale_info;
typedef struct { float* x; float* y; float* response; int capacity; int* d_count; } cs_keypoints;
typedef struct { int type, dim, row_stride_bytes; void* data; } cs_descriptors;
typedef struct { int32_t* idx_a; int32_t* idx_b; float* distance; int capacity; int* d_count; } cs_matches;
typedef struct {
    double H[9];
    double H_s[9];
    int num_matches, num_inliers;
    double inlier_ratio, inlier_coverage, reproj_residual;
    int status;
} cs_homography_result;
typedef struct { double first_distance, second_distance; int first_index, second_index; } cs_top2;
typedef struct { int fast_threshold, max_keypoints; float harris_k; } cs_s2_config;
typedef struct { double ratio_threshold; int cross_check; } cs_s4_config;
typedef struct {
    int hypotheses, previous_hypotheses, batch_size, refinement_steps;
    double threshold;
    uint64_t seed, generation;
} cs_s5_config;
typedef struct {
    double denominator_tolerance, inverse_tolerance, snap_tolerance;
    uint64_t max_canvas_pixels;
} cs_geometry_policy;
typedef struct {
    int width, height, origin_x, origin_y, width_a, height_a, width_b, height_b;
    double denominator_tolerance, snap_tolerance;
    double inverse[9];
} cs_geometry;
typedef struct { uint64_t union_pixels, overlap_pixels; } cs_render_statistics;

/* Library ----------------------------------------------------------------- */
const char* cs_version(void);
const char* cs_status_name(int status);
int cs_make_scale_info(int width, int height, double scale, cs_scale_info* out);

/* S1: RGB -> GRAY + area resize; probe gradient. */
int cs_s1_workspace_bytes(int orig_w, int orig_h, double scale, size_t* bytes);
int cs_s1_preprocess(const cs_image* a, const cs_image* b, double scale, cs_image* out_a, cs_image* out_b,
                     cs_scale_info* scale_a, cs_scale_info* scale_b, void* stream);
int cs_s1_preprocess_single(const cs_image* in, double scale, cs_image* out, cs_scale_info* scale_info,
                            void* stream);
int cs_s1_gradient_workspace_bytes(size_t* bytes);
int cs_s1_compute_gradient(const cs_image* gray_a, const cs_image* gray_b, void* workspace,
                           double* pinned_a, double* pinned_b, void* stream);

/* S2: FAST + Harris top-K. */
int cs_s2_workspace_bytes(int width, int height, const cs_s2_config* config, size_t* bytes);
int cs_s2_detect(const cs_image* image, const cs_s2_config* config, const cs_keypoints* output,
                 void* workspace, size_t workspace_bytes, void* stream);

/* S3: rBRIEF descriptors. */
int cs_s3_workspace_bytes(int width, int height, size_t* bytes);
int cs_s3_describe(const cs_image* image, const cs_keypoints* keypoints, const cs_descriptors* output,
                   void* workspace, size_t workspace_bytes, void* stream);

/* S4: brute-force top-2 + ratio + cross-check. */
int cs_s4_workspace_bytes(int capacity_a, int capacity_b, const cs_s4_config* config, size_t* bytes);
int cs_s4_knn_top2(const cs_keypoints* a, const cs_descriptors* da, const cs_keypoints* b,
                   const cs_descriptors* db, cs_top2* output, void* stream);
int cs_s4_match(const cs_keypoints* a, const cs_descriptors* da, const cs_keypoints* b,
                const cs_descriptors* db, const cs_s4_config* config, const cs_matches* output,
                void* workspace, size_t workspace_bytes, void* stream);
int cs_s4_forward_neighbors(int capacity_a, int capacity_b, const cs_s4_config* config, const void* workspace,
                            size_t workspace_bytes, const cs_top2** output);
/* cs_s4_match without a workspace argument and without any allocation. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The workspace is a fixed device array of
 * CS_S4_NOALLOC_SCRATCH_BYTES held by each GPU until the process ends, enough
 * for 16384 x 16384 descriptors with cross-check; capacities whose
 * cs_s4_workspace_bytes exceed it return CS_INSUFFICIENT_WORKSPACE. Results
 * equal cs_s4_match's. All calls on one GPU share the array: calls that
 * overlap on one GPU race. */
#define CS_S4_NOALLOC_SCRATCH_BYTES ((size_t)36 << 20)
int cs_s4_match_noalloc(const cs_keypoints* a, const cs_descriptors* da, const cs_keypoints* b,
                        const cs_descriptors* db, const cs_s4_config* config, const cs_matches* output,
                        void* stream);
/* Train rows per LDS segment of the S4 fast path (64..512, multiple of 64). */
int cs_s4_set_segment(int rows);

/* S5: MSAC homography B -> A; output must be pinned host memory, valid after a
 * stream synchronization. */
int cs_s5_workspace_bytes(const cs_s5_config* config, size_t* bytes);
int cs_s5_estimate(const cs_keypoints* a, const cs_keypoints* b, const cs_matches* matches,
                   const cs_scale_info* scale_a, const cs_scale_info* scale_b, const cs_s5_config* config,
                   uint8_t* inlier_mask, cs_homography_result* output, void* workspace, size_t workspace_bytes,
                   void* stream);

/* S6: host canvas geometry, then warp + feather / multiband render.
 * policy may be NULL for the default GeometryPolicy. */
int cs_s6_prepare_geometry(int width_a, int height_a, int width_b, int height_b, const double* h,
                           const cs_geometry_policy* policy, cs_geometry* out);
int cs_s6_effective_reductions(const cs_geometry* geometry, int mode);
int cs_s6_workspace_bytes(const cs_geometry* geometry, int mode, size_t* bytes);
int cs_s6_render(const cs_image* a, const cs_image* b, const cs_geometry* geometry, int mode,
                 const cs_image* output, const cs_image* mask, void* workspace, size_t workspace_bytes,
                 void* stream, cs_render_statistics* statistics);

/* S6 over a chain of n frames (1..64; no counterpart in the reference, which
 * renders pairs; the pair render's blend generalised to n frames, equal to
 * cs_s6_render for two). h_adj holds n - 1 homographies, h_adj[9k..] mapping
 * frame k + 1 to frame k. out (n entries) share the chain canvas and frame
 * 0's size; out[k].inverse maps frame-0-original to frame-k-original
 * coordinates. policy may be NULL: default tolerances and a canvas budget of
 * 2^28 pixels. Frames 1.. are sampled through their inverse maps; ownership
 * goes to the highest border score (lowest frame on ties), feather averages
 * by score, multiband mixes every frame's bands. render_chain uploads two
 * small tables from host memory, so unlike cs_s6_render it is not meant for
 * graph capture. frames[k] is RGB8 of out[k].width_b x height_b. */
int cs_s6_prepare_chain(int n, const int* widths, const int* heights, const double* h_adj,
                        const cs_geometry_policy* policy, cs_geometry* out);
int cs_s6_chain_workspace_bytes(int n, const cs_geometry* geometry, int mode, size_t* bytes);
int cs_s6_render_chain(int n, const cs_image* frames, const cs_geometry* geometry, int mode,
                       const cs_image* output, const cs_image* mask, void* workspace, size_t workspace_bytes,
                       void* stream, cs_render_statistics* statistics);

/* Multi-level detection (no counterpart in the reference, whose S2/S3 are
 * single-level): level l is S1 of the RGB input at scale S / f^l, detected
 * and described with S2/S3 on its share of K (OpenCV's ORB split: geometric
 * series with ratio 1/f, last level the remainder); keypoints are returned in
 * level-0 work coordinates, level by level. levels (K bytes) and level0 may
 * be NULL. Levels under 64 px on a side are skipped. */
typedef struct { int levels; double scale_factor; int fast_threshold, max_keypoints; float harris_k; } cs_pyramid_config;
int cs_pyramid_level_counts(const cs_pyramid_config* config, int* counts);
int cs_pyramid_workspace_bytes(int width, int height, double scale, const cs_pyramid_config* config, size_t* bytes);
int cs_pyramid_detect_describe(const cs_image* rgb, double scale, const cs_pyramid_config* config,
                               const cs_keypoints* keypoints, const cs_descriptors* descriptors, uint8_t* levels,
                               cs_scale_info* level0, void* workspace, size_t workspace_bytes, void* stream);

/* Controller observations (src/controller/features.hpp:10-32,
 * relational_features.hpp:8-24): scene features of two GRAY8 images before
 * matching and gate features of a finished S5. device holds
 * CS_FEATURE_DEVICE_DOUBLES doubles of device memory; host receives 32
 * (scene) or 38 (gate) doubles, valid once the stream is synchronized.
 * extract_scene replaces controller::extract_scene (features.cu:125);
 * extract_gate replaces controller::extract_gate (features.cu:136), with
 * relation_mask bits 1 matching, 2 geometry, 4 photometric. An unusable
 * homography or fewer than 4 matches returns INVALID_ARGUMENT (nothing
 * enqueued); inputs the relations reject return CUDA_ERROR. */
#define CS_SCENE_FEATURES 32
#define CS_GATE_FEATURES 38
#define CS_FEATURE_DEVICE_DOUBLES 136
int cs_extract_scene(const cs_image* a, const cs_image* b, double* device, double* host, void* stream,
                     int relations);
int cs_extract_gate(const cs_image* a, const cs_image* b, const cs_keypoints* keypoints_a,
                    const cs_keypoints* keypoints_b, const cs_matches* matches, const uint8_t* inlier_mask,
                    const cs_homography_result* homography, const cs_scale_info* scale_a,
                    const cs_scale_info* scale_b, double* device, double* host, void* stream,
                    const cs_top2* forward, unsigned relation_mask);

/* Plain per-pair runner: harness/pipeline_runner.hpp:32-81 (PipelineRunner),
 * harness/harness_types.hpp:14-106 (HarnessConfig, RunRecord). */
typedef struct cs_runner cs_runner;
typedef struct cs_controller cs_controller;
typedef struct {
    double S;
    int K, R, blend_mode, fast_threshold;
    float harris_k;
    double ratio_threshold;
    int cross_check, msac_batch_size, msac_refinements;
    double msac_threshold;
    uint64_t seed;
    int pyramid_levels;           /* 0 or 1: single-level S2/S3 (the reference's) */
    double pyramid_scale_factor;  /* used when pyramid_levels > 1 */
} cs_harness_config;
typedef struct {
    double S;
    int K, R, blend_mode;
    double t_s1, t_s2, t_s3, t_s4, t_s5, t_s6, t_controller, t_e2e, t_probe, t_stream_wait, t_gpu_span;
    double probe_gradient_a, probe_gradient_b;
    int attempt_count, num_kp_a, num_kp_b, num_matches, num_inliers;
    double inlier_ratio, inlier_coverage, reproj_residual;
    double H[9];
    uint64_t canvas_pixels, overlap_pixels;
    int canvas_width, canvas_height, num_levels, cache_hit;
    char status[32];
    char fail_stage[16];
} cs_run_record;

cs_runner* cs_runner_create(void);
void cs_runner_destroy(cs_runner* runner);
int cs_runner_initialize(cs_runner* runner, int max_width, int max_height, int max_k, int max_r, int canvas_width,
                         int canvas_height, int workspace_mode, uint64_t device_reserve);
int cs_runner_query_memory(int width, int height, int k, int r, int canvas_width, int canvas_height, int mode,
                           uint64_t* device_bytes, uint64_t* workspace_bytes, uint64_t* s6_workspace_bytes);
int cs_runner_upload_pair(cs_runner* runner, const uint8_t* rgb_a, int width_a, int height_a, const uint8_t* rgb_b,
                          int width_b, int height_b);
int cs_runner_use_device_pair(cs_runner* runner, const cs_image* a, const cs_image* b);
int cs_runner_execute_run(cs_runner* runner, const cs_harness_config* config, cs_run_record* record);
int cs_runner_execute_adaptive_run(cs_runner* runner, cs_controller* controller, uint64_t seed,
                                   cs_run_record* record);
int cs_runner_set_flags(cs_runner* runner, int reuse_enabled, int probe_enabled);
/* Feature collection of the plain runner (PipelineRunner::set_collect_features,
 * harness/pipeline_runner.hpp:56): scene features once a pair is on the
 * device, gate features after each successful S5 of execute_run. */
int cs_runner_set_collect_features(cs_runner* runner, int enabled);
/* Observations of the last execute_run (harness/harness_types.hpp:72-76):
 * scene (32 doubles) and gate (38 doubles) may be NULL. */
int cs_runner_features(cs_runner* runner, double* scene, double* gate, int* features_present, double* t_features);
int cs_runner_outputs(cs_runner* runner, cs_image* canvas, cs_image* mask, cs_keypoints* keypoints_a,
                      cs_keypoints* keypoints_b, cs_descriptors* descriptors_a, cs_descriptors* descriptors_b,
                      cs_matches* matches, uint8_t** inlier_mask, cs_homography_result* homography);
void* cs_runner_stream(cs_runner* runner);
/* Per-attempt trace of the last adaptive run (PipelineRunner::attempt_trace,
 * harness/pipeline_runner.hpp:78 and harness/attempt_trace.hpp:10-31): up to
 * 16 records; *count = attempts of that run. */
typedef struct {
    double S;
    int K, R, blend_mode;
    int action, reason;           /* controller decision (ACCEPT 0 / RETRY 1 / REJECT 2) */
    double next_S;
    int next_K, next_R, next_blend_mode;
    int previous_R;               /* hypotheses reused from the previous attempt */
    uint64_t generation;
    cs_homography_result homography;
    double t_s1, t_s2, t_s3, t_s4, t_s5, t_attempt;
    char restart_from[8];         /* "none", "S1", "S2", "S5" */
    char fail_stage[8];
    /* has_feedback: S5 ran (its result is recorded); has_homography: it also
     * returned a usable H (harness/pipeline_runner.cu:777-779). */
    int execution_status, has_decision, has_homography, has_feedback;
} cs_attempt_record;
int cs_runner_attempt_trace(cs_runner* runner, cs_attempt_record* out, int capacity, int* count);

/* Rule-based controller with the built-in legacy-v1 ladder (blind = 1: the
 * probe-free S=0.5 variant), and plan_retry. cs_controller_walk runs
 * begin_pair then observe over n feedbacks; trace rows are (action, reason,
 * S, K, R), the first row the initial knobs. Returns the row count. */
cs_controller* cs_controller_create_legacy(int blind);
void cs_controller_destroy(cs_controller* controller);
int cs_plan_retry(double s0, int k0, int r0, int b0, double s1, int k1, int r1, int b1, int* plan4);
int cs_controller_walk(cs_controller* controller, double gradient_a, double gradient_b, int width_a, int height_a,
                       int width_b, int height_b, const cs_homography_result* feedback, const int* exec_status,
                       int n, double* trace);
/* Controller from a policy document: schema controller-policy-v2 (rules),
 * -v3 (selector and gate trees) or -v4 (pooled boosted models), as
 * harness/controller_policy.hpp:123-206 parses it. On failure returns NULL
 * and writes the reason to error (error_len bytes, may be 0). */
cs_controller* cs_controller_create_policy(const char* json_text, char* error, size_t error_len);
/* What the configured policy reads (controller.hpp:72-77): learned features,
 * the pair relations of the scene features, the gate relation mask, and
 * whether the current rung needs gate features. */
int cs_controller_observations(cs_controller* controller, int* uses_learned_features, int* uses_pair_relations,
                               unsigned* relation_mask, int* needs_gate_features);
/* cs_controller_walk with observations: scene (32 doubles, NULL = absent) goes
 * to begin_pair; gate row i (38 doubles) goes with feedback i when
 * gate_present[i] is nonzero (gate or gate_present NULL = absent). */
int cs_controller_walk_features(cs_controller* controller, double gradient_a, double gradient_b, int width_a,
                                int height_a, int width_b, int height_b, const double* scene,
                                const cs_homography_result* feedback, const int* exec_status, const double* gate,
                                const int* gate_present, int n, double* trace);

/* Batched form (no counterpart in the reference): ragged batches of pairs,
 * each stage one persistent launch over the whole batch, captured in HIP
 * graphs that are replayed with new job tables. Device memory comes from one
 * pool of pool_bytes allocated at creation; the per-pair outputs in
 * cs_pair_outcome are device views that stay valid until the next run.
 * Knobs per pair mirror cs_harness_config; images are device RGB8. */
typedef struct cs_batch cs_batch;
typedef struct {
    cs_image a, b;
    double S;
    int K, R, blend_mode, fast_threshold;
    float harris_k;
    double ratio_threshold;
    int cross_check, msac_batch_size, msac_refinements;
    double msac_threshold;
    uint64_t seed;
} cs_pair_spec;
typedef struct {
    int status;
    char fail_stage[16];
    int num_kp_a, num_kp_b;
    cs_scale_info scale_a, scale_b;
    cs_homography_result homography;
    cs_geometry geometry;
    cs_render_statistics statistics;
    cs_image canvas, mask;
    cs_keypoints keypoints_a, keypoints_b;
    cs_descriptors descriptors_a, descriptors_b;
    cs_matches matches;
    uint8_t* inlier_mask;
} cs_pair_outcome;
typedef struct { double feature_ms, render_ms, host_ms, total_ms; } cs_batch_timings;
/* Per-kernel sums over the renders completed while probing: launches,
 * summed event-timed duration, summed compulsory HBM bytes. */
typedef struct {
    char kernel[32];
    int level;
    int pad;
    int64_t launches;
    double total_ms;
    double algorithmic_bytes;
} cs_kernel_probe;
/* Launch shapes: the S4 train segment, workgroups per CU (eighths of each
 * kernel's default) for every stage and, where non-zero, per stage S1..S6,
 * graph replay, and the S1 tile height (16, 32 or 64 rows). */
typedef struct {
    int s4_segment, workgroups_per_cu, graphs;
    int stage_workgroups_per_cu[6];
    int s1_tile_rows;
} cs_tuning;

int cs_batch_create(int max_pairs, uint64_t pool_bytes, int use_graphs, cs_batch** out);
void cs_batch_destroy(cs_batch* batch);
int cs_batch_run(cs_batch* batch, const cs_pair_spec* pairs, int n, cs_pair_outcome* outcomes,
                 cs_batch_timings* timings);
/* Stream of batches. cs_batch_submit returns once the batch's S6 is enqueued:
 * its S1-S5 ran on the feature stream while the previous batch's S6 was still
 * on the render stream. The previous batch completes inside the call; its
 * outcomes go to previous[0..previous_n) (previous_n must equal its size;
 * previous may be NULL to drop them). cs_batch_finish completes the last
 * submitted batch and writes its n outcomes (n = 0 and nothing written when
 * no batch is pending). A batch's device outputs stay
 * valid until the second submit after its own. timings->render_ms of a
 * submit is the previous batch's S6. Beyond cs_batch_run's contract, no
 * counterpart in the reference (stream form of the batched runner). */
int cs_batch_submit(cs_batch* batch, const cs_pair_spec* pairs, int n, cs_pair_outcome* previous, int previous_n,
                    cs_batch_timings* timings);
int cs_batch_finish(cs_batch* batch, cs_pair_outcome* outcomes, int n);
/* Times the launch-shape knobs on this batch, keeps the fastest and returns
 * it; cache_path (may be NULL) is a JSON file of earlier choices. */
int cs_batch_autotune(cs_batch* batch, const cs_pair_spec* pairs, int n, const char* cache_path, cs_tuning* chosen);
int cs_batch_apply(cs_batch* batch, const cs_tuning* tuning);
int cs_batch_pool(cs_batch* batch, uint64_t* capacity, uint64_t* in_use, uint64_t* peak);
/* Pool bytes one pair takes for S1-S5 (width/height: max over A and B) and
 * one canvas with its S6 workspace: a batch slot holds pool_bytes / 2, a
 * quarter of it for pairs and the rest for canvases. Sizing queries with no
 * counterpart in the reference. */
int cs_batch_pair_bytes(int width, int height, int K, int R, uint64_t* bytes);
/* Per-kernel timing of the S6 launches (measurement; no counterpart in the
 * reference): while on, renders launch directly with HIP events around every
 * kernel on the render stream; enabling clears the sums. probe_read writes up
 * to capacity entries and the number available to *count. */
int cs_batch_set_probe(cs_batch* batch, int on);
int cs_batch_probe_read(cs_batch* batch, cs_kernel_probe* out, int capacity, int* count);
int cs_batch_render_bytes(int canvas_width, int canvas_height, int mode, uint64_t* bytes);

/* Served form (no counterpart in the reference): a request loop replaying an
 * arrival trace with continuous batching, HBM-budget admission and the
 * legacy-v1 adaptive controller per request (or fixed knobs per request). */
typedef struct cs_server cs_server;
typedef struct {
    uint64_t hbm_budget;
    int max_batch, max_live, adaptive, use_graphs;
    double canvas_factor;
} cs_server_config;
typedef struct {
    uint64_t id;
    cs_image a, b;
    double arrival_ms;
    uint64_t seed;
    cs_pair_spec fixed;
    int keep_output;
} cs_request;
typedef struct {
    uint64_t id;
    char status[32];
    char fail_stage[16];
    double arrival_ms, admitted_ms, done_ms;
    int attempts;
    double S;
    int K, R, blend_mode;
    int num_kp_a, num_kp_b;
    cs_homography_result homography;
    cs_geometry geometry;
    cs_render_statistics statistics;
    cs_image canvas, mask;
    uint64_t reserved_bytes;  /* HBM budget held from admission to completion */
} cs_completion;
typedef struct {
    int completed, failed;
    double makespan_ms, p50_ms, p99_ms, mean_ms;
    int feature_launches, render_launches;
    double mean_feature_batch, mean_render_batch;
    uint64_t peak_pool_bytes;
    double feature_busy_ms, render_busy_ms;
} cs_serve_stats;

int cs_server_create(const cs_server_config* config, cs_server** out);
/* A server whose adaptive controller runs a policy document
 * (controller-policy-v2/v3/v4; NULL: legacy-v1). Learned policies get their
 * scene and gate observations per request. A rejected document returns
 * CS_INVALID_ARGUMENT. */
int cs_server_create_policy(const cs_server_config* config, const char* policy_json, cs_server** server);
void cs_server_destroy(cs_server* server);
/* Requests sorted by arrival; completions (n of them) in completion order. */
int cs_server_run_trace(cs_server* server, const cs_request* requests, int n, cs_completion* completions,
                        cs_serve_stats* stats);
int cs_server_release(cs_server* server, uint64_t id);

/* Strip stitching (no counterpart in the reference, whose stitching stops at
 * pairs; its large-frame stress driver harness/stress_multires.cu:81-200 runs
 * each adjacent pair on its own canvas of at most 16,777,216 pixels): n
 * frames in a row, S1-S5 of every adjacent pair in one batched phase (pairs[k]:
 * frame k + 1 -> frame k; each equal to cs_runner_execute_run's), the chain
 * canvas from the composed homographies (budget 2^28 pixels) and one
 * cs_s6_render_chain. Frames are device RGB8; config as for
 * cs_runner_execute_run (single-level S2/S3). Pair buffers are allocated by
 * cs_strip_create; the
 * canvas and the render workspace grow on demand and stay until the next
 * growth or cs_strip_destroy. cs_strip_run returns the outcome's status (the
 * record says which stage and pair failed). */
typedef struct cs_strip cs_strip;
typedef struct {
    int status;
    char fail_stage[16];
    int fail_pair;
    int canvas_width, canvas_height, origin_x, origin_y;
    uint64_t union_pixels, overlap_pixels;
    double t_pairs, t_geometry, t_render, t_total; /* ms: S1-S5 phase, canvas layout, S6, wall clock */
} cs_strip_record;
int cs_strip_create(int max_frames, int max_w, int max_h, int max_k, int max_r, cs_strip** out);
void cs_strip_destroy(cs_strip* strip);
int cs_strip_run(cs_strip* strip, int n, const cs_image* frames, const cs_harness_config* config,
                 cs_strip_record* record, cs_homography_result* pairs, int* keypoints);
/* Canvas, mask (device views) and the n frame geometries of the last run;
 * any output may be NULL. */
int cs_strip_outputs(const cs_strip* strip, cs_image* canvas, cs_image* mask, cs_geometry* geometry, int n);
void* cs_strip_stream(cs_strip* strip);

/* Seeded synthetic pair (no counterpart in the reference; bench and test
 * input): host RGB8 buffers of width*height*3, overlap drawn in [lo, hi],
 * kind 0 textured, 1 low texture, 2 repetitive; h_ba = true H (B -> A). */
int cs_synth_pair(uint64_t seed, int width, int height, double overlap_lo, double overlap_hi, int kind,
                  uint8_t* rgb_a, uint8_t* rgb_b, double* h_ba);
/* Seeded synthetic strip of n frames (1..64; no counterpart in the
 * reference): frames holds n * width * height * 3 bytes, frame k + 1 sees
 * `overlap` (0..1) of frame k; h_adj (may be NULL) gets the n - 1 true
 * adjacent homographies, h_adj[9k..] mapping frame k + 1 to frame k. */
int cs_synth_strip(uint64_t seed, int n, int width, int height, double overlap, int kind, uint8_t* frames,
                   double* h_adj);

#ifdef __cplusplus
}
#endif

#endif /* CS_STITCH_H */
