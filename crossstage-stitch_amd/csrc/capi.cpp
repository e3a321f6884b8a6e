// extern "C" layer over the C++ stage API: converts the POD C structs of
// include/cs_stitch.h to the stitch:: types field by field and returns the
// Status as int. No exception crosses this boundary.
#include "cs_stitch.h"

#include <exception>

#include "stitch/features.hpp"
#include "stitch/pyramid.hpp"
#include "stitch/stages.hpp"
#include "stitch/tuning.hpp"

using namespace stitch;

namespace {

ImageView img(const cs_image* v) {
    return {v->data, v->width, v->height, v->pitch_bytes, static_cast<PixelFormat>(v->format)};
}
Keypoints kps(const cs_keypoints* k) { return {k->x, k->y, k->response, k->capacity, k->d_count}; }
Descriptors desc(const cs_descriptors* d) {
    return {static_cast<DescriptorType>(d->type), d->dim, d->row_stride_bytes, d->data};
}
Matches mt(const cs_matches* m) { return {m->idx_a, m->idx_b, m->distance, m->capacity, m->d_count}; }
void put_scale(const ScaleInfo& s, cs_scale_info* o) {
    o->orig_w = s.orig_w;
    o->orig_h = s.orig_h;
    o->work_w = s.work_w;
    o->work_h = s.work_h;
    o->sx = s.sx;
    o->sy = s.sy;
    for (int i = 0; i < 9; ++i) o->D[i] = s.D[i];
}
hipStream_t st(void* s) { return static_cast<hipStream_t>(s); }
ScaleInfo scale(const cs_scale_info* s) {
    ScaleInfo o{};
    o.orig_w = s->orig_w;
    o.orig_h = s->orig_h;
    o.work_w = s->work_w;
    o.work_h = s->work_h;
    o.sx = s->sx;
    o.sy = s->sy;
    for (int i = 0; i < 9; ++i) o.D[i] = s->D[i];
    return o;
}
s5::Config s5cfg(const cs_s5_config* c) {
    return {c->hypotheses, c->previous_hypotheses, c->batch_size, c->refinement_steps, c->threshold, c->seed,
            c->generation};
}
s6::Geometry geo(const cs_geometry* g) {
    s6::Geometry o{};
    o.width = g->width;
    o.height = g->height;
    o.origin_x = g->origin_x;
    o.origin_y = g->origin_y;
    o.width_a = g->width_a;
    o.height_a = g->height_a;
    o.width_b = g->width_b;
    o.height_b = g->height_b;
    o.denominator_tolerance = g->denominator_tolerance;
    o.snap_tolerance = g->snap_tolerance;
    for (int i = 0; i < 9; ++i) o.inverse[i] = g->inverse[i];
    return o;
}
void put_geo(const s6::Geometry& g, cs_geometry* o) {
    o->width = g.width;
    o->height = g.height;
    o->origin_x = g.origin_x;
    o->origin_y = g.origin_y;
    o->width_a = g.width_a;
    o->height_a = g.height_a;
    o->width_b = g.width_b;
    o->height_b = g.height_b;
    o->denominator_tolerance = g.denominator_tolerance;
    o->snap_tolerance = g.snap_tolerance;
    for (int i = 0; i < 9; ++i) o->inverse[i] = g.inverse[i];
}

template <class F>
int guard(F&& f) {
    try {
        return static_cast<int>(f());
    } catch (...) {
        return CS_DEVICE_ERROR;
    }
}

static_assert(sizeof(cs_top2) == sizeof(s4::Top2), "Top2 layout");
static_assert(sizeof(cs_homography_result) == sizeof(HomographyResult), "HomographyResult layout");
static_assert(sizeof(cs_render_statistics) == sizeof(s6::RenderStatistics), "RenderStatistics layout");

}  // namespace

extern "C" {

const char* cs_version(void) { return "crossstage-stitch-amd 0.1 (gfx950)"; }

const char* cs_status_name(int s) { return status_name(static_cast<Status>(s)); }

int cs_make_scale_info(int w, int h, double s, cs_scale_info* out) {
    if (!out) return CS_INVALID_ARGUMENT;
    ScaleInfo si{};
    const Status r = make_scale_info(w, h, s, &si);
    if (r == Status::OK) put_scale(si, out);
    return static_cast<int>(r);
}

int cs_s1_workspace_bytes(int w, int h, double scale, size_t* bytes) {
    return static_cast<int>(s1::s1_workspace_bytes(w, h, scale, bytes));
}

int cs_s1_preprocess(const cs_image* a, const cs_image* b, double scale, cs_image* out_a, cs_image* out_b,
                     cs_scale_info* sa, cs_scale_info* sb, void* stream) {
    if (!a || !b || !out_a || !out_b || !sa || !sb) return CS_INVALID_ARGUMENT;
    return guard([&] {
        ImageView oa = img(out_a), ob = img(out_b);
        ScaleInfo xa{}, xb{};
        const Status r = s1::s1_preprocess(img(a), img(b), scale, &oa, &ob, &xa, &xb, st(stream));
        if (r == Status::OK) {
            put_scale(xa, sa);
            put_scale(xb, sb);
        }
        return r;
    });
}

int cs_s1_preprocess_single(const cs_image* in, double scale, cs_image* out, cs_scale_info* si, void* stream) {
    if (!in || !out || !si) return CS_INVALID_ARGUMENT;
    return guard([&] {
        ImageView o = img(out);
        ScaleInfo x{};
        const Status r = s1::s1_preprocess_single(img(in), scale, &o, &x, st(stream));
        if (r == Status::OK) put_scale(x, si);
        return r;
    });
}

int cs_s1_gradient_workspace_bytes(size_t* bytes) { return static_cast<int>(s1::s1_gradient_workspace_bytes(bytes)); }

int cs_s1_compute_gradient(const cs_image* a, const cs_image* b, void* ws, double* pa, double* pb, void* stream) {
    if (!a || !b) return CS_INVALID_ARGUMENT;
    return guard([&] { return s1::s1_compute_gradient(img(a), img(b), ws, pa, pb, st(stream)); });
}

int cs_s2_workspace_bytes(int w, int h, const cs_s2_config* c, size_t* bytes) {
    if (!c) return CS_INVALID_ARGUMENT;
    return static_cast<int>(s2::s2_workspace_bytes(w, h, {c->fast_threshold, c->max_keypoints, c->harris_k}, bytes));
}

int cs_s2_detect(const cs_image* image, const cs_s2_config* c, const cs_keypoints* out, void* ws, size_t bytes,
                 void* stream) {
    if (!image || !c || !out) return CS_INVALID_ARGUMENT;
    return guard([&] {
        return s2::s2_detect(img(image), {c->fast_threshold, c->max_keypoints, c->harris_k}, kps(out), ws, bytes,
                             st(stream));
    });
}

int cs_s3_workspace_bytes(int w, int h, size_t* bytes) { return static_cast<int>(s3::s3_workspace_bytes(w, h, bytes)); }

int cs_s3_describe(const cs_image* image, const cs_keypoints* k, const cs_descriptors* out, void* ws, size_t bytes,
                   void* stream) {
    if (!image || !k || !out) return CS_INVALID_ARGUMENT;
    return guard([&] { return s3::s3_describe(img(image), kps(k), desc(out), ws, bytes, st(stream)); });
}

int cs_s4_workspace_bytes(int cap_a, int cap_b, const cs_s4_config* c, size_t* bytes) {
    if (!c) return CS_INVALID_ARGUMENT;
    return static_cast<int>(s4::s4_workspace_bytes(cap_a, cap_b, {c->ratio_threshold, c->cross_check != 0}, bytes));
}

int cs_s4_knn_top2(const cs_keypoints* a, const cs_descriptors* da, const cs_keypoints* b, const cs_descriptors* db,
                   cs_top2* out, void* stream) {
    if (!a || !da || !b || !db) return CS_INVALID_ARGUMENT;
    return guard([&] {
        return s4::s4_knn_top2(kps(a), desc(da), kps(b), desc(db), reinterpret_cast<s4::Top2*>(out), st(stream));
    });
}

int cs_s4_match(const cs_keypoints* a, const cs_descriptors* da, const cs_keypoints* b, const cs_descriptors* db,
                const cs_s4_config* c, const cs_matches* out, void* ws, size_t bytes, void* stream) {
    if (!a || !da || !b || !db || !c || !out) return CS_INVALID_ARGUMENT;
    return guard([&] {
        return s4::s4_match(kps(a), desc(da), kps(b), desc(db), {c->ratio_threshold, c->cross_check != 0}, mt(out),
                            ws, bytes, st(stream));
    });
}

int cs_s4_forward_neighbors(int cap_a, int cap_b, const cs_s4_config* c, const void* ws, size_t bytes,
                            const cs_top2** out) {
    if (!c) return CS_INVALID_ARGUMENT;
    const s4::Top2* p = nullptr;
    const Status r =
        s4::s4_forward_neighbors(cap_a, cap_b, {c->ratio_threshold, c->cross_check != 0}, ws, bytes, &p);
    if (r == Status::OK && out) *out = reinterpret_cast<const cs_top2*>(p);
    return static_cast<int>(r);
}

int cs_s4_set_segment(int rows) {
    if (rows < 64 || rows > 512 || rows % 64) return CS_INVALID_ARGUMENT;
    jobs::s4_set_segment(rows);
    return CS_OK;
}

int cs_s5_workspace_bytes(const cs_s5_config* c, size_t* bytes) {
    if (!c) return CS_INVALID_ARGUMENT;
    return static_cast<int>(s5::s5_workspace_bytes(s5cfg(c), bytes));
}

int cs_s5_estimate(const cs_keypoints* a, const cs_keypoints* b, const cs_matches* m, const cs_scale_info* sa,
                   const cs_scale_info* sb, const cs_s5_config* c, uint8_t* mask, cs_homography_result* out, void* ws,
                   size_t bytes, void* stream) {
    if (!a || !b || !m || !sa || !sb || !c) return CS_INVALID_ARGUMENT;
    return guard([&] {
        return s5::estimate(kps(a), kps(b), mt(m), scale(sa), scale(sb), s5cfg(c), mask,
                            reinterpret_cast<HomographyResult*>(out), ws, bytes, st(stream));
    });
}

int cs_s6_prepare_geometry(int wa, int ha, int wb, int hb, const double* h, const cs_geometry_policy* p,
                           cs_geometry* out) {
    if (!out) return CS_INVALID_ARGUMENT;
    s6::GeometryPolicy pol{};
    if (p) {
        pol.denominator_tolerance = p->denominator_tolerance;
        pol.inverse_tolerance = p->inverse_tolerance;
        pol.snap_tolerance = p->snap_tolerance;
        pol.max_canvas_pixels = static_cast<std::size_t>(p->max_canvas_pixels);
    }
    s6::Geometry g{};
    const Status r = s6::prepare_geometry(wa, ha, wb, hb, h, pol, &g);
    put_geo(g, out);
    return static_cast<int>(r);
}

int cs_s6_effective_reductions(const cs_geometry* g, int mode) {
    if (!g) return -1;
    return s6::effective_reductions(geo(g), static_cast<s6::BlendMode>(mode));
}

int cs_s6_workspace_bytes(const cs_geometry* g, int mode, size_t* bytes) {
    if (!g) return CS_INVALID_ARGUMENT;
    return static_cast<int>(s6::s6_workspace_bytes(geo(g), static_cast<s6::BlendMode>(mode), bytes));
}

int cs_s6_render(const cs_image* a, const cs_image* b, const cs_geometry* g, int mode, const cs_image* out,
                 const cs_image* mask, void* ws, size_t bytes, void* stream, cs_render_statistics* stats) {
    if (!a || !b || !g || !out || !mask) return CS_INVALID_ARGUMENT;
    return guard([&] {
        return s6::render(img(a), img(b), geo(g), static_cast<s6::BlendMode>(mode), img(out), img(mask), ws, bytes,
                          st(stream), reinterpret_cast<s6::RenderStatistics*>(stats));
    });
}

int cs_s6_prepare_chain(int n, const int* widths, const int* heights, const double* h_adj,
                        const cs_geometry_policy* p, cs_geometry* out) {
    if (!out || n < 1 || n > s6::kMaxChainFrames) return CS_INVALID_ARGUMENT;
    s6::GeometryPolicy pol{};
    pol.max_canvas_pixels = s6::kChainCanvasPixels;
    if (p) {
        pol.denominator_tolerance = p->denominator_tolerance;
        pol.inverse_tolerance = p->inverse_tolerance;
        pol.snap_tolerance = p->snap_tolerance;
        pol.max_canvas_pixels = static_cast<std::size_t>(p->max_canvas_pixels);
    }
    s6::Geometry g[s6::kMaxChainFrames]{};
    const Status r = s6::prepare_chain_geometry(n, widths, heights, h_adj, pol, g);
    for (int k = 0; k < n; ++k) put_geo(g[k], &out[k]);
    return static_cast<int>(r);
}

int cs_s6_chain_workspace_bytes(int n, const cs_geometry* g, int mode, size_t* bytes) {
    if (!g || n < 1 || n > s6::kMaxChainFrames) return CS_INVALID_ARGUMENT;
    s6::Geometry v[s6::kMaxChainFrames]{};
    for (int k = 0; k < n; ++k) v[k] = geo(&g[k]);
    return guard([&] { return s6::chain_workspace_bytes(n, v, static_cast<s6::BlendMode>(mode), bytes); });
}

int cs_s6_render_chain(int n, const cs_image* frames, const cs_geometry* g, int mode, const cs_image* out,
                       const cs_image* mask, void* ws, size_t bytes, void* stream, cs_render_statistics* stats) {
    if (!frames || !g || !out || !mask || n < 1 || n > s6::kMaxChainFrames) return CS_INVALID_ARGUMENT;
    s6::Geometry v[s6::kMaxChainFrames]{};
    ImageView f[s6::kMaxChainFrames]{};
    for (int k = 0; k < n; ++k) {
        v[k] = geo(&g[k]);
        f[k] = img(&frames[k]);
    }
    return guard([&] {
        return s6::render_chain(n, f, v, static_cast<s6::BlendMode>(mode), img(out), img(mask), ws, bytes,
                                st(stream), reinterpret_cast<s6::RenderStatistics*>(stats));
    });
}

static pyramid::Config pyramid_config(const cs_pyramid_config* c) {
    return {c->levels, c->scale_factor, c->fast_threshold, c->max_keypoints, c->harris_k};
}

int cs_pyramid_level_counts(const cs_pyramid_config* c, int* counts) {
    if (!c) return CS_INVALID_ARGUMENT;
    return static_cast<int>(pyramid::level_counts(pyramid_config(c), counts));
}

int cs_pyramid_workspace_bytes(int w, int h, double scale, const cs_pyramid_config* c, size_t* bytes) {
    if (!c) return CS_INVALID_ARGUMENT;
    return static_cast<int>(pyramid::workspace_bytes(w, h, scale, pyramid_config(c), bytes));
}

int cs_pyramid_detect_describe(const cs_image* rgb, double scale, const cs_pyramid_config* c, const cs_keypoints* kp,
                               const cs_descriptors* d, uint8_t* levels, cs_scale_info* level0, void* ws,
                               size_t bytes, void* stream) {
    if (!rgb || !c || !kp || !d) return CS_INVALID_ARGUMENT;
    return guard([&] {
        ScaleInfo s0{};
        const Status r = pyramid::detect_describe(img(rgb), scale, pyramid_config(c), kps(kp), desc(d), levels, &s0,
                                                  ws, bytes, st(stream));
        if (r == Status::OK && level0) put_scale(s0, level0);
        return r;
    });
}

static_assert(CS_SCENE_FEATURES == features::kSceneCount && CS_GATE_FEATURES == features::kGateCount &&
                  CS_FEATURE_DEVICE_DOUBLES == features::kDeviceDoubles,
              "feature counts");
static_assert(sizeof(cs_top2) == sizeof(s4::Top2), "Top2 layout");

int cs_extract_scene(const cs_image* a, const cs_image* b, double* device, double* host, void* stream, int relations) {
    if (!a || !b) return CS_INVALID_ARGUMENT;
    return guard([&] { return features::extract_scene(img(a), img(b), device, host, st(stream), relations != 0); });
}

int cs_extract_gate(const cs_image* a, const cs_image* b, const cs_keypoints* ka, const cs_keypoints* kb,
                    const cs_matches* m, const uint8_t* mask, const cs_homography_result* h, const cs_scale_info* sa,
                    const cs_scale_info* sb, double* device, double* host, void* stream, const cs_top2* forward,
                    unsigned relation_mask) {
    if (!a || !b || !ka || !kb || !m || !h || !sa || !sb) return CS_INVALID_ARGUMENT;
    return guard([&] {
        return features::extract_gate(img(a), img(b), kps(ka), kps(kb), mt(m), mask,
                                      *reinterpret_cast<const HomographyResult*>(h), scale(sa), scale(sb), device,
                                      host, st(stream), reinterpret_cast<const s4::Top2*>(forward), relation_mask);
    });
}

}  // extern "C"
