// extern "C" layer over the C++ stage API: converts the POD C structs of
// include/cs_stitch.h to the stitch:: types field by field and returns the
// Status as int. No exception crosses this boundary.
#include "cs_stitch.h"

#include <exception>

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

}  // extern "C"
