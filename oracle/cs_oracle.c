/* CPU oracle for the per-pair stitching path (S1-S6).
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, in plain scalar C, the
 * arithmetic of the reference's stages so the HIP implementation can be
 * checked against it. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load the library built from it, and only as a checker
 * or as the timed CPU baseline -- never as a fallback for the product path.
 *
 * Every function cites the reference file:line whose semantics it follows
 * (paths relative to the reference repository root). Floating-point
 * expressions keep the reference's operation order and are compiled with
 * -ffp-contract=off so no multiply-add is fused, which is the semantics of
 * the reference's host build of the same arithmetic.
 *
 * Pinning (how each part is checked against something other than itself) is
 * listed in the header comment of each section and in the tests that cite it.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orb_pairs.h"

#define ORC_OK 0
#define ORC_INVALID_ARGUMENT 1
#define ORC_INSUFFICIENT_WORKSPACE 2
#define ORC_NOT_ENOUGH_POINTS 4
#define ORC_DEGENERATE 5

/* ======================================================================
 * Common geometry (src/common/geometry.hpp:13-48)
 * ====================================================================== */

typedef struct {
    int orig_w, orig_h, work_w, work_h;
    double sx, sy;
    double D[9];
} orc_scale_info;

/* make_scale_info, src/common/geometry.hpp:13-25. */
int orc_make_scale_info(int w, int h, double s, orc_scale_info* out) {
    if (!out || w <= 0 || h <= 0 || !(s > 0.0) || s > 1.0) return ORC_INVALID_ARGUMENT;
    long ww = lround((double)w * s), hh = lround((double)h * s);
    out->orig_w = w;
    out->orig_h = h;
    out->work_w = (int)(ww > 1 ? ww : 1);
    out->work_h = (int)(hh > 1 ? hh : 1);
    out->sx = (double)out->work_w / w;
    out->sy = (double)out->work_h / h;
    const double d[9] = {out->sx, 0, 0.5 * (out->sx - 1.0), 0, out->sy, 0.5 * (out->sy - 1.0), 0, 0, 1};
    memcpy(out->D, d, sizeof d);
    return ORC_OK;
}

/* mat3_mul, src/common/geometry.hpp:27-33 (also s5_math.hpp:20-26). */
static void mat3_mul(const double* a, const double* b, double* out) {
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    memcpy(out, r, sizeof r);
}

/* mat3_inverse, src/common/geometry.hpp:35-48. */
static int mat3_inverse(const double* m, double* out) {
    const double c00 = m[4] * m[8] - m[5] * m[7];
    const double c01 = m[5] * m[6] - m[3] * m[8];
    const double c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (fabs(det) < 1e-300) return 0;
    const double inv = 1.0 / det;
    const double r[9] = {c00 * inv, (m[2] * m[7] - m[1] * m[8]) * inv, (m[1] * m[5] - m[2] * m[4]) * inv,
                         c01 * inv, (m[0] * m[8] - m[2] * m[6]) * inv, (m[2] * m[3] - m[0] * m[5]) * inv,
                         c02 * inv, (m[1] * m[6] - m[0] * m[7]) * inv, (m[0] * m[4] - m[1] * m[3]) * inv};
    memcpy(out, r, sizeof r);
    return 1;
}

/* ======================================================================
 * S1: RGB -> GRAY + area resize (src/s1/s1.cu:31-103, :113-134, :140-203)
 * Pinned by: integer closed form (modes 0-2), known answers of
 * tests/test_s1_cuda.cu (exact at S in {1, .5, .25}), gradient known answers
 * (constant image -> 0, checkerboard -> 1).
 * ====================================================================== */

static int rgb_to_gray(int r, int g, int b) { return (9798 * r + 19235 * g + 3735 * b + 16384) >> 15; }

/* s1_preprocess_kernel for one image; mode choice per determine_mode (s1.cu:105-110). */
int orc_s1_preprocess(const uint8_t* src, int w, int h, int pitch, double scale,
                      uint8_t* dst, int dst_pitch, int* work_w, int* work_h) {
    orc_scale_info si;
    if (!src || !dst || pitch < 3 * w) return ORC_INVALID_ARGUMENT;
    if (orc_make_scale_info(w, h, scale, &si) != ORC_OK) return ORC_INVALID_ARGUMENT;
    if (dst_pitch < si.work_w) return ORC_INVALID_ARGUMENT;
    const int ww = si.work_w, wh = si.work_h;
    int mode = 3;
    if (w == ww && h == wh) mode = 0;
    else if (w == 2 * ww && h == 2 * wh) mode = 1;
    else if (w == 4 * ww && h == 4 * wh) mode = 2;
    const float inv_sx = (float)w / (float)ww;
    const float inv_sy = (float)h / (float)wh;
    const float scale_area = 1.0f / (inv_sx * inv_sy);
    for (int yd = 0; yd < wh; ++yd) {
        for (int xd = 0; xd < ww; ++xd) {
            int out;
            if (mode == 0) {
                const uint8_t* p = src + (size_t)yd * pitch + 3 * xd;
                out = rgb_to_gray(p[0], p[1], p[2]);
            } else if (mode == 1) {
                const uint8_t* r0 = src + (size_t)(2 * yd) * pitch + 3 * (2 * xd);
                const uint8_t* r1 = src + (size_t)(2 * yd + 1) * pitch + 3 * (2 * xd);
                out = (rgb_to_gray(r0[0], r0[1], r0[2]) + rgb_to_gray(r0[3], r0[4], r0[5]) +
                       rgb_to_gray(r1[0], r1[1], r1[2]) + rgb_to_gray(r1[3], r1[4], r1[5]) + 2) >> 2;
            } else if (mode == 2) {
                int sum = 0;
                for (int dy = 0; dy < 4; ++dy) {
                    const uint8_t* r = src + (size_t)(4 * yd + dy) * pitch + 3 * (4 * xd);
                    for (int dx = 0; dx < 4; ++dx) sum += rgb_to_gray(r[3 * dx], r[3 * dx + 1], r[3 * dx + 2]);
                }
                out = (int)rintf((float)sum * 0.0625f);
            } else {
                const float fsy1 = (float)yd * inv_sy, fsy2 = (float)(yd + 1) * inv_sy;
                const float fsx1 = (float)xd * inv_sx, fsx2 = (float)(xd + 1) * inv_sx;
                const int y_start = (int)floorf(fsy1), y_end = (int)ceilf(fsy2);
                const int x_start = (int)floorf(fsx1), x_end = (int)ceilf(fsx2);
                float total = 0.0f;
                for (int ys = y_start; ys < y_end && ys < h; ++ys) {
                    const float y_lo = fmaxf((float)ys, fsy1);
                    const float y_hi = fminf((float)(ys + 1), fsy2);
                    const float wy = fmaxf(0.0f, y_hi - y_lo);
                    if (wy <= 0.0f) continue;
                    const uint8_t* row = src + (size_t)ys * pitch;
                    for (int xs = x_start; xs < x_end && xs < w; ++xs) {
                        const float x_lo = fmaxf((float)xs, fsx1);
                        const float x_hi = fminf((float)(xs + 1), fsx2);
                        const float wx = fmaxf(0.0f, x_hi - x_lo);
                        if (wx <= 0.0f) continue;
                        const uint8_t* rgb = row + 3 * xs;
                        total += (float)rgb_to_gray(rgb[0], rgb[1], rgb[2]) * (wx * wy);
                    }
                }
                out = (int)rintf(total * scale_area);
            }
            dst[(size_t)yd * dst_pitch + xd] = (uint8_t)out;
        }
    }
    if (work_w) *work_w = ww;
    if (work_h) *work_h = wh;
    return ORC_OK;
}

/* Mean absolute gradient, s1_gradient_accum/normalize (s1.cu:140-203). */
double orc_s1_gradient(const uint8_t* g, int w, int h, int pitch) {
    unsigned long long sum = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int c = g[(size_t)y * pitch + x];
            if (x + 1 < w) sum += (unsigned long long)abs(g[(size_t)y * pitch + x + 1] - c);
            if (y + 1 < h) sum += (unsigned long long)abs(g[(size_t)(y + 1) * pitch + x] - c);
        }
    const unsigned long long edges = (unsigned long long)h * (w - 1) + (unsigned long long)w * (h - 1);
    return edges == 0 ? 0.0 : (double)sum / (255.0 * (double)edges);
}

/* ======================================================================
 * S2: FAST-9 -> strict 3x3 NMS in the 31-px margin -> top 2K by FAST score ->
 * Harris -> top K (src/s2_s3/s2.cu:39-310; tools/s2_s3/reference.py:50-147).
 * Pinned by: golden vectors written by the reference's NumPy module
 * (tests/golden), brute-force rescans on tiny images.
 * ====================================================================== */

#define S2_EDGE 31
static const int kCircleX[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
static const int kCircleY[16] = {3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1, 0, 1, 2, 3};

/* fast_score, s2.cu:39-67 (cornerScore<16>). */
int orc_fast_score(const uint8_t* img, int pitch, int x, int y, int threshold) {
    const int v = img[(size_t)y * pitch + x];
    int d[16];
    for (int k = 0; k < 16; ++k) d[k] = v - img[(size_t)(y + kCircleY[k]) * pitch + x + kCircleX[k]];
    int best = -2147483647 - 1;
    for (int k = 0; k < 16; ++k) {
        int lo = 2147483647, hi = -2147483647 - 1;
        for (int j = 0; j < 9; ++j) {
            const int e = d[(k + j) & 15];
            if (e < lo) lo = e;
            if (e > hi) hi = e;
        }
        const int m = lo > -hi ? lo : -hi;
        if (m > best) best = m;
    }
    return best > threshold ? best - 1 : 0;
}

static uint32_t order_bits(float value) {
    uint32_t u;
    memcpy(&u, &value, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* harris_response, s2.cu:248-269; reference.py:103-128. */
float orc_harris_response(const uint8_t* img, int pitch, int x0, int y0, float k) {
    int a = 0, b = 0, c = 0;
    for (int i = 0; i < 7; ++i) {
        const uint8_t* row = img + (size_t)(y0 - 3 + i) * pitch + (x0 - 3);
        for (int j = 0; j < 7; ++j) {
            const uint8_t* q = row + j;
            const int ix = (q[1] - q[-1]) * 2 + (q[-pitch + 1] - q[-pitch - 1]) + (q[pitch + 1] - q[pitch - 1]);
            const int iy = (q[pitch] - q[-pitch]) * 2 + (q[pitch - 1] - q[-pitch - 1]) + (q[pitch + 1] - q[-pitch + 1]);
            a += ix * ix;
            b += iy * iy;
            c += ix * iy;
        }
    }
    const float fa = (float)a, fb = (float)b, fc = (float)c;
    const float trace = fa + fb;
    const float det = fa * fb - fc * fc;
    const float penalty = (k * trace) * trace;
    return (det - penalty) * 0x1.bb9da2p-52f;
}

static int cmp_u64_desc(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

/* s2_detect (s2.cu:333-387): writes up to K keypoints sorted by Harris response
 * desc, ties by row-major index. Returns the count, or -1 on invalid input.
 * scores_out / maxima_out (optional, w*h bytes) receive the FAST and NMS maps. */
int orc_s2_detect(const uint8_t* img, int w, int h, int pitch, int fast_threshold, int max_keypoints,
                  float harris_k, float* xs, float* ys, float* resp, uint8_t* scores_out, uint8_t* maxima_out) {
    if (!img || w <= 0 || h <= 0 || pitch < w || fast_threshold < 1 || fast_threshold > 254 ||
        max_keypoints < 0 || !(harris_k >= 0.0f) || !isfinite(harris_k)) return -1;
    if (max_keypoints == 0 || w <= 2 * S2_EDGE || h <= 2 * S2_EDGE) return 0;
    const size_t n = (size_t)w * h;
    uint8_t* scores = (uint8_t*)calloc(n, 1);
    uint8_t* maxima = (uint8_t*)calloc(n, 1);
    for (int y = 3; y < h - 3; ++y)
        for (int x = 3; x < w - 3; ++x) scores[(size_t)y * w + x] = (uint8_t)orc_fast_score(img, pitch, x, y, fast_threshold);
    uint32_t hist[256] = {0};
    for (int y = S2_EDGE; y < h - S2_EDGE; ++y)
        for (int x = S2_EDGE; x < w - S2_EDGE; ++x) {
            const int s = scores[(size_t)y * w + x];
            int keep = s > 0;
            for (int dy = -1; keep && dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    if ((dx || dy) && s <= scores[(size_t)(y + dy) * w + x + dx]) { keep = 0; break; }
            if (keep) { maxima[(size_t)y * w + x] = (uint8_t)s; hist[s]++; }
        }
    /* select_kernel (s2.cu:111-139): boundary score of the 2K pre-selection. */
    const uint32_t cap = (uint32_t)(2 * max_keypoints);
    uint32_t suffix[258] = {0};
    for (int s = 255; s >= 1; --s) suffix[s] = suffix[s + 1] + hist[s];
    uint32_t threshold = 1, above, ties;
    const uint32_t total = suffix[1];
    if (total <= cap) { threshold = 1; above = total - hist[1]; ties = hist[1]; }
    else {
        int s = 255;
        while (!(suffix[s] >= cap && suffix[s + 1] < cap)) --s;
        threshold = (uint32_t)s; above = suffix[s + 1]; ties = cap - suffix[s + 1];
    }
    const uint32_t selected = above + ties;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (selected ? selected : 1));
    uint32_t n_above = 0, n_tie = 0;
    for (size_t p = 0; p < n; ++p) {
        const uint32_t s = maxima[p];
        if (!s) continue;
        if (s > threshold) keys[n_above++] = p;
        else if (s == threshold && n_tie < ties) keys[above + n_tie++] = p;
    }
    /* harris_kernel + rank_kernel (s2.cu:272-310). */
    for (uint32_t i = 0; i < selected; ++i) {
        const uint32_t px = (uint32_t)keys[i];
        const float r = orc_harris_response(img, pitch, (int)(px % (uint32_t)w), (int)(px / (uint32_t)w), harris_k);
        keys[i] = ((uint64_t)order_bits(r) << 32) | (0xffffffffu - px);
    }
    qsort(keys, selected, sizeof(uint64_t), cmp_u64_desc);
    const int count = (int)(selected < (uint32_t)max_keypoints ? selected : (uint32_t)max_keypoints);
    for (int i = 0; i < count; ++i) {
        const uint32_t px = 0xffffffffu - (uint32_t)keys[i];
        const uint32_t ob = (uint32_t)(keys[i] >> 32);
        const uint32_t u = (ob & 0x80000000u) ? (ob & 0x7fffffffu) : ~ob;
        float r;
        memcpy(&r, &u, 4);
        xs[i] = (float)(px % (uint32_t)w);
        ys[i] = (float)(px / (uint32_t)w);
        if (resp) resp[i] = r;
    }
    if (scores_out) memcpy(scores_out, scores, n);
    if (maxima_out) memcpy(maxima_out, maxima, n);
    free(keys); free(scores); free(maxima);
    return count;
}

/* ======================================================================
 * S3: float32 7x7 sigma-2 Gaussian + IC angle + rBRIEF
 * (src/s2_s3/s3.cu:26-157; tools/s2_s3/reference.py:152-284).
 * Pinned by: tap values of getGaussianKernel(7, 2) (computed independently in
 * tests), u_max known values, pattern rows 0 and 255, golden vectors from the
 * reference's NumPy module.
 * ====================================================================== */

static const float kTaps[7] = {0x1.1f5f62p-4f, 0x1.0c70fcp-3f, 0x1.869472p-3f, 0x1.ba95c0p-3f,
                               0x1.869472p-3f, 0x1.0c70fcp-3f, 0x1.1f5f62p-4f};
static const int kUMax[16] = {15, 15, 15, 15, 14, 14, 14, 13, 13, 12, 11, 10, 9, 8, 6, 3};
static const int kPairs[CS_ORB_PAIR_COUNT * 4] = CS_ORB_PAIRS_INIT;

static int reflect101_i(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i : 2 * n - 2 - i;
    return i;
}

/* blur_rows_kernel + blur_cols_kernel, s3.cu:47-73. out is w*h, unpitched. */
void orc_s3_blur(const uint8_t* img, int w, int h, int pitch, uint8_t* out) {
    float* rows = (float*)malloc(sizeof(float) * (size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float acc = 0.0f;
            for (int i = 0; i < 7; ++i) acc = acc + kTaps[i] * (float)img[(size_t)y * pitch + reflect101_i(x + i - 3, w)];
            rows[(size_t)y * w + x] = acc;
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float acc = 0.0f;
            for (int i = 0; i < 7; ++i) acc = acc + kTaps[i] * rows[(size_t)reflect101_i(y + i - 3, h) * w + x];
            int v = (int)rintf(acc);
            out[(size_t)y * w + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    free(rows);
}

/* cv::fastAtan2 in float32, s3.cu:76-90; reference.py:216-228. */
float orc_fast_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const int use_x = ax >= ay;
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

/* describe_kernel for keypoints [0, n), s3.cu:105-157. desc is n*32 bytes.
 * blurred may be NULL (computed here). */
void orc_s3_describe(const uint8_t* img, int w, int h, int pitch, const float* kx, const float* ky, int n,
                     const uint8_t* blurred_in, uint8_t* desc) {
    uint8_t* blurred = (uint8_t*)blurred_in;
    if (!blurred && n > 0) {
        blurred = (uint8_t*)malloc((size_t)w * h);
        orc_s3_blur(img, w, h, pitch, blurred);
    }
    for (int k = 0; k < n; ++k) {
        uint8_t* row = desc + (size_t)k * 32;
        const int x = (int)rintf(kx[k]), y = (int)rintf(ky[k]);
        if (x < S2_EDGE || x >= w - S2_EDGE || y < S2_EDGE || y >= h - S2_EDGE) { memset(row, 0, 32); continue; }
        const uint8_t* centre = img + (size_t)y * pitch + x;
        int m10 = 0, m01 = 0;
        for (int u = -15; u <= 15; ++u) m10 += u * centre[u];
        for (int v = 1; v <= 15; ++v) {
            const int d = kUMax[v];
            int v_sum = 0;
            for (int u = -d; u <= d; ++u) {
                const int plus = centre[u + v * pitch], minus = centre[u - v * pitch];
                v_sum += plus - minus;
                m10 += u * (plus + minus);
            }
            m01 += v * v_sum;
        }
        const float angle = orc_fast_atan2((float)m01, (float)m10) * 0x1.1df46ap-6f;
        const float a = (float)cos((double)angle), b = (float)sin((double)angle);
        const uint8_t* smooth = blurred + (size_t)y * w + x;
        for (int l = 0; l < 32; ++l) {
            int value = 0;
            for (int j = 0; j < 8; ++j) {
                const int* p = &kPairs[(l * 8 + j) * 4];
                int t[2];
                for (int s = 0; s < 2; ++s) {
                    const float fx = (float)p[2 * s], fy = (float)p[2 * s + 1];
                    const int ix = (int)rintf(fx * a - fy * b);
                    const int iy = (int)rintf(fx * b + fy * a);
                    t[s] = smooth[iy * w + ix];
                }
                value |= (t[0] < t[1]) << j;
            }
            row[l] = (uint8_t)value;
        }
    }
    if (blurred != blurred_in) free(blurred);
}

/* ======================================================================
 * S4: exact top-2 + ratio test + cross-check (src/s4/s4.cu:57-263).
 * Pinned by: the brute-force sort oracle and edge answers of
 * tests/test_s4_cuda.cu:60-94, :185-215 (restated as tests here).
 * ====================================================================== */

typedef struct {
    double first_distance, second_distance;
    int first_index, second_index;
} orc_top2;

/* top2_kernel semantics: exact first/second by (distance, index). dim bytes
 * (binary) or floats (float32). Rows [0, nb) of train are searched. */
void orc_s4_top2(int binary, const void* qa, int na, int stride_a, const void* tb, int nb, int stride_b,
                 int dim, orc_top2* out) {
    for (int q = 0; q < na; ++q) {
        orc_top2 t = {INFINITY, INFINITY, -1, -1};
        for (int j = 0; j < nb; ++j) {
            double d = 0;
            if (binary) {
                const uint8_t* x = (const uint8_t*)qa + (size_t)q * stride_a;
                const uint8_t* y = (const uint8_t*)tb + (size_t)j * stride_b;
                for (int k = 0; k < dim; ++k) d += __builtin_popcount((unsigned)(x[k] ^ y[k]));
            } else {
                const float* x = (const float*)((const uint8_t*)qa + (size_t)q * stride_a);
                const float* y = (const float*)((const uint8_t*)tb + (size_t)j * stride_b);
                for (int k = 0; k < dim; ++k) {
                    const double delta = (double)x[k] - y[k];
                    d = d + delta * delta;
                }
            }
            if (t.first_index < 0 || d < t.first_distance) {
                t.second_distance = t.first_distance; t.second_index = t.first_index;
                t.first_distance = d; t.first_index = j;
            } else if (t.second_index < 0 || d < t.second_distance) {
                t.second_distance = d; t.second_index = j;
            }
        }
        out[q] = t;
    }
}

/* s4_match: forward (and reverse) top-2, accept rule s4.cu:221-225, A-ordered
 * compaction s4.cu:251-261. Returns the match count. */
int orc_s4_match(int binary, const void* da, int na, int stride_a, const void* db, int nb, int stride_b, int dim,
                 double ratio, int cross_check, int32_t* idx_a, int32_t* idx_b, float* distance,
                 orc_top2* forward_out) {
    orc_top2* fwd = (orc_top2*)malloc(sizeof(orc_top2) * (na ? na : 1));
    orc_top2* rev = cross_check ? (orc_top2*)malloc(sizeof(orc_top2) * (nb ? nb : 1)) : NULL;
    orc_s4_top2(binary, da, na, stride_a, db, nb, stride_b, dim, fwd);
    if (rev) orc_s4_top2(binary, db, nb, stride_b, da, na, stride_a, dim, rev);
    const double r = binary ? ratio : ratio * ratio;
    int count = 0;
    for (int q = 0; q < na; ++q) {
        const orc_top2 p = fwd[q];
        int accept = p.second_index >= 0 &&
                     (p.first_distance == 0 ? p.second_distance > 0 : p.first_distance < r * p.second_distance);
        if (accept && rev) accept = rev[p.first_index].first_index == q;
        if (!accept) continue;
        idx_a[count] = q;
        idx_b[count] = p.first_index;
        const double d = binary ? p.first_distance : sqrt(p.first_distance);
        distance[count] = d > FLT_MAX ? INFINITY : (float)d;
        ++count;
    }
    if (forward_out) memcpy(forward_out, fwd, sizeof(orc_top2) * na);
    free(fwd); free(rev);
    return count;
}

/* ======================================================================
 * S5: MSAC homography B -> A (src/s5/s5_math.hpp:19-184; src/s5/s5.cu:42-212).
 * Pinned by: the reference's own s5_math.hpp compiled from its sources
 * (oracle/_ref), the known-answer recovery of tests/test_s5_cuda.cu:99-116,
 * sample distinctness, split-R == full-R, and an independent FP64 DLT.
 * ====================================================================== */

typedef struct { double ax, ay, bx, by; } orc_point;

static int finite_d(double x) { return x == x && x <= 1.7976931348623157e308 && x >= -1.7976931348623157e308; }

/* normalize, s5_math.hpp:30-42. */
int orc_s5_normalize(double* h) {
    double scale = 0.0;
    for (int i = 0; i < 9; ++i) {
        if (!finite_d(h[i])) return 0;
        if (fabs(h[i]) > scale) scale = fabs(h[i]);
    }
    if (!(scale > 0.0)) return 0;
    for (int i = 0; i < 9; ++i) h[i] /= scale;
    const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                       h[2] * (h[3] * h[7] - h[4] * h[6]);
    return finite_d(det) && fabs(det) > 1e-15;
}

/* residual_squared, s5_math.hpp:44-51. */
double orc_s5_residual(const double* h, orc_point p) {
    const double z = h[6] * p.bx + h[7] * p.by + h[8];
    if (!finite_d(z) || fabs(z) < 1e-15) return 1.7976931348623157e308;
    const double dx = (h[0] * p.bx + h[1] * p.by + h[2]) / z - p.ax;
    const double dy = (h[3] * p.bx + h[4] * p.by + h[5]) / z - p.ay;
    const double r = dx * dx + dy * dy;
    return finite_d(r) ? r : 1.7976931348623157e308;
}

/* SplitMix64 finalizer on a counter, s5_math.hpp:53-60 (public-domain generator). */
uint64_t orc_s5_random_word(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

/* sample_four, s5_math.hpp:62-79: k-th draw uniform over the n-k remaining
 * indices, mapped back by skipping the already drawn ones in sorted order. */
void orc_s5_sample_four(int n, uint64_t seed, int hypothesis, int* sample) {
    for (int k = 0; k < 4; ++k) {
        int v = (int)(orc_s5_random_word(seed ^ ((uint64_t)hypothesis * 4 + (uint64_t)k)) % (uint64_t)(n - k));
        for (int j = 0; j < k; ++j) {
            int smallest = n;
            for (int l = 0; l < k; ++l) {
                int rank = 0;
                for (int m = 0; m < k; ++m) rank += sample[m] < sample[l];
                if (rank == j) smallest = sample[l];
            }
            if (v >= smallest) ++v;
        }
        sample[k] = v;
    }
}

/* nondegenerate_four, s5_math.hpp:81-103. */
int orc_s5_nondegenerate(const orc_point* p) {
    for (int image = 0; image < 2; ++image) {
        double scale = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < i; ++j) {
                const double dx = image ? p[i].bx - p[j].bx : p[i].ax - p[j].ax;
                const double dy = image ? p[i].by - p[j].by : p[i].ay - p[j].ay;
                const double d = dx * dx + dy * dy;
                if (d > scale) scale = d;
            }
        if (!(scale > 0.0) || !finite_d(scale)) return 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j)
                for (int k = j + 1; k < 4; ++k) {
                    const double x1 = image ? p[j].bx - p[i].bx : p[j].ax - p[i].ax;
                    const double y1 = image ? p[j].by - p[i].by : p[j].ay - p[i].ay;
                    const double x2 = image ? p[k].bx - p[i].bx : p[k].ax - p[i].ax;
                    const double y2 = image ? p[k].by - p[i].by : p[k].ay - p[i].ay;
                    if (fabs(x1 * y2 - y1 * x2) <= 1e-8 * scale) return 0;
                }
    }
    return 1;
}

/* fit, s5_math.hpp:110-172: Hartley-normalised 8x9 normal equations solved by
 * Gauss-Jordan with partial pivoting; keep[i] (NULL = all) selects points. */
int orc_s5_fit(int n, const orc_point* pts, const uint8_t* keep, double* out) {
    double tax = 0, tay = 0, tbx = 0, tby = 0;
    int count = 0;
    for (int i = 0; i < n; ++i) {
        if (keep && !keep[i]) continue;
        tax += pts[i].ax; tay += pts[i].ay; tbx += pts[i].bx; tby += pts[i].by;
        ++count;
    }
    if (count < 4) return 0;
    tax /= count; tay /= count; tbx /= count; tby /= count;
    double va = 0.0, vb = 0.0;
    for (int i = 0; i < n; ++i) {
        if (keep && !keep[i]) continue;
        const orc_point p = pts[i];
        va += (p.ax - tax) * (p.ax - tax) + (p.ay - tay) * (p.ay - tay);
        vb += (p.bx - tbx) * (p.bx - tbx) + (p.by - tby) * (p.by - tby);
    }
    if (!(va > 0.0) || !(vb > 0.0) || !finite_d(va) || !finite_d(vb)) return 0;
    const double sa = sqrt(2.0 * count / va);
    const double sb = sqrt(2.0 * count / vb);
    double a[8][9];
    memset(a, 0, sizeof a);
    for (int i = 0; i < n; ++i) {
        if (keep && !keep[i]) continue;
        const orc_point p = pts[i];
        const double x = (p.bx - tbx) * sb, y = (p.by - tby) * sb;
        const double u = (p.ax - tax) * sa, v = (p.ay - tay) * sa;
        const double rows[2][9] = {{x, y, 1, 0, 0, 0, -u * x, -u * y, u}, {0, 0, 0, x, y, 1, -v * x, -v * y, v}};
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 9; ++c) a[r][c] += rows[0][r] * rows[0][c] + rows[1][r] * rows[1][c];
    }
    double max_entry = 0.0;
    for (int i = 0; i < 8; ++i)
        if (fabs(a[i][i]) > max_entry) max_entry = fabs(a[i][i]);
    for (int c = 0; c < 8; ++c) {
        int pivot = c;
        for (int r = c + 1; r < 8; ++r)
            if (fabs(a[r][c]) > fabs(a[pivot][c])) pivot = r;
        if (!finite_d(a[pivot][c]) || fabs(a[pivot][c]) <= 1e-12 * max_entry) return 0;
        for (int j = c; j < 9; ++j) { const double t = a[c][j]; a[c][j] = a[pivot][j]; a[pivot][j] = t; }
        const double divisor = a[c][c];
        for (int j = c; j < 9; ++j) a[c][j] /= divisor;
        for (int r = 0; r < 8; ++r) {
            if (r == c) continue;
            const double factor = a[r][c];
            for (int j = c; j < 9; ++j) a[r][j] -= factor * a[c][j];
        }
    }
    double h[9];
    for (int i = 0; i < 8; ++i) h[i] = a[i][8];
    h[8] = 1.0;
    const double inv_a[9] = {1.0 / sa, 0, tax, 0, 1.0 / sa, tay, 0, 0, 1};
    const double b[9] = {sb, 0, -sb * tbx, 0, sb, -sb * tby, 0, 0, 1};
    double tmp[9];
    mat3_mul(inv_a, h, tmp);
    mat3_mul(tmp, b, out);
    return orc_s5_normalize(out);
}

/* valid_domain, s5_math.hpp:178-184. */
int orc_s5_valid_domain(const double* h, int width, int height) {
    const double z[4] = {h[8], h[6] * (width - 1) + h[8], h[7] * (height - 1) + h[8],
                         h[6] * (width - 1) + h[7] * (height - 1) + h[8]};
    for (int i = 0; i < 4; ++i)
        if (!finite_d(z[i]) || fabs(z[i]) < 1e-15 || (z[i] > 0.0) != (z[0] > 0.0)) return 0;
    return 1;
}

typedef struct {
    double H[9];
    double H_s[9];
    int num_matches;
    int num_inliers;
    double inlier_ratio;
    double inlier_coverage;
    double reproj_residual;
    int status;
} orc_homography_result;

typedef struct {
    double h[9];
    double loss;
    int hypothesis;
    int valid;
} orc_model;

/* Persistent search state across R extensions (s5_internal.hpp:14-34). */
typedef struct {
    orc_model best;
    int evaluated;
    int status;
} orc_s5_state;

static int better(const orc_model* a, const orc_model* b) {
    return a->valid && (!b->valid || a->loss < b->loss || (a->loss == b->loss && a->hypothesis < b->hypothesis));
}

/* MSAC loss in the order of the reference score kernel (s5.cu:101-122):
 * 128 lane-strided partial sums, then a halving tree. */
double orc_s5_tree_loss(int nm, const orc_point* pts, const double* h, double t2) {
    double part[128];
    for (int l = 0; l < 128; ++l) {
        double v = 0.0;
        for (int i = l; i < nm; i += 128) {
            const double r = orc_s5_residual(h, pts[i]);
            v += r < t2 ? r : t2;
        }
        part[l] = v;
    }
    for (int stride = 64; stride > 0; stride /= 2)
        for (int l = 0; l < stride; ++l) part[l] += part[l + stride];
    return part[0];
}

/* serial_loss, s5.cu:139-146. */
double orc_s5_serial_loss(int nm, const orc_point* pts, const double* h, double t2) {
    double v = 0.0;
    for (int i = 0; i < nm; ++i) {
        const double r = orc_s5_residual(h, pts[i]);
        v += r < t2 ? r : t2;
    }
    return v;
}

/* s5::estimate on already-gathered correspondences pts[0..nm) (the Points
 * functor of s5.cu:15-21 reads keypoints through match indices; the caller
 * does that gather). state carries the raw search across R extensions;
 * previous_hypotheses == 0 restarts it. mask has nm bytes. */
int orc_s5_estimate(int nm, const orc_point* pts, const orc_scale_info* sa, const orc_scale_info* sb,
                    int hypotheses, int previous_hypotheses, int refinement_steps, double threshold,
                    uint64_t seed, orc_s5_state* state, uint8_t* mask, orc_homography_result* result) {
    double inverse_a[9];
    if (!mat3_inverse(sa->D, inverse_a)) return ORC_INVALID_ARGUMENT;
    if (previous_hypotheses == 0) {
        memset(&state->best, 0, sizeof state->best);
        state->evaluated = 0;
        state->status = ORC_OK;
        for (int i = 0; i < nm; ++i)
            if (!finite_d(pts[i].ax) || !finite_d(pts[i].ay) || !finite_d(pts[i].bx) || !finite_d(pts[i].by))
                state->status = ORC_INVALID_ARGUMENT;
        if (state->status == ORC_OK && nm < 4) state->status = ORC_NOT_ENOUGH_POINTS;
    } else if (state->evaluated != previous_hypotheses) {
        state->status = ORC_INVALID_ARGUMENT;
    }
    const double t2 = threshold * threshold;
    if (state->status == ORC_OK) {
        for (int hyp = previous_hypotheses; hyp < hypotheses; ++hyp) {
            orc_model m;
            memset(&m, 0, sizeof m);
            m.hypothesis = hyp;
            int idx[4];
            orc_s5_sample_four(nm, seed, hyp, idx);
            orc_point f[4];
            for (int j = 0; j < 4; ++j) f[j] = pts[idx[j]];
            m.valid = orc_s5_nondegenerate(f) && orc_s5_fit(4, f, NULL, m.h) &&
                      orc_s5_valid_domain(m.h, sb->work_w, sb->work_h);
            if (m.valid) {
                m.loss = orc_s5_tree_loss(nm, pts, m.h, t2);
                if (!finite_d(m.loss)) m.valid = 0;
            }
            if (better(&m, &state->best)) state->best = m;
        }
    }
    /* finish, s5.cu:148-212. */
    orc_homography_result r;
    memset(&r, 0, sizeof r);
    for (int i = 0; i < 9; ++i) r.H[i] = r.H_s[i] = NAN;
    r.reproj_residual = NAN;
    r.num_matches = nm;
    r.status = state->status;
    memset(mask, 0, (size_t)nm);
    if (state->status == ORC_OK && !state->best.valid) r.status = ORC_DEGENERATE;
    if (r.status == ORC_OK) {
        orc_model model = state->best;
        model.loss = orc_s5_serial_loss(nm, pts, model.h, t2);
        uint8_t* keep = (uint8_t*)malloc((size_t)nm);
        for (int step = 0; step < refinement_steps; ++step) {
            orc_model cand;
            memset(&cand, 0, sizeof cand);
            for (int i = 0; i < nm; ++i) keep[i] = orc_s5_residual(model.h, pts[i]) <= t2;
            cand.valid = orc_s5_fit(nm, pts, keep, cand.h) && orc_s5_valid_domain(cand.h, sb->work_w, sb->work_h);
            if (!cand.valid) break;
            cand.loss = orc_s5_serial_loss(nm, pts, cand.h, t2);
            if (!finite_d(cand.loss) || !(cand.loss < model.loss)) break;
            model = cand;
        }
        free(keep);
        double min_x = 1.7976931348623157e308, min_y = min_x, max_x = -min_x, max_y = -min_x, sum = 0.0;
        for (int i = 0; i < nm; ++i) {
            const double res = orc_s5_residual(model.h, pts[i]);
            if (res > t2) continue;
            mask[i] = 1;
            ++r.num_inliers;
            sum += res;
            min_x = fmin(min_x, pts[i].ax); max_x = fmax(max_x, pts[i].ax);
            min_y = fmin(min_y, pts[i].ay); max_y = fmax(max_y, pts[i].ay);
        }
        if (r.num_inliers < 4) {
            r.status = ORC_DEGENERATE;
            memset(mask, 0, (size_t)nm);
            r.num_inliers = 0;
        } else {
            r.inlier_ratio = (double)r.num_inliers / nm;
            r.reproj_residual = sqrt(sum / r.num_inliers);
            const double width = fmax(1.0, (double)(sa->work_w - 1));
            const double height = fmax(1.0, (double)(sa->work_h - 1));
            r.inlier_coverage = fmin(1.0, (max_x - min_x) / width) * fmin(1.0, (max_y - min_y) / height);
            for (int i = 0; i < 9; ++i) r.H_s[i] = model.h[i];
            double tmp[9];
            mat3_mul(inverse_a, model.h, tmp);
            mat3_mul(tmp, sb->D, r.H);
            if (!orc_s5_normalize(r.H) || !orc_s5_valid_domain(r.H, sb->orig_w, sb->orig_h)) {
                r.status = ORC_DEGENERATE;
                memset(mask, 0, (size_t)nm);
                r.num_inliers = 0;
                r.inlier_ratio = r.inlier_coverage = 0.0;
                r.reproj_residual = NAN;
                for (int i = 0; i < 9; ++i) r.H[i] = r.H_s[i] = NAN;
            }
        }
    }
    state->evaluated = hypotheses;
    *result = r;
    return ORC_OK;
}

/* prepare (s5.cu:42-83) + estimate on device-layout inputs: keypoint SoA,
 * match indices and counts. Validates counts, indices and finiteness the way
 * the reference's prepare kernel does, gathers the correspondences and runs
 * orc_s5_estimate. mask receives capacity_m bytes. */
int orc_s5_estimate_kp(const float* ax, const float* ay, int na, int cap_a, const float* bx, const float* by, int nb,
                       int cap_b, const int32_t* idx_a, const int32_t* idx_b, int nm, int cap_m,
                       const orc_scale_info* sa, const orc_scale_info* sb, int hypotheses, int previous_hypotheses,
                       int refinement_steps, double threshold, uint64_t seed, orc_s5_state* state, uint8_t* mask,
                       orc_homography_result* result) {
    int bad = na < 0 || na > cap_a || nb < 0 || nb > cap_b || nm < 0 || nm > cap_m;
    for (int i = 0; !bad && i < nm; ++i) bad = idx_a[i] < 0 || idx_a[i] >= na || idx_b[i] < 0 || idx_b[i] >= nb;
    if (cap_m > 0) memset(mask, 0, (size_t)cap_m);
    if (bad) {
        if (previous_hypotheses == 0 || state->status == ORC_OK) state->status = ORC_INVALID_ARGUMENT;
        memset(result, 0, sizeof *result);
        for (int i = 0; i < 9; ++i) result->H[i] = result->H_s[i] = NAN;
        result->reproj_residual = NAN;
        result->num_matches = nm >= 0 && nm <= cap_m ? nm : 0;
        result->status = state->status;
        state->evaluated = hypotheses;
        return ORC_OK;
    }
    orc_point* pts = (orc_point*)malloc(sizeof(orc_point) * (size_t)(nm > 0 ? nm : 1));
    for (int i = 0; i < nm; ++i) {
        pts[i].ax = ax[idx_a[i]]; pts[i].ay = ay[idx_a[i]];
        pts[i].bx = bx[idx_b[i]]; pts[i].by = by[idx_b[i]];
    }
    const int st = orc_s5_estimate(nm, pts, sa, sb, hypotheses, previous_hypotheses, refinement_steps, threshold,
                                   seed, state, mask, result);
    free(pts);
    return st;
}

/* ======================================================================
 * S6 geometry (src/s6/geometry.cpp:106-200), long double as the reference.
 * Pinned by: the reference's geometry.cpp compiled from its sources
 * (oracle/_ref); negative-origin canvas 7x4 @(-2,-1) (tests/test_s6_cuda.cu:140-165).
 * ====================================================================== */

typedef long double Real;

typedef struct {
    int width, height, origin_x, origin_y, width_a, height_a, width_b, height_b;
    double denominator_tolerance, snap_tolerance;
    double inverse[9];
} orc_geometry;

static Real maxr(Real a, Real b) { return a < b ? b : a; } /* std::max */
static Real minr(Real a, Real b) { return b < a ? b : a; } /* std::min */

static void rmul(const Real* a, const Real* b, Real* out) {
    Real t[9] = {0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) t[3 * r + c] += a[3 * r + k] * b[3 * k + c];
    memcpy(out, t, sizeof t);
}

static int rnormalize(Real* m) {
    Real largest = 0;
    for (int i = 0; i < 9; ++i) {
        if (!isfinite(m[i])) return 0;
        largest = maxr(largest, fabsl(m[i]));
    }
    if (largest == 0) return 0;
    for (int i = 0; i < 9; ++i) m[i] /= largest;
    return 1;
}

static Real rinf_norm(const Real* m) {
    Real norm = 0;
    for (int r = 0; r < 3; ++r) {
        Real s = 0;
        for (int c = 0; c < 3; ++c) s += fabsl(m[3 * r + c]);
        norm = maxr(norm, s);
    }
    return norm;
}

static int rinvert(const Real* input, Real tol, Real* inverse) {
    Real a[3][6] = {{0}};
    Real scale[3] = {0};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            a[r][c] = input[3 * r + c];
            scale[r] = maxr(scale[r], fabsl(a[r][c]));
        }
        if (scale[r] == 0) return 0;
        a[r][r + 3] = 1;
    }
    for (int col = 0; col < 3; ++col) {
        int pivot = col;
        Real best = 0;
        for (int r = col; r < 3; ++r) {
            const Real score = fabsl(a[r][col]) / scale[r];
            if (score > best) { best = score; pivot = r; }
        }
        if (!isfinite(best) || best <= tol) return 0;
        if (pivot != col) {
            for (int c = 0; c < 6; ++c) { Real t = a[pivot][c]; a[pivot][c] = a[col][c]; a[col][c] = t; }
            Real t = scale[pivot]; scale[pivot] = scale[col]; scale[col] = t;
        }
        const Real divisor = a[col][col];
        for (int c = 0; c < 6; ++c) a[col][c] /= divisor;
        for (int r = 0; r < 3; ++r) {
            if (r == col) continue;
            const Real factor = a[r][col];
            for (int c = 0; c < 6; ++c) a[r][c] -= factor * a[col][c];
        }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const Real v = a[r][c + 3];
            if (!isfinite(v)) return 0;
            inverse[3 * r + c] = v;
        }
    const Real condition = rinf_norm(input) * rinf_norm(inverse);
    return isfinite(condition) && condition * tol < 1;
}

static Real snap_integer(Real x, double snap_tol) {
    const Real rounded = roundl(x);
    const Real numerical_floor = (16 * DBL_EPSILON) * maxr((Real)1, fabsl(x));
    const Real tolerance = maxr((Real)snap_tol, numerical_floor);
    return fabsl(x - rounded) <= tolerance ? rounded : x;
}

int orc_s6_prepare_geometry(int wa, int ha, int wb, int hb, const double* h, double den_tol, double inv_tol,
                            double snap_tol, unsigned long long max_pixels, orc_geometry* out) {
    Real m[9] = {0};
    if (!out) return ORC_INVALID_ARGUMENT;
    if (h) for (int i = 0; i < 9; ++i) m[i] = h[i];
    memset(out, 0, sizeof *out);
    out->denominator_tolerance = 1e-12;
    out->snap_tolerance = 1e-10;
    if (!h || wa <= 0 || ha <= 0 || wb <= 0 || hb <= 0) return ORC_INVALID_ARGUMENT;
    if (!rnormalize(m)) return ORC_DEGENERATE;
    const Real bx = (Real)wb - 1, by = (Real)hb - 1;
    const Real dscale = fabsl(m[6]) * bx + fabsl(m[7]) * by + fabsl(m[8]);
    if (!(dscale > 0) || !isfinite(dscale)) return ORC_DEGENERATE;
    const Real dfloor = den_tol * dscale;
    Real xmin = 0, ymin = 0, xmax = (Real)wa - 1, ymax = (Real)ha - 1;
    int dsign = 0;
    for (int corner = 0; corner < 4; ++corner) {
        const Real x = (corner & 1) ? bx : 0, y = (corner & 2) ? by : 0;
        const Real w = m[6] * x + m[7] * y + m[8];
        if (!isfinite(w) || fabsl(w) <= dfloor) return ORC_DEGENERATE;
        const int sign = w > 0 ? 1 : -1;
        if (dsign && sign != dsign) return ORC_DEGENERATE;
        dsign = sign;
        const Real px = (m[0] * x + m[1] * y + m[2]) / w;
        const Real py = (m[3] * x + m[4] * y + m[5]) / w;
        if (!isfinite(px) || !isfinite(py)) return ORC_DEGENERATE;
        xmin = minr(xmin, px); ymin = minr(ymin, py);
        xmax = maxr(xmax, px); ymax = maxr(ymax, py);
    }
    xmin = floorl(snap_integer(xmin, snap_tol));
    ymin = floorl(snap_integer(ymin, snap_tol));
    xmax = ceill(snap_integer(xmax, snap_tol));
    ymax = ceill(snap_integer(ymax, snap_tol));
    const Real width = xmax - xmin + 1, height = ymax - ymin + 1;
    if (xmin < -2147483648.0L || ymin < -2147483648.0L || xmin > 2147483647.0L || ymin > 2147483647.0L ||
        !isfinite(width) || !isfinite(height) || width < 1 || height < 1 || width > 2147483647.0L ||
        height > 2147483647.0L)
        return ORC_INSUFFICIENT_WORKSPACE;
    const unsigned long long cw = (unsigned long long)width, ch = (unsigned long long)height;
    if (cw > max_pixels / ch) return ORC_INSUFFICIENT_WORKSPACE;
    const Real sa = maxr(maxr((Real)1, (Real)wa - 1), (Real)ha - 1);
    const Real sb = maxr(maxr((Real)1, bx), by);
    const Real axc = ((Real)wa - 1) / 2, ayc = ((Real)ha - 1) / 2;
    const Real na[9] = {1 / sa, 0, -axc / sa, 0, 1 / sa, -ayc / sa, 0, 0, 1};
    const Real nbi[9] = {sb, 0, bx / 2, 0, sb, by / 2, 0, 0, 1};
    Real t[9], nh[9], ninv[9], oinv[9];
    rmul(na, m, t);
    rmul(t, nbi, nh);
    if (!rnormalize(nh) || !rinvert(nh, inv_tol, ninv)) return ORC_DEGENERATE;
    rmul(nbi, ninv, t);
    rmul(t, na, oinv);
    if (!rnormalize(oinv)) return ORC_DEGENERATE;
    orc_geometry g;
    g.width = (int)width; g.height = (int)height;
    g.origin_x = (int)xmin; g.origin_y = (int)ymin;
    g.width_a = wa; g.height_a = ha; g.width_b = wb; g.height_b = hb;
    g.denominator_tolerance = den_tol;
    g.snap_tolerance = snap_tol;
    for (int i = 0; i < 9; ++i) {
        g.inverse[i] = (double)oinv[i];
        if (!isfinite(g.inverse[i])) return ORC_DEGENERATE;
    }
    *out = g;
    return ORC_OK;
}

/* ======================================================================
 * S6 render (src/s6/s6.cu:122-306, orchestration :370-433): FP64 inverse map,
 * bilinear, feather or Laplacian multiband with FP32 pyramids. Band counts
 * beyond the reference's 3 and 5 use the same depth-generic algorithm.
 * Filter taps, expansions and the band mix accumulate with fused
 * multiply-add (one rounding per tap), the contraction nvcc applies to
 * s6.cu's `value += tap * tap * source` by default.
 * Pinned by: golden vectors from the reference's FP64 NumPy renderer
 * (tests/golden, <= 1 LSB, mask exact), analytic cases of
 * tests/test_s6_cuda.cu:105-191 and tests/python/test_s6_reference.py:12-79.
 * ====================================================================== */

static int reflect101_l(long long x, long long n) {
    if (n == 1) return 0;
    const long long period = 2 * n - 2;
    x %= period;
    if (x < 0) x += period;
    return (int)(x < n ? x : period - x);
}

static float tap(int x) {
    const int t = x < 0 ? -x : x;
    return t == 0 ? 0.375f : (t == 1 ? 0.25f : 0.0625f);
}

static float expanded(const float* src, int cw, int ch, int channels, int x, int y, int c) {
    float sum = 0;
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101_l((long long)y + dy, 2LL * ch);
        if (yy & 1) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int xx = reflect101_l((long long)x + dx, 2LL * cw);
            if (xx & 1) continue;
            sum = fmaf(tap(dx) * tap(dy), src[((size_t)(yy / 2) * cw + xx / 2) * channels + c], sum);
        }
    }
    return 4 * sum;
}

static int center_coordinate(double* x, int n, double tol) {
    if (!isfinite(*x)) return 0;
    const double nearest = nearbyint(*x);
    const double roundoff = 16 * 2.2204460492503131e-16 * fmax(1.0, fabs(*x));
    if (fabs(*x - nearest) <= fmax(tol, roundoff)) *x = nearest;
    return *x >= 0 && *x <= n - 1.0;
}

static int sample_rgb(const uint8_t* img, int w, int h, int pitch, double x, double y, float* rgb, double* score,
                      double tol) {
    if (!center_coordinate(&x, w, tol) || !center_coordinate(&y, h, tol)) return 0;
    const int x0 = (int)floor(x), y0 = (int)floor(y);
    const int x1 = x0 < w - 1 ? x0 + 1 : x0, y1 = y0 < h - 1 ? y0 + 1 : y0;
    const double ax = x - x0, ay = y - y0;
    const uint8_t* r0 = img + (size_t)y0 * pitch;
    const uint8_t* r1 = img + (size_t)y1 * pitch;
    for (int c = 0; c < 3; ++c) {
        const double v0 = (1 - ax) * r0[3 * x0 + c] + ax * r0[3 * x1 + c];
        const double v1 = (1 - ax) * r1[3 * x0 + c] + ax * r1[3 * x1 + c];
        rgb[c] = (float)(((1 - ay) * v0 + ay * v1) / 255.0);
    }
    *score = fmin(fmin((x + 0.5) / w, (w - 0.5 - x) / w), fmin((y + 0.5) / h, (h - 0.5 - y) / h));
    return 1;
}

static uint8_t quantize(float x) { return (uint8_t)floorf(fminf(255.0f, fmaxf(0.0f, 255.0f * x)) + 0.5f); }

static void reduce_level(const float* src, int w, int h, float* dst, int dw, int dh, int channels) {
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x)
            for (int c = 0; c < channels; ++c) {
                float v = 0;
                for (int dy = -2; dy <= 2; ++dy) {
                    const int yy = reflect101_l(2LL * y + dy, h);
                    for (int dx = -2; dx <= 2; ++dx) {
                        const int xx = reflect101_l(2LL * x + dx, w);
                        v = fmaf(tap(dx) * tap(dy), src[((size_t)yy * w + xx) * channels + c], v);
                    }
                }
                dst[((size_t)y * dw + x) * channels + c] = v;
            }
}

/* Number of reductions requested by a blend mode: 0 feather, 3, 5, 6, 7. */
int orc_s6_requested_reductions(int mode) {
    static const int req[5] = {0, 3, 5, 6, 7};
    return mode >= 0 && mode < 5 ? req[mode] : -1;
}

/* effective_reductions, s6.cu:309-318. */
int orc_s6_effective_reductions(int cw, int ch, int mode) {
    const int requested = orc_s6_requested_reductions(mode);
    int r = 0, w = cw, h = ch;
    while ((w > 1 || h > 1) && r < requested) { w = w / 2 + w % 2; h = h / 2 + h % 2; ++r; }
    return r;
}

/* render: out is width*height*3 (unpitched), mask width*height; stats[2] =
 * union, overlap pixel counts. a/b are pitched RGB8. */
int orc_s6_render(const uint8_t* a, int pitch_a, const uint8_t* b, int pitch_b, const orc_geometry* g, int mode,
                  uint8_t* out, uint8_t* mask, unsigned long long* stats) {
    const int requested = orc_s6_requested_reductions(mode);
    if (requested < 0) return ORC_INVALID_ARGUMENT;
    const int W = g->width, H = g->height;
    const size_t n = (size_t)W * H;
    float *ca = NULL, *cb = NULL, *va = NULL, *vb = NULL, *ua = NULL, *ub = NULL;
    if (mode != 0) {
        ca = (float*)malloc(n * 12); cb = (float*)malloc(n * 12);
        va = (float*)malloc(n * 4); vb = (float*)malloc(n * 4);
        ua = (float*)malloc(n * 4); ub = (float*)malloc(n * 4);
    }
    unsigned long long uni = 0, ovl = 0;
    for (size_t p = 0; p < n; ++p) {
        const int x = (int)(p % W), y = (int)(p / W);
        const double xa = (double)x + g->origin_x, ya = (double)y + g->origin_y;
        float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
        double sa = 0, sb = 0;
        const int ma = sample_rgb(a, g->width_a, g->height_a, pitch_a, xa, ya, ra, &sa, g->snap_tolerance);
        const double* inv = g->inverse;
        const double den = inv[6] * xa + inv[7] * ya + inv[8];
        const double dsc = fabs(inv[6] * xa) + fabs(inv[7] * ya) + fabs(inv[8]);
        int mb = 0;
        if (isfinite(den) && dsc > 0 && fabs(den) > g->denominator_tolerance * dsc) {
            const double xb = (inv[0] * xa + inv[1] * ya + inv[2]) / den;
            const double yb = (inv[3] * xa + inv[4] * ya + inv[5]) / den;
            mb = sample_rgb(b, g->width_b, g->height_b, pitch_b, xb, yb, rb, &sb, g->snap_tolerance);
        }
        mask[p] = ma || mb ? 1 : 0;
        uni += ma || mb;
        ovl += ma && mb;
        if (ca) {
            va[p] = ma ? 1.0f : 0.0f;
            vb[p] = mb ? 1.0f : 0.0f;
            ua[p] = ma && (!mb || sa >= sb) ? 1.0f : 0.0f;
            ub[p] = mb && (!ma || sb > sa) ? 1.0f : 0.0f;
            for (int c = 0; c < 3; ++c) { ca[3 * p + c] = ra[c]; cb[3 * p + c] = rb[c]; }
        } else {
            for (int c = 0; c < 3; ++c) {
                const float color = ma && mb ? (float)((sa * ra[c] + sb * rb[c]) / (sa + sb)) : (ma ? ra[c] : rb[c]);
                out[3 * p + c] = quantize(color);
            }
        }
    }
    if (stats) { stats[0] = uni; stats[1] = ovl; }
    if (mode == 0) return ORC_OK;

    /* Pyramid layout: level 0 = canvas, halving with ceil to 1x1 (s6.cu:60-88). */
    int lw[40], lh[40], top = 0;
    lw[0] = W; lh[0] = H;
    while (lw[top] > 1 || lh[top] > 1) {
        lw[top + 1] = lw[top] / 2 + lw[top] % 2;
        lh[top + 1] = lh[top] / 2 + lh[top] % 2;
        ++top;
    }
    const int blend_top = requested < top ? requested : top;
    float *A[40], *B[40], *VA[40], *VB[40], *UA[40], *UB[40], *Z[40];
    A[0] = ca; B[0] = cb; VA[0] = va; VB[0] = vb; UA[0] = ua; UB[0] = ub;
    for (int l = 1; l <= top; ++l) {
        const size_t m = (size_t)lw[l] * lh[l];
        A[l] = (float*)malloc(m * 12); B[l] = (float*)malloc(m * 12);
        VA[l] = (float*)malloc(m * 4); VB[l] = (float*)malloc(m * 4);
        UA[l] = l <= blend_top ? (float*)malloc(m * 4) : NULL;
        UB[l] = l <= blend_top ? (float*)malloc(m * 4) : NULL;
    }
    for (int l = 0; l <= blend_top; ++l) Z[l] = (float*)malloc((size_t)lw[l] * lh[l] * 12);
    /* Pull raw numerators and validity to 1x1 (s6.cu:389-398). */
    for (int l = 1; l <= top; ++l) {
        reduce_level(A[l - 1], lw[l - 1], lh[l - 1], A[l], lw[l], lh[l], 3);
        reduce_level(VA[l - 1], lw[l - 1], lh[l - 1], VA[l], lw[l], lh[l], 1);
        reduce_level(B[l - 1], lw[l - 1], lh[l - 1], B[l], lw[l], lh[l], 3);
        reduce_level(VB[l - 1], lw[l - 1], lh[l - 1], VB[l], lw[l], lh[l], 1);
    }
    /* normalize_color on every level (s6.cu:399-404). */
    for (int l = 0; l <= top; ++l) {
        const size_t m = (size_t)lw[l] * lh[l];
        for (size_t p = 0; p < m; ++p)
            for (int c = 0; c < 3; ++c) {
                A[l][3 * p + c] = VA[l][p] > 0 ? A[l][3 * p + c] / VA[l][p] : 0;
                B[l][3 * p + c] = VB[l][p] > 0 ? B[l][3 * p + c] / VB[l][p] : 0;
            }
    }
    /* push_missing top-down (s6.cu:405-410). */
    for (int l = top - 1; l >= 0; --l) {
        for (int y = 0; y < lh[l]; ++y)
            for (int x = 0; x < lw[l]; ++x) {
                const size_t p = (size_t)y * lw[l] + x;
                for (int c = 0; c < 3; ++c) {
                    if (!(VA[l][p] > 0)) A[l][3 * p + c] = expanded(A[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c);
                    if (!(VB[l][p] > 0)) B[l][3 * p + c] = expanded(B[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c);
                }
            }
    }
    /* Gaussian pyramids of the filled images and ownership (s6.cu:411-419). */
    for (int l = 1; l <= blend_top; ++l) {
        reduce_level(A[l - 1], lw[l - 1], lh[l - 1], A[l], lw[l], lh[l], 3);
        reduce_level(B[l - 1], lw[l - 1], lh[l - 1], B[l], lw[l], lh[l], 3);
        reduce_level(UA[l - 1], lw[l - 1], lh[l - 1], UA[l], lw[l], lh[l], 1);
        reduce_level(UB[l - 1], lw[l - 1], lh[l - 1], UB[l], lw[l], lh[l], 1);
    }
    /* mix_band (s6.cu:271-284, :420-425). */
    for (int l = 0; l <= blend_top; ++l) {
        const int coarse = l < blend_top;
        for (int y = 0; y < lh[l]; ++y)
            for (int x = 0; x < lw[l]; ++x) {
                const size_t p = (size_t)y * lw[l] + x;
                const float total = UA[l][p] + UB[l][p];
                for (int c = 0; c < 3; ++c) {
                    const float la = A[l][3 * p + c] - (coarse ? expanded(A[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c) : 0);
                    const float lb = B[l][3 * p + c] - (coarse ? expanded(B[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c) : 0);
                    Z[l][3 * p + c] = total > 0 ? fmaf(UA[l][p], la, UB[l][p] * lb) / total : 0;
                }
            }
    }
    /* reconstruct bottom-up collapse (s6.cu:285-293, :426-429). */
    for (int l = blend_top - 1; l >= 0; --l)
        for (int y = 0; y < lh[l]; ++y)
            for (int x = 0; x < lw[l]; ++x) {
                const size_t p = (size_t)y * lw[l] + x;
                for (int c = 0; c < 3; ++c) Z[l][3 * p + c] += expanded(Z[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c);
            }
    /* finish (s6.cu:294-306). */
    for (size_t p = 0; p < n; ++p)
        for (int c = 0; c < 3; ++c) {
            const float color = va[p] > 0 && vb[p] > 0 ? Z[0][3 * p + c]
                                : (va[p] > 0 ? ca[3 * p + c] : (vb[p] > 0 ? cb[3 * p + c] : 0));
            out[3 * p + c] = quantize(color);
        }
    for (int l = 0; l <= top; ++l) {
        free(A[l]); free(B[l]); free(VA[l]); free(VB[l]);
        if (l <= blend_top) { free(UA[l]); free(UB[l]); free(Z[l]); }
    }
    return ORC_OK;
}

/* ======================================================================
 * S6 over a chain of frames (no counterpart in the reference, which renders
 * pairs): frame 0 sits on the canvas at an integer offset, as A does, and
 * every other frame k is sampled through its own inverse map g[k].inverse
 * (frame-0-original -> frame-k-original). The pair algorithm of
 * orc_s6_render generalised to n frames:
 *   - ownership goes to the valid frame with the largest border score, the
 *     lowest index on ties (for two frames: A if s_a >= s_b);
 *   - feather: sum_k s_k c_k / sum_k s_k over the valid frames, summed in
 *     frame order;
 *   - multiband: every frame's pull-push extension and Gaussian pyramid as
 *     in the pair render, ownership pyramids per frame, and the band mix
 *     sum_k U_k L_k / sum_k U_k with the numerator fused from the last frame
 *     down (for two frames: fmaf(U_a, L_a, U_b L_b)) and the denominator
 *     summed in frame order;
 *   - single-source pixels keep their frame's colour, overlap (two or more
 *     frames valid) takes the collapsed mix.
 * For n = 2 every expression is the pair render's, so the result equals
 * orc_s6_render bit for bit (checked by the tests); beyond two frames the
 * blend is this build's definition. All g[k] share the canvas (width,
 * height, origin) and frame 0's size (width_a, height_a); g[k].width_b /
 * height_b give frame k's size. out / mask as in orc_s6_render.
 * ====================================================================== */
/* Canvas of a chain of n frames (no counterpart in the reference): h holds
 * the n - 1 adjacent homographies, h[k] mapping frame k + 1 to frame k.
 * Frame k maps to frame 0 through G_k = G_{k-1} h[k-1] (G_0 = I), each
 * product normalised by its largest coefficient; frame k's pairwise
 * geometry with frame 0 is orc_s6_prepare_geometry of G_k (unbounded
 * budget), and the chain canvas is the union of those boxes, at most
 * max_pixels. out[k] gets the common canvas, frame 0's size, frame k's size
 * and frame k's inverse map (the identity for k = 0). */
int orc_s6_prepare_chain(int n, const int* w, const int* hgt, const double* h, double den_tol, double inv_tol,
                         double snap_tol, unsigned long long max_pixels, orc_geometry* out) {
    if (n < 1 || n > 64) return ORC_INVALID_ARGUMENT;
    double G[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    long long x0 = 0, y0 = 0, x1 = w[0], y1 = hgt[0];
    for (int k = 0; k < n; ++k) {
        memset(&out[k], 0, sizeof out[k]);
        if (w[k] <= 0 || hgt[k] <= 0) return ORC_INVALID_ARGUMENT;
    }
    for (int k = 1; k < n; ++k) {
        double P[9], peak = 0;
        const double* a = h + 9 * (k - 1);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
                for (int i = 0; i < 3; ++i) acc += G[3 * r + i] * a[3 * i + c];
                P[3 * r + c] = acc;
            }
        for (int i = 0; i < 9; ++i) {
            if (!isfinite(P[i])) return ORC_DEGENERATE;
            peak = fmax(peak, fabs(P[i]));
        }
        if (peak == 0) return ORC_DEGENERATE;
        for (int i = 0; i < 9; ++i) G[i] = P[i] / peak;
        const int st = orc_s6_prepare_geometry(w[0], hgt[0], w[k], hgt[k], G, den_tol, inv_tol, snap_tol,
                                               ~0ull, &out[k]);
        if (st != ORC_OK) return st;
        x0 = out[k].origin_x < x0 ? out[k].origin_x : x0;
        y0 = out[k].origin_y < y0 ? out[k].origin_y : y0;
        x1 = out[k].origin_x + (long long)out[k].width > x1 ? out[k].origin_x + (long long)out[k].width : x1;
        y1 = out[k].origin_y + (long long)out[k].height > y1 ? out[k].origin_y + (long long)out[k].height : y1;
    }
    const unsigned long long cw = (unsigned long long)(x1 - x0), ch = (unsigned long long)(y1 - y0);
    if (x1 - x0 > 2147483647LL || y1 - y0 > 2147483647LL || cw * ch > max_pixels) return ORC_INSUFFICIENT_WORKSPACE;
    for (int k = 0; k < n; ++k) {
        out[k].width = (int)cw;
        out[k].height = (int)ch;
        out[k].origin_x = (int)x0;
        out[k].origin_y = (int)y0;
        out[k].width_a = w[0];
        out[k].height_a = hgt[0];
        out[k].width_b = w[k];
        out[k].height_b = hgt[k];
        out[k].denominator_tolerance = den_tol;
        out[k].snap_tolerance = snap_tol;
        if (k == 0) {
            memset(out[0].inverse, 0, sizeof out[0].inverse);
            out[0].inverse[0] = out[0].inverse[4] = out[0].inverse[8] = 1;
        }
    }
    return ORC_OK;
}

int orc_s6_render_chain(int n, const uint8_t* const* img, const int* pitch, const orc_geometry* g, int mode,
                        uint8_t* out, uint8_t* mask, unsigned long long* stats) {
    const int requested = orc_s6_requested_reductions(mode);
    if (requested < 0 || n < 1 || n > 64) return ORC_INVALID_ARGUMENT;
    const int W = g[0].width, H = g[0].height;
    const size_t np = (size_t)W * H;
    float* c0 = (float*)calloc(np * 3 * (size_t)n, sizeof(float));  /* level-0 colours, per frame */
    float* v0 = (float*)calloc(np * (size_t)n, sizeof(float));      /* level-0 validity */
    float* u0 = (float*)calloc(np * (size_t)n, sizeof(float));      /* level-0 ownership */
    int* count = (int*)calloc(np, sizeof(int));
    unsigned long long uni = 0, ovl = 0;
    for (size_t p = 0; p < np; ++p) {
        const int x = (int)(p % W), y = (int)(p / W);
        const double xa = (double)x + g[0].origin_x, ya = (double)y + g[0].origin_y;
        double best = 0, fs = 0, fc[3] = {0, 0, 0};
        int owner = -1, valid = 0;
        float single[3] = {0, 0, 0};
        for (int k = 0; k < n; ++k) {
            float rgb[3] = {0, 0, 0};
            double s = 0;
            int m = 0;
            if (k == 0) {
                m = sample_rgb(img[0], g[0].width_a, g[0].height_a, pitch[0], xa, ya, rgb, &s, g[0].snap_tolerance);
            } else {
                const double* inv = g[k].inverse;
                const double den = inv[6] * xa + inv[7] * ya + inv[8];
                const double dsc = fabs(inv[6] * xa) + fabs(inv[7] * ya) + fabs(inv[8]);
                if (isfinite(den) && dsc > 0 && fabs(den) > g[k].denominator_tolerance * dsc) {
                    const double xb = (inv[0] * xa + inv[1] * ya + inv[2]) / den;
                    const double yb = (inv[3] * xa + inv[4] * ya + inv[5]) / den;
                    m = sample_rgb(img[k], g[k].width_b, g[k].height_b, pitch[k], xb, yb, rgb, &s,
                                   g[k].snap_tolerance);
                }
            }
            if (!m) continue;
            ++valid;
            if (owner < 0 || s > best) {
                owner = k;
                best = s;
            }
            for (int c = 0; c < 3; ++c) {
                c0[((size_t)k * np + p) * 3 + c] = rgb[c];
                fc[c] = valid == 1 ? s * rgb[c] : fc[c] + s * rgb[c];
                single[c] = rgb[c];
            }
            fs = valid == 1 ? s : fs + s;
            v0[(size_t)k * np + p] = 1.0f;
        }
        if (owner >= 0) u0[(size_t)owner * np + p] = 1.0f;
        count[p] = valid;
        mask[p] = valid ? 1 : 0;
        uni += valid > 0;
        ovl += valid > 1;
        if (mode == 0)
            for (int c = 0; c < 3; ++c)
                out[3 * p + c] = quantize(valid > 1 ? (float)(fc[c] / fs) : single[c]);
    }
    if (stats) { stats[0] = uni; stats[1] = ovl; }
    if (mode == 0) {
        free(c0); free(v0); free(u0); free(count);
        return ORC_OK;
    }
    int lw[40], lh[40], top = 0;
    lw[0] = W; lh[0] = H;
    while (lw[top] > 1 || lh[top] > 1) {
        lw[top + 1] = lw[top] / 2 + lw[top] % 2;
        lh[top + 1] = lh[top] / 2 + lh[top] % 2;
        ++top;
    }
    const int bt = requested < top ? requested : top;
    /* Per frame k and level l: colour C (3), validity V, ownership U (l <= bt). */
    float** C = (float**)calloc((size_t)n * 40, sizeof(float*));
    float** V = (float**)calloc((size_t)n * 40, sizeof(float*));
    float** U = (float**)calloc((size_t)n * 40, sizeof(float*));
    float* Z[40];
    for (int k = 0; k < n; ++k) {
        C[k * 40] = c0 + (size_t)k * np * 3;
        V[k * 40] = v0 + (size_t)k * np;
        U[k * 40] = u0 + (size_t)k * np;
        for (int l = 1; l <= top; ++l) {
            const size_t m = (size_t)lw[l] * lh[l];
            C[k * 40 + l] = (float*)malloc(m * 12);
            V[k * 40 + l] = (float*)malloc(m * 4);
            U[k * 40 + l] = l <= bt ? (float*)malloc(m * 4) : NULL;
        }
    }
    for (int l = 0; l <= bt; ++l) Z[l] = (float*)malloc((size_t)lw[l] * lh[l] * 12);
    for (int k = 0; k < n; ++k) {
        float** Ck = C + k * 40;
        float** Vk = V + k * 40;
        float** Uk = U + k * 40;
        for (int l = 1; l <= top; ++l) {
            reduce_level(Ck[l - 1], lw[l - 1], lh[l - 1], Ck[l], lw[l], lh[l], 3);
            reduce_level(Vk[l - 1], lw[l - 1], lh[l - 1], Vk[l], lw[l], lh[l], 1);
        }
        for (int l = 0; l <= top; ++l) {
            const size_t m = (size_t)lw[l] * lh[l];
            for (size_t p = 0; p < m; ++p)
                for (int c = 0; c < 3; ++c) Ck[l][3 * p + c] = Vk[l][p] > 0 ? Ck[l][3 * p + c] / Vk[l][p] : 0;
        }
        for (int l = top - 1; l >= 0; --l)
            for (int y = 0; y < lh[l]; ++y)
                for (int x = 0; x < lw[l]; ++x) {
                    const size_t p = (size_t)y * lw[l] + x;
                    if (!(Vk[l][p] > 0))
                        for (int c = 0; c < 3; ++c) Ck[l][3 * p + c] = expanded(Ck[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c);
                }
        for (int l = 1; l <= bt; ++l) {
            reduce_level(Ck[l - 1], lw[l - 1], lh[l - 1], Ck[l], lw[l], lh[l], 3);
            reduce_level(Uk[l - 1], lw[l - 1], lh[l - 1], Uk[l], lw[l], lh[l], 1);
        }
    }
    for (int l = 0; l <= bt; ++l) {
        const int coarse = l < bt;
        for (int y = 0; y < lh[l]; ++y)
            for (int x = 0; x < lw[l]; ++x) {
                const size_t p = (size_t)y * lw[l] + x;
                float total = U[l][p];
                for (int k = 1; k < n; ++k) total += U[k * 40 + l][p];
                for (int c = 0; c < 3; ++c) {
                    float acc = 0;
                    for (int k = n - 1; k >= 0; --k) {
                        const float* Ck = C[k * 40 + l];
                        const float lk = Ck[3 * p + c] -
                                         (coarse ? expanded(C[k * 40 + l + 1], lw[l + 1], lh[l + 1], 3, x, y, c) : 0);
                        const float uk = U[k * 40 + l][p];
                        acc = k == n - 1 ? uk * lk : fmaf(uk, lk, acc);
                    }
                    Z[l][3 * p + c] = total > 0 ? acc / total : 0;
                }
            }
    }
    for (int l = bt - 1; l >= 0; --l)
        for (int y = 0; y < lh[l]; ++y)
            for (int x = 0; x < lw[l]; ++x) {
                const size_t p = (size_t)y * lw[l] + x;
                for (int c = 0; c < 3; ++c) Z[l][3 * p + c] += expanded(Z[l + 1], lw[l + 1], lh[l + 1], 3, x, y, c);
            }
    for (size_t p = 0; p < np; ++p) {
        int only = -1;
        for (int k = 0; k < n && only < 0; ++k)
            if (v0[(size_t)k * np + p] > 0) only = k;
        for (int c = 0; c < 3; ++c) {
            const float color = count[p] > 1 ? Z[0][3 * p + c] : (only >= 0 ? c0[((size_t)only * np + p) * 3 + c] : 0);
            out[3 * p + c] = quantize(color);
        }
    }
    for (int k = 0; k < n; ++k)
        for (int l = 1; l <= top; ++l) {
            free(C[k * 40 + l]);
            free(V[k * 40 + l]);
            free(U[k * 40 + l]);
        }
    for (int l = 0; l <= bt; ++l) free(Z[l]);
    free(C); free(V); free(U);
    free(c0); free(v0); free(u0); free(count);
    return ORC_OK;
}

/* ======================================================================
 * Controller observations: scene and gate features
 * (src/controller/features.cu:11-171, src/controller/relational_features.cu:6-190)
 *
 * Every FP64 statistic is summed in the reference's order: N partial sums
 * (partial t takes elements t, t + N, t + 2N, ... in increasing order), then
 * the N partials added pairwise (v[t] += v[t + N/2] for t < N/2, then N/4,
 * ... 1). N is 128, and 64 for the 7x7 patches.
 *
 * Pinning: closed forms in tests/test_features.py (constant, ramp and
 * checkerboard images give exact means, deviations and gradients; identical
 * images give NCC 1 at zero shift and zero photometric error under H = I;
 * a pure translation puts the NCC peak at its shift; hand-counted inlier
 * grids give the occupancy, largest-bin and entropy values; a brute-force
 * double sum agrees with the tree sums to rounding). log and hypot come from
 * the C library; their last-bit agreement with the device's is unpinned.
 * ====================================================================== */

typedef struct {
    const uint8_t* data;
    int w, h, pitch, format; /* format: 0 RGB8, 1 GRAY8 (src/common/types.hpp) */
} orc_view;

#define OBS_N 128
#define OBS_PATCH_N 64
#define OBS_DEVICE_ERROR 3

/* features.cu:12-16 */
static double obs_gray(const orc_view* im, int x, int y) {
    const uint8_t* p = im->data + (size_t)y * im->pitch + (size_t)x * (im->format == 0 ? 3 : 1);
    if (im->format == 1) return p[0] / 255.;
    return (9798. * p[0] + 19235. * p[1] + 3735. * p[2]) / (32768. * 255.);
}

/* features.cu:17-23 */
static double obs_sample(const orc_view* im, double x, double y) {
    const int ix = (int)x, iy = (int)y;
    const int jx = ix + 1 < im->w - 1 ? ix + 1 : im->w - 1, jy = iy + 1 < im->h - 1 ? iy + 1 : im->h - 1;
    const double dx = x - ix, dy = y - iy;
    return (1 - dy) * ((1 - dx) * obs_gray(im, ix, iy) + dx * obs_gray(im, jx, iy)) +
           dy * ((1 - dx) * obs_gray(im, ix, jy) + dx * obs_gray(im, jx, jy));
}

/* features.cu:24-33: pairwise tree over n partials, n a power of two. */
static double obs_tree(double* v, int n) {
    for (int step = n / 2; step; step /= 2)
        for (int t = 0; t < step; ++t) v[t] += v[t + step];
    return v[0];
}

/* relational_features.cu:17-20 */
static double obs_correlation(double a, double b, double aa, double bb, double ab, double n) {
    const double va = fmax(0., aa - a * a / n), vb = fmax(0., bb - b * b / n);
    return fmin(1., fmax(-1., (ab - a * b / n) / fmax(1e-12, sqrt(va * vb))));
}

/* Adjugate inverse, features.cu:115-123. */
static int obs_inverse(const double* h, double* out) {
    const double d = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                     h[2] * (h[3] * h[7] - h[4] * h[6]);
    if (!finite_d(d) || fabs(d) < 1e-18) return 0;
    const double adj[9] = {h[4] * h[8] - h[5] * h[7], h[2] * h[7] - h[1] * h[8], h[1] * h[5] - h[2] * h[4],
                           h[5] * h[6] - h[3] * h[8], h[0] * h[8] - h[2] * h[6], h[2] * h[3] - h[0] * h[5],
                           h[3] * h[7] - h[4] * h[6], h[1] * h[6] - h[0] * h[7], h[0] * h[4] - h[1] * h[3]};
    for (int i = 0; i < 9; ++i) {
        out[i] = adj[i] / d;
        if (!finite_d(out[i])) return 0;
    }
    return 1;
}

/* relational_features.cu:155-159 */
static int obs_valid_image(const orc_view* im) {
    return im->data && im->w >= 2 && im->h >= 2 && im->pitch > 0 && (im->format == 0 || im->format == 1) &&
           im->pitch >= im->w * (im->format == 0 ? 3 : 1);
}

/* scene_kernel, features.cu:34-62: region 0/1 whole-image gradient of A/B,
 * regions 2..9 quadrant (mean, std, gradient). */
static void obs_scene_region(const orc_view* a, const orc_view* b, int region, double* out) {
    const int image = region < 2 ? region : (region - 2) / 4;
    const orc_view* im = image ? b : a;
    const int tile = region < 2 ? -1 : (region - 2) % 4;
    const int x0 = tile < 0 ? 0 : (tile % 2) * im->w / 2, y0 = tile < 0 ? 0 : (tile / 2) * im->h / 2;
    const int x1 = tile < 0 ? im->w : ((tile % 2) + 1) * im->w / 2;
    const int y1 = tile < 0 ? im->h : ((tile / 2) + 1) * im->h / 2;
    const int w = x1 - x0, h = y1 - y0;
    double sum[OBS_N], sq[OBS_N], grad[OBS_N];
    for (int t = 0; t < OBS_N; ++t) {
        double s = 0, q = 0, g = 0;
        for (int i = t; i < w * h; i += OBS_N) {
            const int x = x0 + i % w, y = y0 + i / w;
            const double v = obs_gray(im, x, y);
            s += v;
            q += v * v;
            if (x + 1 < x1) g += fabs(v - obs_gray(im, x + 1, y));
            if (y + 1 < y1) g += fabs(v - obs_gray(im, x, y + 1));
        }
        sum[t] = s;
        sq[t] = q;
        grad[t] = g;
    }
    const double total = obs_tree(sum, OBS_N), squares = obs_tree(sq, OBS_N), gradient = obs_tree(grad, OBS_N);
    const double mean = total / (w * h), edges = h * (w - 1) + w * (h - 1);
    if (tile < 0) {
        out[region] = edges ? gradient / edges : 0.;
    } else {
        double* o = out + 2 + image * 12 + tile * 3;
        o[0] = mean;
        o[1] = sqrt(fmax(0., squares / (w * h) - mean * mean));
        o[2] = edges ? gradient / edges : 0.;
    }
}

/* pair_kernel and pair_finish, relational_features.cu:27-49: NCC of a 32x32
 * grid of A against B shifted by (dx, dy) in {-3..3}/8; out[0..5]. */
static void obs_pair_relations(const orc_view* a, const orc_view* b, double* out) {
    double score[49], overlap[49];
    for (int shift = 0; shift < 49; ++shift) {
        const double dx = (shift % 7 - 3) / 8., dy = (shift / 7 - 3) / 8.;
        double p[6][OBS_N];
        for (int t = 0; t < OBS_N; ++t) {
            double n = 0, sa = 0, sb = 0, aa = 0, bb = 0, ab = 0;
            for (int i = t; i < 1024; i += OBS_N) {
                const double x = (i % 32) / 31., y = (i / 32) / 31., u = x + dx, v = y + dy;
                if (u < 0 || u > 1 || v < 0 || v > 1) continue;
                const double av = obs_sample(a, x * (a->w - 1), y * (a->h - 1));
                const double bv = obs_sample(b, u * (b->w - 1), v * (b->h - 1));
                ++n;
                sa += av;
                sb += bv;
                aa += av * av;
                bb += bv * bv;
                ab += av * bv;
            }
            p[0][t] = n;
            p[1][t] = sa;
            p[2][t] = sb;
            p[3][t] = aa;
            p[4][t] = bb;
            p[5][t] = ab;
        }
        double s[6];
        for (int k = 0; k < 6; ++k) s[k] = obs_tree(p[k], OBS_N);
        score[shift] = obs_correlation(s[1], s[2], s[3], s[4], s[5], s[0]);
        overlap[shift] = s[0] / 1024.;
    }
    int best = 0;
    for (int j = 1; j < 49; ++j)
        if (score[j] > score[best]) best = j;
    double second = -2;
    for (int j = 0; j < 49; ++j)
        if (j != best) second = fmax(second, score[j]);
    out[0] = score[best];
    out[1] = score[best] - second;
    out[2] = (best % 7 - 3) / 8.;
    out[3] = (best / 7 - 3) / 8.;
    out[4] = score[24];
    out[5] = overlap[best];
}

/* extract_scene, features.cu:125-135: 32 doubles. Returns 0 (OK),
 * 1 (INVALID_ARGUMENT) or 3 (the reference's CUDA_ERROR for images the
 * pair relations reject). */
int orc_scene_features(const uint8_t* da, int wa, int ha, int pa, int fa, const uint8_t* db, int wb, int hb, int pb,
                       int fb, int relations, double* out) {
    const orc_view a = {da, wa, ha, pa, fa}, b = {db, wb, hb, pb, fb};
    if (!da || !db || !out || wa < 2 || wb < 2 || ha < 2 || hb < 2 || fa != 1 || fb != 1) return ORC_INVALID_ARGUMENT;
    for (int region = 0; region < 10; ++region) obs_scene_region(&a, &b, region, out);
    for (int k = 26; k < 32; ++k) out[k] = 0;
    if (relations) {
        if (!obs_valid_image(&a) || !obs_valid_image(&b)) return OBS_DEVICE_ERROR;
        obs_pair_relations(&a, &b, out + 26);
    }
    return ORC_OK;
}

/* photo_kernel, features.cu:64-90: out[10..17]. */
static void obs_photo(const orc_view* a, const orc_view* b, const double* m, double* out) {
    double p[8][OBS_N];
    for (int t = 0; t < OBS_N; ++t) {
        double n = 0, sum_a = 0, sum_b = 0, sq_a = 0, sq_b = 0, cross = 0, abs_error = 0, sq_error = 0;
        for (int i = t; i < 4096; i += OBS_N) {
            const double x = (i % 64) * (a->w - 1.) / 63., y = (i / 64) * (a->h - 1.) / 63.;
            const double z = m[6] * x + m[7] * y + m[8];
            if (fabs(z) < 1e-12) continue;
            const double u = (m[0] * x + m[1] * y + m[2]) / z, v = (m[3] * x + m[4] * y + m[5]) / z;
            if (u < 0 || v < 0 || u > b->w - 1 || v > b->h - 1 || !finite_d(u) || !finite_d(v)) continue;
            const double av = obs_sample(a, fmax(0., x), fmax(0., y)), bv = obs_sample(b, u, v), d = av - bv;
            n += 1;
            sum_a += av;
            sum_b += bv;
            sq_a += av * av;
            sq_b += bv * bv;
            cross += av * bv;
            abs_error += fabs(d);
            sq_error += d * d;
        }
        p[0][t] = n;
        p[1][t] = sum_a;
        p[2][t] = sum_b;
        p[3][t] = sq_a;
        p[4][t] = sq_b;
        p[5][t] = cross;
        p[6][t] = abs_error;
        p[7][t] = sq_error;
    }
    double s[8];
    for (int k = 0; k < 8; ++k) s[k] = obs_tree(p[k], OBS_N);
    const double den = fmax(1., s[0]), ma = s[1] / den, mb = s[2] / den;
    const double va = fmax(0., s[3] / den - ma * ma), vb = fmax(0., s[4] / den - mb * mb);
    const double sd = sqrt(va) + sqrt(vb);
    out[10] = s[0] / 4096.;
    out[11] = s[6] / den;
    out[12] = sqrt(s[7] / den);
    out[13] = fmin(1., fmax(-1., (s[5] / den - ma * mb) / fmax(1e-12, sqrt(va * vb))));
    out[14] = fabs(ma - mb);
    out[15] = sqrt(va);
    out[16] = sqrt(vb);
    out[17] = fmin(1e6, s[6] / den / fmax(1e-6, sd));
}

/* inlier_kernel, features.cu:91-114: out[18..23]; bins from float
 * coordinates. */
static void obs_inliers(const float* ax, const float* ay, int cap_a, const float* bx, const float* by, int cap_b,
                        const int32_t* ia_, const int32_t* ib_, const uint8_t* mask, int count, int wa, int ha,
                        int wb, int hb, double* out) {
    int hist[32] = {0};
    for (int i = 0; i < count; ++i) {
        if (!mask[i]) continue;
        const int ia = ia_[i], ib = ib_[i];
        if (ia < 0 || ib < 0 || ia >= cap_a || ib >= cap_b) continue;
        int xa = (int)(4 * ax[ia] / wa), ya = (int)(4 * ay[ia] / ha);
        int xb = (int)(4 * bx[ib] / wb), yb = (int)(4 * by[ib] / hb);
        xa = xa < 0 ? 0 : xa > 3 ? 3 : xa;
        ya = ya < 0 ? 0 : ya > 3 ? 3 : ya;
        xb = xb < 0 ? 0 : xb > 3 ? 3 : xb;
        yb = yb < 0 ? 0 : yb > 3 ? 3 : yb;
        ++hist[ya * 4 + xa];
        ++hist[16 + yb * 4 + xb];
    }
    for (int image = 0; image < 2; ++image) {
        int total = 0, occupied = 0, largest = 0;
        for (int j = 0; j < 16; ++j) {
            const int c = hist[image * 16 + j];
            total += c;
            occupied += c > 0;
            largest = c > largest ? c : largest;
        }
        double entropy = 0;
        if (total)
            for (int j = 0; j < 16; ++j)
                if (hist[image * 16 + j]) {
                    const double p = (double)hist[image * 16 + j] / total;
                    entropy -= p * log(p) / log(16.);
                }
        out[18 + image] = occupied / 16.;
        out[20 + image] = total ? (double)largest / total : 0.;
        out[22 + image] = entropy;
    }
}

/* statistics_kernel, relational_features.cu:103-154: out[0..7] of the
 * alignment relations. */
static void obs_statistics(const float* ax, const float* ay, int cap_a, const float* bx, const float* by, int cap_b,
                           const int32_t* ia_, const int32_t* ib_, const float* distance, const uint8_t* mask,
                           int num_matches, const double* H, const orc_scale_info* sa, const orc_scale_info* sb,
                           const orc_top2* forward, unsigned groups, double* out) {
    static double regional[16][OBS_N];
    int residual_hist[257] = {0}, distance_hist[257] = {0}, counts[16] = {0};
    double dsum[OBS_N], inl[OBS_N];
    memset(regional, 0, sizeof regional);
    for (int t = 0; t < OBS_N; ++t) {
        int inliers = 0;
        double distance_sum = 0;
        for (int i = t; i < num_matches; i += OBS_N) {
            const int ia = ia_[i], ib = ib_[i];
            if (ia < 0 || ib < 0 || ia >= cap_a || ib >= cap_b) continue;
            if (groups & 1) {
                const double d = fmin(256., fmax(0., (double)distance[i]));
                distance_sum += d;
                ++distance_hist[(int)d];
            }
            if (!(groups & 2) || !mask[i]) continue;
            const double x = (bx[ib] + .5) / sb->sx - .5, y = (by[ib] + .5) / sb->sy - .5;
            const double z = H[6] * x + H[7] * y + H[8];
            if (!finite_d(z) || fabs(z) < 1e-12) continue;
            const double u = (H[0] * x + H[1] * y + H[2]) / z, v = (H[3] * x + H[4] * y + H[5]) / z;
            const double e = hypot(u - ((ax[ia] + .5) / sa->sx - .5), v - ((ay[ia] + .5) / sa->sy - .5));
            if (!finite_d(e)) continue;
            ++residual_hist[(int)(fmin(32., e) * 8)];
            ++inliers;
            int tx = (int)(ax[ia] * 4 / sa->work_w), ty = (int)(ay[ia] * 4 / sa->work_h);
            tx = tx < 0 ? 0 : tx > 3 ? 3 : tx;
            ty = ty < 0 ? 0 : ty > 3 ? 3 : ty;
            ++counts[tx + 4 * ty];
            regional[tx + 4 * ty][t] += fmin(1e6, e);
        }
        dsum[t] = distance_sum;
        inl[t] = inliers;
    }
    const double distance_sum = obs_tree(dsum, OBS_N);
    const int inliers = (int)obs_tree(inl, OBS_N);
    double maximum = 0, sum = 0, sq = 0;
    int regions = 0;
    if (groups & 2)
        for (int j = 0; j < 16; ++j) {
            const double total = obs_tree(regional[j], OBS_N);
            if (counts[j]) {
                const double v = total / counts[j];
                maximum = fmax(maximum, v);
                sum += v;
                sq += v * v;
                ++regions;
            }
        }
    double rs[OBS_N], rc[OBS_N], am[OBS_N];
    for (int t = 0; t < OBS_N; ++t) {
        double ratio_sum = 0;
        int ratios = 0, ambiguous = 0;
        if (forward && (groups & 1))
            for (int j = t; j < cap_a; j += OBS_N) {
                const orc_top2 p = forward[j];
                if (p.first_index < 0 || p.second_index < 0 || !finite_d(p.second_distance)) continue;
                const double ratio = p.second_distance > 0 ? p.first_distance / p.second_distance : 1.;
                ratio_sum += ratio;
                ++ratios;
                ambiguous += ratio >= .8;
            }
        rs[t] = ratio_sum;
        rc[t] = ratios;
        am[t] = ambiguous;
    }
    const double ratio_sum = obs_tree(rs, OBS_N);
    const int ratios = (int)obs_tree(rc, OBS_N), ambiguous = (int)obs_tree(am, OBS_N);
    int seen = 0, p50 = 256, p90 = 256, dp90 = 256;
    for (int j = 0; j < 257; ++j) {
        seen += residual_hist[j];
        if (seen >= ceil(.5 * inliers) && p50 == 256) p50 = j;
        if (seen >= ceil(.9 * inliers)) {
            p90 = j;
            break;
        }
    }
    seen = 0;
    for (int j = 0; j < 257; ++j) {
        seen += distance_hist[j];
        if (seen >= ceil(.9 * num_matches)) {
            dp90 = j;
            break;
        }
    }
    out[0] = inliers ? p50 / 8. : 32;
    out[1] = inliers ? p90 / 8. : 32;
    out[2] = maximum;
    out[3] = regions ? sqrt(fmax(0., sq / regions - (sum / regions) * (sum / regions))) : 0;
    out[4] = distance_sum / fmax(1., (double)num_matches) / 256.;
    out[5] = dp90 / 256.;
    out[6] = ratios ? ratio_sum / ratios : 1;
    out[7] = ratios ? (double)ambiguous / ratios : 1;
}

/* relational_features.cu:60-66 */
static int obs_warp(const orc_view* b, const double* m, double x, double y, double* value) {
    const double z = m[6] * x + m[7] * y + m[8];
    if (!finite_d(z) || fabs(z) < 1e-12) return 0;
    const double u = (m[0] * x + m[1] * y + m[2]) / z, v = (m[3] * x + m[4] * y + m[5]) / z;
    if (!finite_d(u) || !finite_d(v) || u < 0 || v < 0 || u > b->w - 1 || v > b->h - 1) return 0;
    *value = obs_sample(b, u, v);
    return 1;
}

/* patch_kernel and patch_finish, relational_features.cu:67-102: out[8..13]
 * of the alignment relations. */
static void obs_patches(const orc_view* a, const orc_view* b, const double* m, double* out) {
    double ws[16][5];
    for (int patch = 0; patch < 16; ++patch) {
        const double cx = (patch % 4 + .5) * a->w / 4., cy = (patch / 4 + .5) * a->h / 4.;
        double p[11][OBS_PATCH_N];
        for (int t = 0; t < OBS_PATCH_N; ++t) {
            double n = 0, sa = 0, sb = 0, aa = 0, bb = 0, ab = 0, mae = 0, edge = 0, dot = 0, ga2 = 0, gb2 = 0;
            for (int i = t; i < 49; i += OBS_PATCH_N) {
                const int dx = i % 7 - 3, dy = i / 7 - 3;
                const double x = cx + dx, y = cy + dy;
                double bv, bxv, byv;
                if (x < 0 || y < 0 || x + 1 > a->w - 1 || y + 1 > a->h - 1 || !obs_warp(b, m, x, y, &bv) ||
                    !obs_warp(b, m, x + 1, y, &bxv) || !obs_warp(b, m, x, y + 1, &byv))
                    continue;
                const double av = obs_sample(a, x, y), gx = obs_sample(a, x + 1, y) - av,
                             gy = obs_sample(a, x, y + 1) - av;
                bxv -= bv;
                byv -= bv;
                ++n;
                sa += av;
                sb += bv;
                aa += av * av;
                bb += bv * bv;
                ab += av * bv;
                mae += fabs(av - bv);
                edge += fabs(gx - bxv) + fabs(gy - byv);
                dot += gx * bxv + gy * byv;
                ga2 += gx * gx + gy * gy;
                gb2 += bxv * bxv + byv * byv;
            }
            p[0][t] = n;
            p[1][t] = sa;
            p[2][t] = sb;
            p[3][t] = aa;
            p[4][t] = bb;
            p[5][t] = ab;
            p[6][t] = mae;
            p[7][t] = edge;
            p[8][t] = dot;
            p[9][t] = ga2;
            p[10][t] = gb2;
        }
        double s[11];
        for (int k = 0; k < 11; ++k) s[k] = obs_tree(p[k], OBS_PATCH_N);
        const double den = fmax(s[0], 1.), ma = s[1] / den, mb = s[2] / den;
        const double va = fmax(0., s[3] / den - ma * ma), vb = fmax(0., s[4] / den - mb * mb);
        ws[patch][0] = ((2 * ma * mb + .0001) * (2 * (s[5] / den - ma * mb) + .0009)) /
                       ((ma * ma + mb * mb + .0001) * (va + vb + .0009));
        ws[patch][1] = s[7] / (2 * den);
        ws[patch][2] = s[8] / fmax(1e-12, sqrt(s[9] * s[10]));
        ws[patch][3] = fmin(1e6, s[6] / den / fmax(1e-6, sqrt(va) + sqrt(vb)));
        ws[patch][4] = s[0] == 49;
    }
    int n = 0;
    double sum = 0, low = 1, eg = 0, co = 0, errors[16];
    for (int j = 0; j < 16; ++j)
        if (ws[j][4]) {
            errors[n++] = ws[j][3];
            sum += ws[j][0];
            low = fmin(low, ws[j][0]);
            eg += ws[j][1];
            co += ws[j][2];
        }
    for (int j = 1; j < n; ++j) {
        const double v = errors[j];
        int k = j;
        while (k && errors[k - 1] > v) {
            errors[k] = errors[k - 1];
            --k;
        }
        errors[k] = v;
    }
    out[8] = n ? sum / n : -1;
    out[9] = n ? low : -1;
    out[10] = n ? eg / n : 1;
    out[11] = n ? co / n : 0;
    out[12] = n / 16.;
    out[13] = n ? errors[(int)ceil(.9 * n) - 1] : 1e6;
}

/* extract_gate, features.cu:136-170 with alignment_relations,
 * relational_features.cu:168-189: 38 doubles. The homography result is
 * passed field by field (status must be OK). Returns 0, 1 (INVALID_ARGUMENT:
 * nothing computed) or 3 (inputs the alignment relations reject). */
int orc_gate_features(const uint8_t* da, int wa, int ha, int pa, int fa, const uint8_t* db, int wb, int hb, int pb,
                      int fb, const float* ax, const float* ay, int cap_a, const float* bx, const float* by,
                      int cap_b, const int32_t* idx_a, const int32_t* idx_b, const float* distance, int capacity,
                      const uint8_t* mask, int num_matches, int num_inliers, double inlier_ratio,
                      double inlier_coverage, double reproj_residual, const double* H, const orc_scale_info* sa,
                      const orc_scale_info* sb, const orc_top2* forward, unsigned relation_mask, double* out) {
    const orc_view a = {da, wa, ha, pa, fa}, b = {db, wb, hb, pb, fb};
    double inv[9];
    if (!da || !db || !out || !mask || num_matches < 4 || num_matches > capacity || !obs_inverse(H, inv))
        return ORC_INVALID_ARGUMENT;
    double x[4], y[4], zmin = 1e300, zmax = 0.;
    out[0] = num_matches / 4000.;
    out[1] = num_inliers / 4000.;
    out[2] = inlier_ratio;
    out[3] = inlier_coverage;
    out[4] = reproj_residual;
    out[5] = reproj_residual / (sa->sx < sa->sy ? sa->sy : sa->sx);
    for (int i = 0; i < 4; ++i) {
        const double u = (i == 1 || i == 2) ? wb - 1. : 0., v = i >= 2 ? hb - 1. : 0.;
        const double z = H[6] * u + H[7] * v + H[8];
        if (!finite_d(z) || fabs(z) < 1e-12) return ORC_INVALID_ARGUMENT;
        x[i] = (H[0] * u + H[1] * v + H[2]) / z;
        y[i] = (H[3] * u + H[4] * v + H[5]) / z;
        if (!finite_d(x[i]) || !finite_d(y[i])) return ORC_INVALID_ARGUMENT;
        zmin = fmin(zmin, fabs(z));
        zmax = fmax(zmax, fabs(z));
    }
    double area = 0.;
    for (int i = 0; i < 4; ++i) area += x[i] * y[(i + 1) % 4] - y[i] * x[(i + 1) % 4];
    double xl = x[0], xh = x[0], yl = y[0], yh = y[0];
    for (int i = 1; i < 4; ++i) {
        xl = x[i] < xl ? x[i] : xl;
        xh = xh < x[i] ? x[i] : xh;
        yl = y[i] < yl ? y[i] : yl;
        yh = yh < y[i] ? y[i] : yh;
    }
    out[6] = fmin(1e6, fabs(area) / 2. / ((double)wb * hb));
    out[7] = fmin(1e6, (xh - xl) / wa);
    out[8] = fmin(1e6, (yh - yl) / ha);
    out[9] = fmin(1e6, zmax / zmin);
    obs_photo(&a, &b, inv, out);
    obs_inliers(ax, ay, cap_a, bx, by, cap_b, idx_a, idx_b, mask, num_matches, sa->work_w, sa->work_h, sb->work_w,
                sb->work_h, out);
    for (int k = 24; k < 38; ++k) out[k] = 0;
    if (relation_mask) {
        if (!obs_valid_image(&a) || !obs_valid_image(&b) || !idx_a || !idx_b || !distance || !ax || !ay || !bx ||
            !by || cap_a < 1 || cap_a > 4000 || cap_b < 1 || cap_b > 4000 || !finite_d(sa->sx) ||
            !finite_d(sa->sy) || !finite_d(sb->sx) || !finite_d(sb->sy) || sa->sx <= 0 || sa->sy <= 0 ||
            sb->sx <= 0 || sb->sy <= 0 || sa->work_w <= 0 || sa->work_h <= 0 || relation_mask > 7)
            return OBS_DEVICE_ERROR;
        if (relation_mask & 3)
            obs_statistics(ax, ay, cap_a, bx, by, cap_b, idx_a, idx_b, distance, mask, num_matches, H, sa, sb,
                           forward, relation_mask, out + 24);
        if (relation_mask & 4) obs_patches(&a, &b, inv, out + 24);
    }
    return ORC_OK;
}
