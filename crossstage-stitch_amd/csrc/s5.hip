// S5: MSAC homography B -> A for any number of pairs per launch.
// Semantics: src/s5/s5_math.hpp:19-184 and src/s5/s5.cu:42-267 -
// SplitMix64 counter sampling, Hartley-normalised 4-point DLT solved by
// Gauss-Jordan in FP64, MSAC loss sum min(r^2, t^2), best by (valid, loss,
// lower hypothesis), refinement over inliers, telemetry, H = D_A^-1 H_s D_B.
//
// Every FP64 expression keeps the reference's operation order with no fused
// multiply-add, and each sum keeps its order: the hypothesis loss is 128
// lane-strided partial sums reduced by a halving tree (one wave64 emulates
// the 128 lanes), the refinement and telemetry sums are serial in match
// order. All R hypotheses of a pair run in one launch (one lane each for the
// solve, one wave each for the loss) - batching hypotheses cannot change the
// winner because (valid, loss, hypothesis) is a strict total order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>

#include "stitch/jobs.hpp"

namespace stitch {
namespace {

using jobs::S5Job;
using jobs::S5Model;
using jobs::S5State;

constexpr double kMax = 1.7976931348623157e308;

struct Pt {
    double ax, ay, bx, by;
};

__device__ __forceinline__ bool finite(double x) { return x == x && x <= kMax && x >= -kMax; }

__device__ void mul3(const double* a, const double* b, double* out) {
    double t[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[j + 3] + a[3 * i + 2] * b[j + 6];
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = t[i];
}

__device__ bool normalize(double* h) {
    double scale = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (!finite(h[i])) return false;
        if (fabs(h[i]) > scale) scale = fabs(h[i]);
    }
    if (!(scale > 0.0)) return false;
#pragma unroll
    for (int i = 0; i < 9; ++i) h[i] /= scale;
    const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                       h[2] * (h[3] * h[7] - h[4] * h[6]);
    return finite(det) && fabs(det) > 1e-15;
}

__device__ __forceinline__ double residual_squared(const double* h, const Pt& p) {
    const double z = h[6] * p.bx + h[7] * p.by + h[8];
    if (!finite(z) || fabs(z) < 1e-15) return kMax;
    const double dx = (h[0] * p.bx + h[1] * p.by + h[2]) / z - p.ax;
    const double dy = (h[3] * p.bx + h[4] * p.by + h[5]) / z - p.ay;
    const double r = dx * dx + dy * dy;
    return finite(r) ? r : kMax;
}

__device__ __forceinline__ std::uint64_t random_word(std::uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

// k-th draw uniform over the n-k remaining indices, mapped back by skipping
// the already drawn ones in sorted order (s5_math.hpp:62-79).
__device__ void sample_four(int n, std::uint64_t seed, int hypothesis, int* s) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int v = static_cast<int>(random_word(seed ^ (static_cast<std::uint64_t>(hypothesis) * 4 + k)) %
                                 static_cast<std::uint64_t>(n - k));
        for (int j = 0; j < k; ++j) {
            int smallest = n;
            for (int l = 0; l < k; ++l) {
                int rank = 0;
                for (int m = 0; m < k; ++m) rank += s[m] < s[l];
                if (rank == j) smallest = s[l];
            }
            if (v >= smallest) ++v;
        }
        s[k] = v;
    }
}

__device__ bool nondegenerate_four(const Pt* p) {
#pragma unroll
    for (int image = 0; image < 2; ++image) {
        double scale = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < i; ++j) {
                const double dx = image ? p[i].bx - p[j].bx : p[i].ax - p[j].ax;
                const double dy = image ? p[i].by - p[j].by : p[i].ay - p[j].ay;
                const double d = dx * dx + dy * dy;
                if (d > scale) scale = d;
            }
        if (!(scale > 0.0) || !finite(scale)) return false;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = i + 1; j < 4; ++j)
#pragma unroll
                for (int k = j + 1; k < 4; ++k) {
                    const double x1 = image ? p[j].bx - p[i].bx : p[j].ax - p[i].ax;
                    const double y1 = image ? p[j].by - p[i].by : p[j].ay - p[i].ay;
                    const double x2 = image ? p[k].bx - p[i].bx : p[k].ax - p[i].ax;
                    const double y2 = image ? p[k].by - p[i].by : p[k].ay - p[i].ay;
                    if (fabs(x1 * y2 - y1 * x2) <= 1e-8 * scale) return false;
                }
    }
    return true;
}

struct Norm {
    double ax, ay, bx, by, sa, sb;
};

// Entry (r, c) of rows[which] for a normalised point: rows[0] =
// {x, y, 1, 0, 0, 0, -u x, -u y, u}, rows[1] = {0, 0, 0, x, y, 1, -v x, -v y, v}.
__device__ __forceinline__ double row_entry(int which, int k, double x, double y, double u, double v) {
    const double w = which ? v : u;
    const int o = which ? 3 : 0;
    if (k == o) return x;
    if (k == o + 1) return y;
    if (k == o + 2) return 1.0;
    if (k == 6) return -w * x;
    if (k == 7) return -w * y;
    if (k == 8) return w;
    return 0.0;
}

// Gauss-Jordan with partial pivoting on the 8x9 normal equations, then
// denormalisation and normalisation (s5_math.hpp:143-172). Static indices
// only, so the matrix stays in registers.
__device__ bool solve_normal(double (&a)[8][9], const Norm& t, double* out) {
    double max_entry = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (fabs(a[i][i]) > max_entry) max_entry = fabs(a[i][i]);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int pivot = c;
        double pv = a[c][c];
#pragma unroll
        for (int r = c + 1; r < 8; ++r)
            if (fabs(a[r][c]) > fabs(pv)) {
                pivot = r;
                pv = a[r][c];
            }
        if (!finite(pv) || fabs(pv) <= 1e-12 * max_entry) return false;
#pragma unroll
        for (int r = c + 1; r < 8; ++r)
            if (r == pivot)
#pragma unroll
                for (int j = c; j < 9; ++j) {
                    const double v = a[c][j];
                    a[c][j] = a[r][j];
                    a[r][j] = v;
                }
        const double divisor = a[c][c];
#pragma unroll
        for (int j = c; j < 9; ++j) a[c][j] /= divisor;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (r == c) continue;
            const double factor = a[r][c];
#pragma unroll
            for (int j = c; j < 9; ++j) a[r][j] -= factor * a[c][j];
        }
    }
    double h[9];
#pragma unroll
    for (int i = 0; i < 8; ++i) h[i] = a[i][8];
    h[8] = 1.0;
    const double inv_a[9] = {1.0 / t.sa, 0, t.ax, 0, 1.0 / t.sa, t.ay, 0, 0, 1};
    const double b[9] = {t.sb, 0, -t.sb * t.bx, 0, t.sb, -t.sb * t.by, 0, 0, 1};
    double tmp[9];
    mul3(inv_a, h, tmp);
    mul3(tmp, b, out);
    return normalize(out);
}

// fit() for exactly four points held in registers (s5_math.hpp:110-172).
__device__ bool fit_four(const Pt* p, double* out) {
    Norm t{0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        t.ax += p[i].ax; t.ay += p[i].ay; t.bx += p[i].bx; t.by += p[i].by;
    }
    t.ax /= 4; t.ay /= 4; t.bx /= 4; t.by /= 4;
    double va = 0.0, vb = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        va += (p[i].ax - t.ax) * (p[i].ax - t.ax) + (p[i].ay - t.ay) * (p[i].ay - t.ay);
        vb += (p[i].bx - t.bx) * (p[i].bx - t.bx) + (p[i].by - t.by) * (p[i].by - t.by);
    }
    if (!(va > 0.0) || !(vb > 0.0) || !finite(va) || !finite(vb)) return false;
    t.sa = sqrt(2.0 * 4 / va);
    t.sb = sqrt(2.0 * 4 / vb);
    double a[8][9];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 9; ++c) a[r][c] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double x = (p[i].bx - t.bx) * t.sb, y = (p[i].by - t.by) * t.sb;
        const double u = (p[i].ax - t.ax) * t.sa, v = (p[i].ay - t.ay) * t.sa;
        const double r0[9] = {x, y, 1, 0, 0, 0, -u * x, -u * y, u};
        const double r1[9] = {0, 0, 0, x, y, 1, -v * x, -v * y, v};
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 9; ++c) a[r][c] += r0[r] * r0[c] + r1[r] * r1[c];
    }
    return solve_normal(a, t, out);
}

__device__ bool valid_domain(const double* h, int width, int height) {
    const double z[4] = {h[8], h[6] * (width - 1) + h[8], h[7] * (height - 1) + h[8],
                         h[6] * (width - 1) + h[7] * (height - 1) + h[8]};
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (!finite(z[i]) || fabs(z[i]) < 1e-15 || (z[i] > 0.0) != (z[0] > 0.0)) return false;
    return true;
}

__device__ __forceinline__ Pt point(const S5Job& j, int i) {
    const int ia = j.idx_a[i], ib = j.idx_b[i];
    return {j.ax[ia], j.ay[ia], j.bx[ib], j.by[ib]};
}

__device__ bool same_scale(const ScaleInfo& a, const ScaleInfo& b) {
    if (a.orig_w != b.orig_w || a.orig_h != b.orig_h || a.work_w != b.work_w || a.work_h != b.work_h ||
        a.sx != b.sx || a.sy != b.sy)
        return false;
    for (int i = 0; i < 9; ++i)
        if (a.D[i] != b.D[i]) return false;
    return true;
}

// prepare (s5.cu:42-83): state reset or reuse check, then validation.
__device__ __forceinline__ void s5_prepare_job(const S5Job& j) {
    S5State* s = j.state;
    __shared__ int proceed, na_s, nb_s, nm_s;
    if (threadIdx.x == 0) {
        const int na = *j.count_a, nb = *j.count_b, nm = *j.count_m;
        const s5::Config& c = j.cfg;
        if (c.previous_hypotheses == 0) {
            s->best = S5Model{};
            s->ax = j.ax; s->ay = j.ay; s->bx = j.bx; s->by = j.by;
            s->idx_a = j.idx_a; s->idx_b = j.idx_b;
            s->count_a = j.count_a; s->count_b = j.count_b; s->count_matches = j.count_m;
            s->na = na; s->nb = nb; s->nm = nm;
            s->capacity_a = j.cap_a; s->capacity_b = j.cap_b; s->capacity_matches = j.cap_m;
            s->evaluated = 0; s->batch_size = c.batch_size; s->threshold = c.threshold;
            s->seed = c.seed; s->generation = c.generation;
            s->scale_a = j.sa; s->scale_b = j.sb;
            s->status = Status::OK;
        } else if (s->evaluated != c.previous_hypotheses || s->generation != c.generation || s->seed != c.seed ||
                   s->threshold != c.threshold || s->batch_size != c.batch_size || s->ax != j.ax ||
                   s->ay != j.ay || s->bx != j.bx || s->by != j.by || s->idx_a != j.idx_a ||
                   s->idx_b != j.idx_b || s->count_a != j.count_a || s->count_b != j.count_b ||
                   s->count_matches != j.count_m || s->capacity_a != j.cap_a || s->capacity_b != j.cap_b ||
                   s->capacity_matches != j.cap_m || s->na != na || s->nb != nb || s->nm != nm ||
                   !same_scale(s->scale_a, j.sa) || !same_scale(s->scale_b, j.sb)) {
            s->status = Status::INVALID_ARGUMENT;
        }
        int go = s->status == Status::OK;
        if (go && (na < 0 || na > j.cap_a || nb < 0 || nb > j.cap_b || nm < 0 || nm > j.cap_m)) {
            s->status = Status::INVALID_ARGUMENT;
            go = 0;
        }
        proceed = go;
        na_s = na; nb_s = nb; nm_s = nm;
    }
    __syncthreads();
    if (!proceed) return;
    int bad = 0;
    for (int i = threadIdx.x; i < nm_s; i += blockDim.x) {
        const int ia = j.idx_a[i], ib = j.idx_b[i];
        if (ia < 0 || ia >= na_s || ib < 0 || ib >= nb_s) { bad = 1; continue; }
        const Pt q = point(j, i);
        if (!finite(q.ax) || !finite(q.ay) || !finite(q.bx) || !finite(q.by)) bad = 1;
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        if (bad) s->status = Status::INVALID_ARGUMENT;
        else if (nm_s < 4) s->status = Status::NOT_ENOUGH_POINTS;
    }
}

__global__ void __launch_bounds__(256) s5_prepare(dev::JobSpan<S5Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s5_prepare_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

// One lane per hypothesis: sample, degeneracy test, 4-point solve, domain.
__global__ void __launch_bounds__(256) s5_generate(dev::JobSpan<S5Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        ji = dev::next_job(span, t, &S5Job::gen_begin, ji, njobs);
        const S5Job& j = dev::job_at(span, ji);
        const int i = static_cast<int>(t - j.gen_begin) * 256 + threadIdx.x;
        if (i >= j.count) continue;
        S5Model m{};
        m.hypothesis = j.first + i;
        const S5State* s = j.state;
        if (s->status == Status::OK) {
            int idx[4];
            sample_four(s->nm, j.cfg.seed, m.hypothesis, idx);
            Pt f[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) f[k] = point(j, idx[k]);
            m.valid = nondegenerate_four(f) && fit_four(f, m.h) && valid_domain(m.h, s->scale_b.work_w, s->scale_b.work_h);
        }
        j.models[i] = m;
    }
}

// One wave per hypothesis: lane l accumulates the reference's virtual lanes l
// and l + 64 (strided by 128, in order), then the halving tree 64, 32, ..., 1.
__global__ void __launch_bounds__(256) s5_score(dev::JobSpan<S5Job> span, dev::Total work) {
    const long long total = work.get();
    const int lane = dev::lane_id(), wave = threadIdx.x / 64;
    int ji = 0;
    const int njobs = span.jobs();
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        ji = dev::next_job(span, t, &S5Job::score_begin, ji, njobs);
        const S5Job& j = dev::job_at(span, ji);
        const int i = static_cast<int>(t - j.score_begin) * 4 + wave;
        if (i >= j.count) continue;
        S5Model* model = &j.models[i];
        const S5State* s = j.state;
        const double t2 = j.cfg.threshold * j.cfg.threshold;
        double lo = 0.0, hi = 0.0;
        const bool active = s->status == Status::OK && model->valid;
        if (active) {
            double h[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) h[k] = model->h[k];
            const int nm = s->nm;
            for (int q = lane; q < nm; q += 128) {
                const double r = residual_squared(h, point(j, q));
                lo += r < t2 ? r : t2;
            }
            for (int q = lane + 64; q < nm; q += 128) {
                const double r = residual_squared(h, point(j, q));
                hi += r < t2 ? r : t2;
            }
        }
        double v = lo + hi;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(v, off, 64);
            v = v + o;
        }
        if (lane == 0) {
            model->loss = v;
            if (!finite(v)) model->valid = 0;
        }
    }
}

__device__ __forceinline__ bool better(const S5Model& a, const S5Model& b) {
    return a.valid && (!b.valid || a.loss < b.loss || (a.loss == b.loss && a.hypothesis < b.hypothesis));
}

// Best model of this call merged into the persistent raw winner.
__device__ __forceinline__ void s5_select_job(const S5Job& j) {
    __shared__ S5Model best[256];
    S5Model mine{};
    for (int i = threadIdx.x; i < j.count; i += blockDim.x)
        if (better(j.models[i], mine)) mine = j.models[i];
    best[threadIdx.x] = mine;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride && better(best[threadIdx.x + stride], best[threadIdx.x]))
            best[threadIdx.x] = best[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0 && j.state->status == Status::OK && better(best[0], j.state->best)) j.state->best = best[0];
}

__global__ void __launch_bounds__(256) s5_select(dev::JobSpan<S5Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s5_select_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

// ---------------------------------------------------------------- finish
// One workgroup per job. Residuals and per-point terms are computed in
// parallel 256 matches at a time; every sum is then accumulated in match
// order by one thread (or, for the 72 normal-equation entries, one thread per
// entry), which reproduces the reference's serial order exactly.

constexpr int kChunk = 256;

struct FinishShared {
    Pt pts[kChunk];
    double term[kChunk];
    double term2[kChunk];
    unsigned char keep[kChunk];
    double a[8][9];
    double h[9];
    double scalars[8];
    int flags[4];
};

__device__ void load_chunk(const S5Job& j, int base, int nm, FinishShared& sh) {
    const int i = base + threadIdx.x;
    if (i < nm) sh.pts[threadIdx.x] = point(j, i);
}

// serial_loss (s5.cu:139-146), result broadcast to all threads.
__device__ double serial_loss(const S5Job& j, int nm, const double* h, double t2, FinishShared& sh) {
    double acc = 0.0;
    for (int base = 0; base < nm; base += kChunk) {
        load_chunk(j, base, nm, sh);
        const int i = base + threadIdx.x;
        if (i < nm) {
            const double r = residual_squared(h, sh.pts[threadIdx.x]);
            sh.term[threadIdx.x] = r < t2 ? r : t2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = min(kChunk, nm - base);
            for (int k = 0; k < n; ++k) acc += sh.term[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) sh.scalars[7] = acc;
    __syncthreads();
    const double v = sh.scalars[7];
    __syncthreads();
    return v;
}

// fit() over the inliers of `model` (s5_math.hpp:110-172, Inliers keep rule).
__device__ bool fit_inliers(const S5Job& j, int nm, const double* model, double t2, double* out, FinishShared& sh) {
    // Centroids and count.
    double sax = 0, say = 0, sbx = 0, sby = 0;
    int count = 0;
    for (int base = 0; base < nm; base += kChunk) {
        load_chunk(j, base, nm, sh);
        const int i = base + threadIdx.x;
        if (i < nm) sh.keep[threadIdx.x] = residual_squared(model, sh.pts[threadIdx.x]) <= t2;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = min(kChunk, nm - base);
            for (int k = 0; k < n; ++k) {
                if (!sh.keep[k]) continue;
                sax += sh.pts[k].ax; say += sh.pts[k].ay; sbx += sh.pts[k].bx; sby += sh.pts[k].by;
                ++count;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sh.flags[0] = count;
        sh.scalars[0] = sax / count; sh.scalars[1] = say / count;
        sh.scalars[2] = sbx / count; sh.scalars[3] = sby / count;
    }
    __syncthreads();
    count = sh.flags[0];
    if (count < 4) return false;
    Norm t{sh.scalars[0], sh.scalars[1], sh.scalars[2], sh.scalars[3], 0, 0};
    __syncthreads();
    // Spreads.
    double va = 0.0, vb = 0.0;
    for (int base = 0; base < nm; base += kChunk) {
        load_chunk(j, base, nm, sh);
        const int i = base + threadIdx.x;
        if (i < nm) {
            const Pt p = sh.pts[threadIdx.x];
            sh.keep[threadIdx.x] = residual_squared(model, p) <= t2;
            sh.term[threadIdx.x] = (p.ax - t.ax) * (p.ax - t.ax) + (p.ay - t.ay) * (p.ay - t.ay);
            sh.term2[threadIdx.x] = (p.bx - t.bx) * (p.bx - t.bx) + (p.by - t.by) * (p.by - t.by);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = min(kChunk, nm - base);
            for (int k = 0; k < n; ++k)
                if (sh.keep[k]) {
                    va += sh.term[k];
                    vb += sh.term2[k];
                }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sh.flags[1] = (va > 0.0) && (vb > 0.0) && finite(va) && finite(vb);
        sh.scalars[4] = sqrt(2.0 * count / va);
        sh.scalars[5] = sqrt(2.0 * count / vb);
    }
    __syncthreads();
    if (!sh.flags[1]) return false;
    t.sa = sh.scalars[4];
    t.sb = sh.scalars[5];
    // Normal equations: thread e < 72 owns entry (e / 9, e % 9).
    const int e = threadIdx.x, er = e / 9, ec = e % 9;
    double acc = 0.0;
    for (int base = 0; base < nm; base += kChunk) {
        load_chunk(j, base, nm, sh);
        const int i = base + threadIdx.x;
        if (i < nm) {
            const Pt p = sh.pts[threadIdx.x];
            sh.keep[threadIdx.x] = residual_squared(model, p) <= t2;
        }
        __syncthreads();
        if (e < 72) {
            const int n = min(kChunk, nm - base);
            for (int k = 0; k < n; ++k) {
                if (!sh.keep[k]) continue;
                const Pt p = sh.pts[k];
                const double x = (p.bx - t.bx) * t.sb, y = (p.by - t.by) * t.sb;
                const double u = (p.ax - t.ax) * t.sa, v = (p.ay - t.ay) * t.sa;
                acc += row_entry(0, er, x, y, u, v) * row_entry(0, ec, x, y, u, v) +
                       row_entry(1, er, x, y, u, v) * row_entry(1, ec, x, y, u, v);
            }
        }
        __syncthreads();
    }
    if (e < 72) sh.a[er][ec] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a[8][9];
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 9; ++c) a[r][c] = sh.a[r][c];
        double h[9];
        sh.flags[2] = solve_normal(a, t, h);
        for (int k = 0; k < 9; ++k) sh.h[k] = h[k];
    }
    __syncthreads();
    const bool ok = sh.flags[2];
    for (int k = 0; k < 9; ++k) out[k] = sh.h[k];
    __syncthreads();
    return ok;
}

__device__ __forceinline__ void s5_finish_job(const S5Job& j) {
    S5State* s = j.state;
    __shared__ FinishShared sh;
    __shared__ HomographyResult result;
    const double invalid = __builtin_nan("");
    if (threadIdx.x == 0) {
        result = HomographyResult{};
        for (int i = 0; i < 9; ++i) result.H[i] = result.H_s[i] = invalid;
        result.reproj_residual = invalid;
        result.num_matches = s->nm >= 0 && s->nm <= j.cap_m ? s->nm : 0;
        result.status = s->status;
        if (s->status == Status::OK && !s->best.valid) result.status = Status::DEGENERATE;
    }
    for (int i = threadIdx.x; i < j.cap_m; i += blockDim.x) j.mask[i] = 0;
    __syncthreads();
    if (result.status == Status::OK) {
        const int nm = s->nm;
        const double t2 = j.cfg.threshold * j.cfg.threshold;
        double h[9];
        for (int k = 0; k < 9; ++k) h[k] = s->best.h[k];
        double loss = serial_loss(j, nm, h, t2, sh);
        for (int step = 0; step < j.cfg.refinement_steps; ++step) {
            double cand[9];
            const bool fitted = fit_inliers(j, nm, h, t2, cand, sh);
            if (!(fitted && valid_domain(cand, s->scale_b.work_w, s->scale_b.work_h))) break;
            const double closs = serial_loss(j, nm, cand, t2, sh);
            if (!finite(closs) || !(closs < loss)) break;
            for (int k = 0; k < 9; ++k) h[k] = cand[k];
            loss = closs;
        }
        // Telemetry, mask and residual sum in match order.
        double sum = 0.0, min_x = kMax, min_y = kMax, max_x = -kMax, max_y = -kMax;
        int inliers = 0;
        for (int base = 0; base < nm; base += kChunk) {
            load_chunk(j, base, nm, sh);
            const int i = base + threadIdx.x;
            if (i < nm) {
                const double r = residual_squared(h, sh.pts[threadIdx.x]);
                sh.term[threadIdx.x] = r;
                const bool in = !(r > t2);
                sh.keep[threadIdx.x] = in;
                if (in) j.mask[i] = 1;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                const int n = min(kChunk, nm - base);
                for (int k = 0; k < n; ++k) {
                    if (!sh.keep[k]) continue;
                    ++inliers;
                    sum += sh.term[k];
                    min_x = fmin(min_x, sh.pts[k].ax); max_x = fmax(max_x, sh.pts[k].ax);
                    min_y = fmin(min_y, sh.pts[k].ay); max_y = fmax(max_y, sh.pts[k].ay);
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            result.num_inliers = inliers;
            if (inliers < 4) {
                result.status = Status::DEGENERATE;
                result.num_inliers = 0;
                sh.flags[3] = 1;
            } else {
                result.inlier_ratio = static_cast<double>(inliers) / nm;
                result.reproj_residual = sqrt(sum / inliers);
                const double width = fmax(1.0, static_cast<double>(s->scale_a.work_w - 1));
                const double height = fmax(1.0, static_cast<double>(s->scale_a.work_h - 1));
                result.inlier_coverage = fmin(1.0, (max_x - min_x) / width) * fmin(1.0, (max_y - min_y) / height);
                for (int k = 0; k < 9; ++k) result.H_s[k] = h[k];
                double tmp[9];
                mul3(j.inverse_a, h, tmp);
                mul3(tmp, j.d_b, result.H);
                sh.flags[3] = 0;
                if (!normalize(result.H) || !valid_domain(result.H, s->scale_b.orig_w, s->scale_b.orig_h)) {
                    result.status = Status::DEGENERATE;
                    result.num_inliers = 0;
                    result.inlier_ratio = result.inlier_coverage = 0.0;
                    result.reproj_residual = invalid;
                    for (int k = 0; k < 9; ++k) result.H[k] = result.H_s[k] = invalid;
                    sh.flags[3] = 1;
                }
            }
        }
        __syncthreads();
        if (sh.flags[3])
            for (int i = threadIdx.x; i < j.cap_m; i += blockDim.x) j.mask[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s->evaluated = j.cfg.hypotheses;
        s->result = result;
    }
}

__global__ void __launch_bounds__(256) s5_finish(dev::JobSpan<S5Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s5_finish_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

bool valid_keypoints(const Keypoints& p) { return p.capacity >= 0 && p.d_count && (p.capacity == 0 || (p.x && p.y)); }

bool valid_scale(const ScaleInfo& s) {
    if (s.orig_w <= 0 || s.orig_h <= 0 || s.work_w <= 0 || s.work_h <= 0 || !(s.sx > 0.0) || !(s.sy > 0.0) ||
        !std::isfinite(s.sx) || !std::isfinite(s.sy))
        return false;
    for (double x : s.D)
        if (!std::isfinite(x)) return false;
    return true;
}

}  // namespace

namespace jobs {

Status make_s5_job(const Keypoints& a, const Keypoints& b, const Matches& m, const ScaleInfo& sa, const ScaleInfo& sb,
                   const s5::Config& c, std::uint8_t* mask, void* ws, std::size_t bytes, S5Job* job) {
    std::size_t required = 0;
    const Status q = s5::s5_workspace_bytes(c, &required);
    if (q != Status::OK) return q;
    if (!valid_keypoints(a) || !valid_keypoints(b) || m.capacity < 0 || !m.d_count ||
        (m.capacity > 0 && (!m.idx_a || !m.idx_b || !mask)) || !valid_scale(sa) || !valid_scale(sb))
        return Status::INVALID_ARGUMENT;
    if (bytes < required) return Status::INSUFFICIENT_WORKSPACE;
    if (!ws || reinterpret_cast<std::uintptr_t>(ws) % 256 != 0) return Status::INVALID_ARGUMENT;
    S5Job j{};
    if (!mat3_inverse(sa.D, j.inverse_a)) return Status::INVALID_ARGUMENT;
    double inverse_b[9];
    if (!mat3_inverse(sb.D, inverse_b)) return Status::INVALID_ARGUMENT;
    for (int i = 0; i < 9; ++i) {
        if (!std::isfinite(j.inverse_a[i]) || !std::isfinite(inverse_b[i])) return Status::INVALID_ARGUMENT;
        j.d_b[i] = sb.D[i];
    }
    j.ax = a.x; j.ay = a.y; j.bx = b.x; j.by = b.y;
    j.count_a = a.d_count; j.count_b = b.d_count;
    j.cap_a = a.capacity; j.cap_b = b.capacity;
    j.idx_a = m.idx_a; j.idx_b = m.idx_b;
    j.count_m = m.d_count; j.cap_m = m.capacity;
    j.sa = sa; j.sb = sb;
    j.cfg = c;
    j.mask = mask;
    j.state = static_cast<S5State*>(ws);
    j.models = reinterpret_cast<S5Model*>(static_cast<std::uint8_t*>(ws) + kS5ModelOffset);
    j.result_dev = &j.state->result;
    j.first = c.previous_hypotheses;
    j.count = c.hypotheses - c.previous_hypotheses;
    *job = j;
    return Status::OK;
}

S5Totals s5_prefix(S5Job* js, int n) {
    S5Totals t{0, 0};
    for (int i = 0; i < n; ++i) {
        js[i].gen_begin = t.gen;
        js[i].score_begin = t.score;
        t.gen += (js[i].count + 255) / 256;
        t.score += (js[i].count + 3) / 4;
    }
    return t;
}

hipError_t launch_s5(const S5Job* table, int count, const S5Totals& t, const S5Job* single, hipStream_t stream,
                     const dev::Replay<S5Totals>* replay) {
    dev::JobSpan<S5Job> span{table, count, single ? *single : S5Job{}, replay ? replay->count : nullptr};
    const auto total = [&](long long S5Totals::*m) {
        return dev::Total(t.*m, replay ? &(replay->totals->*m) : nullptr);
    };
    s5_prepare<<<count, 256, 0, stream>>>(span);
    if (t.gen > 0) s5_generate<<<persistent_grid(t.gen, jobs::Stage::S5), 256, 0, stream>>>(span, total(&S5Totals::gen));
    if (t.score > 0) s5_score<<<persistent_grid(t.score, jobs::Stage::S5, 16), 256, 0, stream>>>(span, total(&S5Totals::score));
    s5_select<<<count, 256, 0, stream>>>(span);
    s5_finish<<<count, 256, 0, stream>>>(span);
    return hipGetLastError();
}

}  // namespace jobs

namespace s5 {

Status s5_workspace_bytes(const Config& c, std::size_t* bytes) {
    if (!bytes || c.hypotheses <= 0 || c.previous_hypotheses < 0 || c.previous_hypotheses > c.hypotheses ||
        c.batch_size <= 0 || c.batch_size > 65535 || c.refinement_steps < 0 || !std::isfinite(c.threshold) ||
        !(c.threshold > 0.0) || !std::isfinite(c.threshold * c.threshold))
        return Status::INVALID_ARGUMENT;
    const int models = c.hypotheses < jobs::kS5MaxModels ? c.hypotheses : jobs::kS5MaxModels;
    *bytes = jobs::kS5ModelOffset + static_cast<std::size_t>(models) * sizeof(jobs::S5Model);
    return Status::OK;
}

Status estimate(const Keypoints& a, const Keypoints& b, const Matches& m, const ScaleInfo& sa, const ScaleInfo& sb,
                const Config& c, std::uint8_t* mask, HomographyResult* output, void* ws, std::size_t bytes,
                hipStream_t stream) {
    if (!output) return Status::INVALID_ARGUMENT;
    jobs::S5Job j{};
    const Status st = jobs::make_s5_job(a, b, m, sa, sb, c, mask, ws, bytes, &j);
    if (st != Status::OK) return st;
    // More than kS5MaxModels new hypotheses run as successive extensions of the
    // saved raw search; the winner cannot depend on the split.
    for (int prev = c.previous_hypotheses; prev < c.hypotheses;) {
        const int next = c.hypotheses - prev > jobs::kS5MaxModels ? prev + jobs::kS5MaxModels : c.hypotheses;
        jobs::S5Job part = j;
        part.cfg.previous_hypotheses = prev;
        part.cfg.hypotheses = next;
        part.first = prev;
        part.count = next - prev;
        const jobs::S5Totals t = jobs::s5_prefix(&part, 1);
        if (jobs::launch_s5(nullptr, 1, t, &part, stream) != hipSuccess) return Status::CUDA_ERROR;
        prev = next;
    }
    if (hipMemcpyAsync(output, j.result_dev, sizeof(*output), hipMemcpyDeviceToHost, stream) != hipSuccess)
        return Status::CUDA_ERROR;
    return Status::OK;
}

}  // namespace s5
}  // namespace stitch
