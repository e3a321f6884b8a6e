// S4: exact brute-force top-2 matching with the strict ratio test, optional
// mutual cross-check and A-ordered compaction, for any number of pairs per
// launch. Semantics: src/s4/s4.cu:57-263 (ties to the lower train index,
// missing neighbours are (-1, +inf), d1 == 0 passes only with d2 > 0).
//
// MI355X fast path (256-bit binary descriptors): integer VALU work, no MFMA.
// A workgroup holds 512 queries in registers (4 waves x 64 lanes x 2) and
// streams a segment of train rows through LDS; per pair it does 8 XOR + 8
// popcount-accumulate, packs (distance << 20 | index) into one u32 key and
// keeps the two smallest keys branch-free - the key order is exactly the
// reference's (distance, index) order. With cross-check the same distances
// also feed a per-train-row minimum over queries (wave DPP reduction), so
// the reverse direction costs no second distance pass.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>

#include "stitch/jobs.hpp"

namespace stitch {
namespace {

using jobs::S4Job;
using jobs::kS4KeyShift;
using jobs::kS4NoKey;
using jobs::kS4QueriesPerLane;
using jobs::kS4QueryBlock;
using s4::Top2;

constexpr int kMinSegment = 64;
constexpr int kMaxSegment = 512;
constexpr unsigned kIndexMask = (1u << kS4KeyShift) - 1u;

__device__ __forceinline__ bool counts_valid(const S4Job& j, int* na, int* nb) {
    *na = *j.count_a;
    *nb = *j.count_b;
    return *na >= 0 && *na <= j.cap_a && *nb >= 0 && *nb <= j.cap_b;
}

__device__ __forceinline__ void insert_key(unsigned& k1, unsigned& k2, unsigned key) {
    const unsigned hi = max(k1, key);
    k1 = min(k1, key);
    k2 = min(k2, hi);
}

__device__ __forceinline__ unsigned dpp_min(unsigned v, unsigned o) { return v < o ? v : o; }

// Minimum over the 64 lanes, valid in lane 63 (quad swaps, row shifts, row
// broadcasts); returned as a wave-uniform value.
__device__ __forceinline__ unsigned wave_min_dpp(unsigned v) {
    v = dpp_min(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xb1, 0xf, 0xf, false)));
    v = dpp_min(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4e, 0xf, 0xf, false)));
    v = dpp_min(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false)));
    v = dpp_min(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false)));
    v = dpp_min(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xf, 0xf, false)));
    v = dpp_min(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xf, 0xf, false)));
    return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Distance tiles of the fast path: (query block of 512, train segment).
template <bool kCrossAny>
__global__ void __launch_bounds__(256) s4_dist(dev::JobSpan<S4Job> span, dev::Total work) {
    const long long total = work.get();
    __shared__ uint4 train[2 * kMaxSegment];
    __shared__ unsigned colmin[kMaxSegment];
    const int lane = dev::lane_id(), wave = threadIdx.x / 64;
    int ji = 0;
    const int njobs = span.jobs();
    // Tiles are sized by the keypoint capacities; only those inside the
    // detected counts hold work.
    dev::sparse_walk(
        total,
        [&](long long tk) {
            const S4Job& jj = dev::lane_job(span, tk, &S4Job::dist_begin, ji, njobs);
            const unsigned local = static_cast<unsigned>(tk - jj.dist_begin), segs = static_cast<unsigned>(jj.segs);
            int na, nb;
            return counts_valid(jj, &na, &nb) && static_cast<int>(local / segs) * kS4QueryBlock < na &&
                   static_cast<int>(local % segs) * jj.segment < nb;
        },
        [&](long long t) {
        ji = dev::next_job(span, t, &S4Job::dist_begin, ji, njobs);
        const S4Job& j = dev::job_at(span, ji);
        const unsigned local = static_cast<unsigned>(t - j.dist_begin), segs = static_cast<unsigned>(j.segs);
        const int qb = static_cast<int>(local / segs), sg = static_cast<int>(local % segs);
        int na, nb;
        if (!counts_valid(j, &na, &nb)) return;
        const int q0 = qb * kS4QueryBlock, t0 = sg * j.segment;
        if (q0 >= na || t0 >= nb) return;
        const bool kCross = kCrossAny && j.cross_check;
        const int rows = min(j.segment, nb - t0);
        for (int i = threadIdx.x; i < 2 * rows; i += 256)
            train[i] = *reinterpret_cast<const uint4*>(j.db + static_cast<size_t>(t0 + (i >> 1)) * j.stride_b + 16 * (i & 1));
        if (kCross)
            for (int i = threadIdx.x; i < rows; i += 256) colmin[i] = kS4NoKey;
        unsigned q[kS4QueriesPerLane][8];
        int qi[kS4QueriesPerLane];
        unsigned k1[kS4QueriesPerLane], k2[kS4QueriesPerLane];
#pragma unroll
        for (int k = 0; k < kS4QueriesPerLane; ++k) {
            qi[k] = q0 + wave * 64 * kS4QueriesPerLane + k * 64 + lane;
            k1[k] = k2[k] = kS4NoKey;
            if (qi[k] < na) {
                const uint4* p = reinterpret_cast<const uint4*>(j.da + static_cast<size_t>(qi[k]) * j.stride_a);
                const uint4 lo = p[0], hi = p[1];
                q[k][0] = lo.x; q[k][1] = lo.y; q[k][2] = lo.z; q[k][3] = lo.w;
                q[k][4] = hi.x; q[k][5] = hi.y; q[k][6] = hi.z; q[k][7] = hi.w;
            } else {
#pragma unroll
                for (int w = 0; w < 8; ++w) q[k][w] = 0;
            }
        }
        __syncthreads();
        for (int r = 0; r < rows; ++r) {
            const uint4 lo = train[2 * r], hi = train[2 * r + 1];
            const unsigned jg = static_cast<unsigned>(t0 + r);
            unsigned cmin = kS4NoKey;
#pragma unroll
            for (int k = 0; k < kS4QueriesPerLane; ++k) {
                unsigned d = __builtin_popcount(q[k][0] ^ lo.x);
                d += __builtin_popcount(q[k][1] ^ lo.y);
                d += __builtin_popcount(q[k][2] ^ lo.z);
                d += __builtin_popcount(q[k][3] ^ lo.w);
                d += __builtin_popcount(q[k][4] ^ hi.x);
                d += __builtin_popcount(q[k][5] ^ hi.y);
                d += __builtin_popcount(q[k][6] ^ hi.z);
                d += __builtin_popcount(q[k][7] ^ hi.w);
                insert_key(k1[k], k2[k], (d << kS4KeyShift) | jg);
                if (kCross) {
                    const unsigned ck = qi[k] < na ? ((d << kS4KeyShift) | static_cast<unsigned>(qi[k])) : kS4NoKey;
                    cmin = min(cmin, ck);
                }
            }
            if (kCross) {
                const unsigned wmin = wave_min_dpp(cmin);
                if (lane == 0 && wmin != kS4NoKey) atomicMin(&colmin[r], wmin);
            }
        }
#pragma unroll
        for (int k = 0; k < kS4QueriesPerLane; ++k)
            if (qi[k] < na)
                reinterpret_cast<uint2*>(j.part_fwd)[static_cast<size_t>(sg) * j.cap_a + qi[k]] = make_uint2(k1[k], k2[k]);
        if (kCross) {
            __syncthreads();
            for (int i = threadIdx.x; i < rows; i += 256)
                j.part_col[static_cast<size_t>(qb) * j.cap_b + t0 + i] = colmin[i];
        }
        __syncthreads();
        });
}

__device__ Top2 empty_top2() { return {INFINITY, INFINITY, -1, -1}; }

// Forward merge over train segments, one thread per query; writes Top2.
__global__ void __launch_bounds__(256) s4_qmerge(dev::JobSpan<S4Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        ji = dev::next_job(span, t, &S4Job::qmerge_begin, ji, njobs);
        const S4Job& j = dev::job_at(span, ji);
        const int q = static_cast<int>(t - j.qmerge_begin) * 256 + threadIdx.x;
        if (q >= j.cap_a) continue;
        int na, nb;
        Top2 out = empty_top2();
        if (counts_valid(j, &na, &nb) && q < na) {
            unsigned k1 = kS4NoKey, k2 = kS4NoKey;
            const int segs = (nb + j.segment - 1) / j.segment;
            for (int s = 0; s < segs; ++s) {
                const uint2 p = reinterpret_cast<const uint2*>(j.part_fwd)[static_cast<size_t>(s) * j.cap_a + q];
                insert_key(k1, k2, p.x);
                insert_key(k1, k2, p.y);
            }
            if (k1 != kS4NoKey) {
                out.first_distance = static_cast<double>(k1 >> kS4KeyShift);
                out.first_index = static_cast<int>(k1 & kIndexMask);
            }
            if (k2 != kS4NoKey) {
                out.second_distance = static_cast<double>(k2 >> kS4KeyShift);
                out.second_index = static_cast<int>(k2 & kIndexMask);
            }
        }
        j.forward[q] = out;
    }
}

// Reverse top-1 per train row: minimum key over query blocks.
__global__ void __launch_bounds__(256) s4_cmerge(dev::JobSpan<S4Job> span, dev::Total work) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        ji = dev::next_job(span, t, &S4Job::cmerge_begin, ji, njobs);
        const S4Job& j = dev::job_at(span, ji);
        const int r = static_cast<int>(t - j.cmerge_begin) * 256 + threadIdx.x;
        if (r >= j.cap_b) continue;
        int na, nb;
        int first = -1;
        if (counts_valid(j, &na, &nb) && r < nb) {
            unsigned m = kS4NoKey;
            const int qbs = (na + kS4QueryBlock - 1) / kS4QueryBlock;
            for (int b = 0; b < qbs; ++b) m = min(m, j.part_col[static_cast<size_t>(b) * j.cap_b + r]);
            if (m != kS4NoKey) first = static_cast<int>(m & kIndexMask);
        }
        j.rev_first[r] = first;
    }
}

__device__ bool before(double d, int i, double od, int oi) {
    return i >= 0 && (oi < 0 || d < od || (d == od && i < oi));
}

__device__ void insert_top2(Top2& p, double d, int i) {
    if (i < 0 || i == p.first_index || i == p.second_index) return;
    if (before(d, i, p.first_distance, p.first_index)) {
        p.second_distance = p.first_distance;
        p.second_index = p.first_index;
        p.first_distance = d;
        p.first_index = i;
    } else if (before(d, i, p.second_distance, p.second_index)) {
        p.second_distance = d;
        p.second_index = i;
    }
}

// Generic path (FLOAT32, other dimensions or alignments): one thread per
// query, distances accumulated in FP64 in dimension order as the reference.
// reverse == 1 searches B rows against A and writes reverse/rev_first.
__global__ void __launch_bounds__(256) s4_generic(dev::JobSpan<S4Job> span, dev::Total work, int reverse,
                                                  Top2* external) {
    const long long total = work.get();
    int ji = 0;
    const int njobs = span.jobs();
    for (long long t = blockIdx.x; t < total; t += gridDim.x) {
        ji = dev::next_job(span, t, reverse ? &S4Job::grev_begin : &S4Job::gfwd_begin, ji, njobs);
        const S4Job& j = dev::job_at(span, ji);
        const long long first_tile = reverse ? j.grev_begin : j.gfwd_begin;
        const int q = static_cast<int>(t - first_tile) * 256 + threadIdx.x;
        const int cap_q = reverse ? j.cap_b : j.cap_a;
        if (q >= cap_q) continue;
        int na, nb;
        Top2 best = empty_top2();
        if (counts_valid(j, &na, &nb)) {
            const int nq = reverse ? nb : na, nt = reverse ? na : nb;
            const std::uint8_t* qd = reverse ? j.db : j.da;
            const std::uint8_t* td = reverse ? j.da : j.db;
            const int qs = reverse ? j.stride_b : j.stride_a, ts = reverse ? j.stride_a : j.stride_b;
            if (q < nq) {
                const std::uint8_t* qrow = qd + static_cast<size_t>(q) * qs;
                for (int r = 0; r < nt; ++r) {
                    const std::uint8_t* trow = td + static_cast<size_t>(r) * ts;
                    double d = 0;
                    if (j.binary) {
                        for (int k = 0; k < j.dim; ++k) d += __builtin_popcount(static_cast<unsigned>(qrow[k] ^ trow[k]));
                    } else {
                        const float* a = reinterpret_cast<const float*>(qrow);
                        const float* b = reinterpret_cast<const float*>(trow);
                        for (int k = 0; k < j.dim; ++k) {
                            const double delta = static_cast<double>(a[k]) - b[k];
                            d = d + delta * delta;
                        }
                    }
                    insert_top2(best, d, r);
                }
            }
        }
        if (external) external[q] = best;
        else if (reverse) {
            j.reverse[q] = best;
            j.rev_first[q] = best.first_index;
        } else {
            j.forward[q] = best;
        }
    }
}

// Accept and compact in A order; one workgroup of 1024 threads per job.
__device__ __forceinline__ void s4_compact_job(const S4Job& j) {
    __shared__ int wave_tot[16];
    const int tid = threadIdx.x, lane = dev::lane_id(), wave = tid / 64;
    const int per = (j.cap_a + 1023) / 1024;
    const int first = tid * per;
    int mine = 0;
    for (int i = 0; i < per; ++i) {
        const int q = first + i;
        if (q >= j.cap_a) break;
        const Top2 p = j.forward[q];
        bool accept = p.second_index >= 0 &&
                      (p.first_distance == 0 ? p.second_distance > 0 : p.first_distance < j.ratio * p.second_distance);
        if (accept && j.cross_check) accept = j.rev_first[p.first_index] == q;
        mine += accept;
    }
    int incl = mine;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int w = 0; w < 16; ++w) {
        if (w < wave) base += wave_tot[w];
        total += wave_tot[w];
    }
    int pos = base + incl - mine;
    for (int i = 0; i < per && mine; ++i) {
        const int q = first + i;
        if (q >= j.cap_a) break;
        const Top2 p = j.forward[q];
        bool accept = p.second_index >= 0 &&
                      (p.first_distance == 0 ? p.second_distance > 0 : p.first_distance < j.ratio * p.second_distance);
        if (accept && j.cross_check) accept = j.rev_first[p.first_index] == q;
        if (!accept) continue;
        j.idx_a[pos] = q;
        j.idx_b[pos] = p.first_index;
        const double d = j.binary ? p.first_distance : sqrt(p.first_distance);
        j.distance[pos] = d > FLT_MAX ? INFINITY : static_cast<float>(d);
        ++pos;
    }
    if (tid == 0) *j.d_count = total;
}

__global__ void __launch_bounds__(1024) s4_compact(dev::JobSpan<S4Job> span) {
    const int jobs = span.jobs();
    for (int ji = blockIdx.x; ji < jobs; ji += gridDim.x) {
        s4_compact_job(dev::job_at(span, ji));
        __syncthreads();
    }
}

bool valid_input(const Keypoints& k, const Descriptors& d) {
    if (k.capacity < 0 || !k.d_count || reinterpret_cast<std::uintptr_t>(k.d_count) % alignof(int) != 0 ||
        d.dim <= 0 || d.row_stride_bytes <= 0)
        return false;
    std::size_t element;
    if (d.type == DescriptorType::BINARY_U8) element = 1;
    else if (d.type == DescriptorType::FLOAT32) element = sizeof(float);
    else return false;
    if (static_cast<std::size_t>(d.dim) > std::numeric_limits<std::size_t>::max() / element) return false;
    const auto row = static_cast<std::size_t>(d.dim) * element;
    if (row > static_cast<std::size_t>(d.row_stride_bytes) || static_cast<std::size_t>(d.row_stride_bytes) % element ||
        (k.capacity && !d.data) || reinterpret_cast<std::uintptr_t>(d.data) % element)
        return false;
    return true;
}

bool valid_pair(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db) {
    return valid_input(a, da) && valid_input(b, db) && da.type == db.type && da.dim == db.dim;
}

std::size_t align256(std::size_t v) { return (v + 255) / 256 * 256; }

int g_segment = 128;  // train rows per LDS segment (set by the autotuner)

}  // namespace

namespace jobs {

void s4_set_segment(int rows) {
    if (rows >= kMinSegment && rows <= kMaxSegment && rows % 64 == 0) g_segment = rows;
}
int s4_segment() { return g_segment; }

namespace {
int g_workgroups[kStages] = {8, 8, 8, 8, 8, 8};
int g_s1_rows = kS1TileH;
}
void set_workgroups_per_cu(int eighths) {
    for (int s = 0; s < kStages; ++s) set_workgroups_per_cu(static_cast<Stage>(s), eighths);
}
void set_workgroups_per_cu(Stage stage, int eighths) {
    const int s = static_cast<int>(stage);
    if (s >= 0 && s < kStages && eighths >= 1 && eighths <= 64) g_workgroups[s] = eighths;
}
int workgroups_per_cu(Stage stage) {
    const int s = static_cast<int>(stage);
    return s >= 0 && s < kStages ? g_workgroups[s] : 8;
}
void s1_set_tile_rows(int rows) {
    if (rows == 16 || rows == 32 || rows == 64) g_s1_rows = rows;
}
int s1_tile_rows() { return g_s1_rows; }

// Offsets: forward, reverse, rev_first, part_fwd, part_col.
std::size_t s4_layout_bytes(int cap_a, int cap_b, bool cc, int segment, std::size_t off[5]) {
    const std::size_t a = cap_a, b = cap_b;
    const std::size_t segs = (b + segment - 1) / segment, qbs = (a + kS4QueryBlock - 1) / kS4QueryBlock;
    const std::size_t sizes[5] = {a * sizeof(Top2), cc ? b * sizeof(Top2) : 0, cc ? b * 4 : 0, segs * a * 8,
                                  cc ? qbs * b * 4 : 0};
    std::size_t cur = 0;
    for (int i = 0; i < 5; ++i) {
        cur = align256(cur);
        off[i] = cur;
        cur += sizes[i];
    }
    return cur;
}

Status make_s4_job(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db,
                   const s4::Config& cfg, const Matches& out, void* ws, std::size_t ws_bytes, S4Job* job) {
    std::size_t required = 0;
    if (s4::s4_workspace_bytes(a.capacity, b.capacity, cfg, &required) != Status::OK || !valid_pair(a, da, b, db) ||
        out.capacity < a.capacity || !out.d_count || reinterpret_cast<std::uintptr_t>(out.d_count) % alignof(int) ||
        (a.capacity && (!out.idx_a || !out.idx_b || !out.distance)))
        return Status::INVALID_ARGUMENT;
    if (ws_bytes < required || (required && !ws)) return Status::INSUFFICIENT_WORKSPACE;
    if (required && reinterpret_cast<std::uintptr_t>(ws) % 256) return Status::INVALID_ARGUMENT;
    S4Job j{};
    j.da = static_cast<const std::uint8_t*>(da.data);
    j.db = static_cast<const std::uint8_t*>(db.data);
    j.stride_a = da.row_stride_bytes;
    j.stride_b = db.row_stride_bytes;
    j.count_a = a.d_count;
    j.count_b = b.d_count;
    j.cap_a = a.capacity;
    j.cap_b = b.capacity;
    j.dim = da.dim;
    j.binary = da.type == DescriptorType::BINARY_U8;
    j.cross_check = cfg.cross_check;
    j.ratio = j.binary ? cfg.ratio_threshold : cfg.ratio_threshold * cfg.ratio_threshold;
    j.fast = j.binary && j.dim == 32 && reinterpret_cast<std::uintptr_t>(j.da) % 16 == 0 &&
             reinterpret_cast<std::uintptr_t>(j.db) % 16 == 0 && j.stride_a % 16 == 0 && j.stride_b % 16 == 0 &&
             j.cap_a <= static_cast<int>(kIndexMask) && j.cap_b <= static_cast<int>(kIndexMask);
    j.segment = g_segment;
    j.qblocks = (j.cap_a + kS4QueryBlock - 1) / kS4QueryBlock;
    j.segs = (j.cap_b + j.segment - 1) / j.segment;
    if (j.cap_a) {
        std::size_t off[5];
        s4_layout_bytes(j.cap_a, j.cap_b, j.cross_check, j.segment, off);
        auto* base = static_cast<std::uint8_t*>(ws);
        j.forward = reinterpret_cast<Top2*>(base + off[0]);
        j.reverse = reinterpret_cast<Top2*>(base + off[1]);
        j.rev_first = reinterpret_cast<int*>(base + off[2]);
        j.part_fwd = reinterpret_cast<std::uint32_t*>(base + off[3]);
        j.part_col = reinterpret_cast<std::uint32_t*>(base + off[4]);
    }
    j.idx_a = out.idx_a;
    j.idx_b = out.idx_b;
    j.distance = out.distance;
    j.d_count = out.d_count;
    *job = j;
    return Status::OK;
}

S4Totals s4_prefix(S4Job* js, int n) {
    S4Totals t{};
    for (int i = 0; i < n; ++i) {
        S4Job& j = js[i];
        j.dist_begin = t.dist;
        j.qmerge_begin = t.qmerge;
        j.cmerge_begin = t.cmerge;
        j.gfwd_begin = t.generic_fwd;
        j.grev_begin = t.generic_rev;
        if (j.fast && j.cap_a) {
            t.dist += static_cast<long long>(j.qblocks) * j.segs;
            t.qmerge += (j.cap_a + 255) / 256;
            if (j.cross_check) t.cmerge += (j.cap_b + 255) / 256;
        } else {
            t.generic_fwd += (j.cap_a + 255) / 256;
            if (j.cross_check) t.generic_rev += (j.cap_b + 255) / 256;
        }
    }
    return t;
}

hipError_t launch_s4(const S4Job* table, int count, const S4Totals& t, const S4Job* single, hipStream_t stream,
                     const dev::Replay<S4Totals>* replay) {
    dev::JobSpan<S4Job> span{table, count, single ? *single : S4Job{}, replay ? replay->count : nullptr};
    const auto total = [&](long long S4Totals::*m) {
        return dev::Total(t.*m, replay ? &(replay->totals->*m) : nullptr);
    };
    // Tables may mix jobs with and without cross-check; the kernel tests each job.
    const bool any_cross = single ? static_cast<bool>(single->cross_check) : true;
    if (t.dist > 0) {
        if (any_cross) s4_dist<true><<<persistent_grid(t.dist, jobs::Stage::S4), 256, 0, stream>>>(span, total(&S4Totals::dist));
        else s4_dist<false><<<persistent_grid(t.dist, jobs::Stage::S4), 256, 0, stream>>>(span, total(&S4Totals::dist));
        s4_qmerge<<<persistent_grid(t.qmerge, jobs::Stage::S4), 256, 0, stream>>>(span, total(&S4Totals::qmerge));
        if (t.cmerge > 0) s4_cmerge<<<persistent_grid(t.cmerge, jobs::Stage::S4), 256, 0, stream>>>(span, total(&S4Totals::cmerge));
    }
    if (t.generic_fwd > 0)
        s4_generic<<<persistent_grid(t.generic_fwd, jobs::Stage::S4), 256, 0, stream>>>(span, total(&S4Totals::generic_fwd), 0, nullptr);
    if (t.generic_rev > 0)
        s4_generic<<<persistent_grid(t.generic_rev, jobs::Stage::S4), 256, 0, stream>>>(span, total(&S4Totals::generic_rev), 1, nullptr);
    s4_compact<<<count, 1024, 0, stream>>>(span);
    return hipGetLastError();
}

}  // namespace jobs

namespace s4 {

Status s4_workspace_bytes(int capacity_a, int capacity_b, const Config& config, std::size_t* bytes) {
    if (!bytes) return Status::INVALID_ARGUMENT;
    *bytes = 0;
    if (capacity_a < 0 || capacity_b < 0 || !std::isfinite(config.ratio_threshold) || config.ratio_threshold <= 0 ||
        config.ratio_threshold >= 1)
        return Status::INVALID_ARGUMENT;
    if (capacity_a == 0) return Status::OK;
    std::size_t off[5];
    // Sized for the smallest segment so any tuned segment fits.
    *bytes = jobs::s4_layout_bytes(capacity_a, capacity_b, config.cross_check, kMinSegment, off);
    return Status::OK;
}

Status s4_knn_top2(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db,
                   Top2* output, hipStream_t stream) {
    if (!valid_pair(a, da, b, db) ||
        (a.capacity && (!output || reinterpret_cast<std::uintptr_t>(output) % alignof(Top2))))
        return Status::INVALID_ARGUMENT;
    if (!a.capacity) return Status::OK;
    S4Job j{};
    j.da = static_cast<const std::uint8_t*>(da.data);
    j.db = static_cast<const std::uint8_t*>(db.data);
    j.stride_a = da.row_stride_bytes;
    j.stride_b = db.row_stride_bytes;
    j.count_a = a.d_count;
    j.count_b = b.d_count;
    j.cap_a = a.capacity;
    j.cap_b = b.capacity;
    j.dim = da.dim;
    j.binary = da.type == DescriptorType::BINARY_U8;
    const long long tiles = (j.cap_a + 255) / 256;
    s4_generic<<<persistent_grid(tiles, jobs::Stage::S4), 256, 0, stream>>>(dev::JobSpan<S4Job>{nullptr, 1, j}, tiles, 0, output);
    return hipGetLastError() == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

Status s4_match(const Keypoints& a, const Descriptors& da, const Keypoints& b, const Descriptors& db,
                const Config& config, Matches output, void* workspace, std::size_t workspace_bytes,
                hipStream_t stream) {
    S4Job j{};
    const Status s = jobs::make_s4_job(a, da, b, db, config, output, workspace, workspace_bytes, &j);
    if (s != Status::OK) return s;
    if (!a.capacity) {
        // Zero matches; a 1-thread fill keeps the call capture-safe.
        S4Job z = j;
        z.cap_a = 0;
        s4_compact<<<1, 1024, 0, stream>>>(dev::JobSpan<S4Job>{nullptr, 1, z});
        return hipGetLastError() == hipSuccess ? Status::OK : Status::CUDA_ERROR;
    }
    const jobs::S4Totals t = jobs::s4_prefix(&j, 1);
    return jobs::launch_s4(nullptr, 1, t, &j, stream) == hipSuccess ? Status::OK : Status::CUDA_ERROR;
}

Status s4_forward_neighbors(int capacity_a, int capacity_b, const Config& config, const void* workspace,
                            std::size_t workspace_bytes, const Top2** output) {
    std::size_t required = 0;
    if (!output || s4_workspace_bytes(capacity_a, capacity_b, config, &required) != Status::OK || !capacity_a ||
        !workspace || workspace_bytes < required || reinterpret_cast<std::uintptr_t>(workspace) % 256)
        return Status::INVALID_ARGUMENT;
    *output = static_cast<const Top2*>(workspace);
    return Status::OK;
}

}  // namespace s4
}  // namespace stitch
