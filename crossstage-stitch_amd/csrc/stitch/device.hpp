// Device-side helpers shared by the stage kernels: job spans for ragged
// multi-image launches, wave64 reductions and launch bookkeeping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "stitch/tuning.hpp"

namespace stitch::dev {

constexpr int kWave = 64;

// Load of launch-constant data (job tables, counts, totals) through the
// constant address space: the compiler then knows the value is the same for
// every lane of a wave when the address is, and uses scalar loads. A load
// through a plain (generic) pointer could reach per-lane scratch memory, so
// everything computed from it would be treated as divergent.
template <class T>
__device__ __forceinline__ T ld_const(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// A launch covers `count` jobs (images, pairs or canvases). One job may be
// passed by value (single-image stage calls, no upload needed); batches pass a
// device table. Each job records, per kernel, the first tile it owns in that
// kernel's flattened tile range, so one persistent grid walks every job.
template <class Job>
struct JobSpan {
    const Job* table;  // device table, or nullptr to use `single`
    int count;
    Job single;
    const int* count_dev = nullptr;  // graph replay: job count read on the device
    __device__ int jobs() const { return count_dev ? ld_const(count_dev) : count; }
};

// Work size (tiles) of one kernel launch. Direct launches pass the value;
// a captured graph is replayed with new job tables, so its kernels read the
// size from device memory written before each replay, and their grids are
// sized for the capacity the graph was captured for.
struct Total {
    long long value;
    const long long* dev;
    __host__ __device__ Total(long long v, const long long* d = nullptr) : value(v), dev(d) {}
    __device__ long long get() const { return dev ? ld_const(dev) : value; }
};

// Device copies of a launch's job count and per-kernel totals (graph replay).
template <class Totals>
struct Replay {
    const int* count;
    const Totals* totals;
};

template <class Job>
__device__ inline const Job& job_at(const JobSpan<Job>& span, int i) {
    // Job tables and the by-value job in the kernel arguments are read-only
    // for the whole launch: address both through the constant address space,
    // so the (wave-uniform) job fields are fetched with scalar loads instead
    // of one vector load per lane through a generic pointer.
    //
    // Every kernel taking a JobSpan takes it as its first argument, at offset
    // 0 of the kernel argument segment, so the by-value job's address is the
    // segment pointer plus offsetof(single). The chosen address goes through
    // readfirstlane: the compiler then reads each field with one scalar load
    // from whichever record applies. (Selecting between the table and
    // &span.single directly lets it keep every field of the by-value copy in
    // SGPRs and branch on table/single at each field access, which costs
    // registers - spills in the larger kernels - and scalar instructions.)
    using CJob = const __attribute__((address_space(4))) Job;
    const std::uint64_t single =
        reinterpret_cast<std::uint64_t>(__builtin_amdgcn_kernarg_segment_ptr()) + offsetof(JobSpan<Job>, single);
    const std::uint64_t chosen = span.table ? reinterpret_cast<std::uint64_t>(span.table + i) : single;
    const std::uint64_t u =
        static_cast<std::uint64_t>(static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(chosen)))) |
        static_cast<std::uint64_t>(static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(chosen >> 32))))
            << 32;
    return *(const Job*)reinterpret_cast<CJob*>(u);
}

// Index of the job owning flattened tile `t` for the kernel whose prefix is
// the member `begin` (jobs are stored in increasing prefix order).
template <class Job>
__device__ inline int find_job(const JobSpan<Job>& span, long long t, long long Job::*begin) {
    if (!span.table) return 0;
    int lo = 0, hi = span.jobs() - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (span.table[mid].*begin <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Job owning tile t when a workgroup visits its tiles in increasing order,
// from the previous answer `ji` (same result as find_job: the last job whose
// prefix is <= t). The common case - t still in job ji - costs one load;
// otherwise a binary search over the jobs after ji. (A linear walk from job
// 0 costs one dependent scalar load per job skipped, which dominates kernels
// whose workgroups take one or two tiles each.)
template <class Begin>
__device__ inline int next_job_at(long long t, Begin begin_of, int ji, int jobs) {
    if (ji + 1 >= jobs || begin_of(ji + 1) > t) return ji;
    int lo = ji + 1, hi = jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin_of(mid) <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class Job>
__device__ inline int next_job(const JobSpan<Job>& span, long long t, long long Job::*begin, int ji, int jobs) {
    if (!span.table) return 0;
    return next_job_at(t, [&](int i) { return ld_const(&(span.table[i].*begin)); }, ji, jobs);
}

// Tile walk of a persistent grid, XCD by XCD. The dispatcher hands
// workgroups to the chip's 8 XCDs in turn (workgroup g to XCD (g + c) % 8),
// and each XCD has its own L2. A plain grid-stride walk puts neighbouring
// tiles, which share halo rows and source lines, on different XCDs, so each
// L2 fetches those lines again. Here the workgroups that share g % 8 take
// one contiguous eighth of the tile range between them, striding by
// gridDim.x / 8 through it: neighbours stay in one L2. Every workgroup still
// visits its tiles in increasing order (next_job relies on that), and the
// tiles visited are the same set, each once. A grid that is not a multiple
// of 8 falls back to the plain walk.
struct TileWalk {
    long long t, end;
    int step;
};

__device__ inline TileWalk tile_walk(long long total) {
    constexpr int kXcds = 8;
    const int g = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    if (g % kXcds != 0) return {b, total, g};
    const int x = b % kXcds;
    return {total * x / kXcds + b / kXcds, total * (x + 1) / kXcds, g / kXcds};
}

__device__ inline int lane_id() { return __lane_id(); }

// A wave-uniform value (a job field) held in scalar registers from here on.
// Loads of launch-constant data are invariant, so the compiler is free to
// sink each one into the conditional block that first uses it, where it
// then waits for that one load on its own; a value passed through here is
// loaded at this point, together with the others, under one wait.
template <class T>
__device__ __forceinline__ T hold(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "hold: 32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        int u = __builtin_bit_cast(int, v);
        asm volatile("" : "+s"(u));
        return __builtin_bit_cast(T, u);
    } else {
        long long u = __builtin_bit_cast(long long, v);
        asm volatile("" : "+s"(u));
        return __builtin_bit_cast(T, u);
    }
}

// The same for a pointer into device (global) memory, which it stays known
// to be, so accesses through it remain global rather than flat ones.
template <class T>
__device__ __forceinline__ T* hold(T* p) {
    long long u = reinterpret_cast<long long>(p);
    asm volatile("" : "+s"(u));
    using G = __attribute__((address_space(1))) T;
    return (T*)reinterpret_cast<G*>(u);
}

// Job owning tile t, looked up by one lane on its own (vector loads, a
// binary search over the jobs from `from`), for predicates that lanes
// evaluate over different tiles at once.
template <class Job>
__device__ inline const Job& lane_job(const JobSpan<Job>& span, long long t, long long Job::*begin, int from, int jobs) {
    if (!span.table) return span.single;
    int lo = from, hi = jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (span.table[mid].*begin <= t) lo = mid;
        else hi = mid - 1;
    }
    return span.table[lo];
}

// Walks the workgroup's tiles blockIdx.x + k * gridDim.x (k = 0, 1, ...) in
// groups of 64: lane k evaluates keep(t) for the group's k-th tile, and
// body(t) runs, in increasing order, for the tiles kept - the same list in
// every wave, so body may use barriers. For kernels whose tile range is
// sized by a capacity and mostly empty: one memory round trip per 64 tiles
// finds the work, instead of one per tile.
template <class Keep, class Body>
__device__ inline void sparse_walk(long long total, Keep keep, Body body) {
    const long long stride = gridDim.x;
    for (long long base = blockIdx.x; base < total; base += 64 * stride) {
        const long long tk = base + static_cast<long long>(lane_id()) * stride;
        const bool kept = tk < total && keep(tk);
        for (unsigned long long todo = __ballot(kept); todo; todo &= todo - 1)
            body(base + static_cast<long long>(__ffsll(static_cast<long long>(todo)) - 1) * stride);
    }
}

template <class T>
__device__ inline T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline unsigned wave_min_u32(unsigned v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

// Exclusive prefix of a per-lane predicate within the wave, and the wave total.
__device__ inline int wave_prefix(bool pred, int* total) {
    const unsigned long long mask = __ballot(pred);
    *total = __popcll(mask);
    return __popcll(mask & ((1ull << lane_id()) - 1ull));
}

}  // namespace stitch::dev

// Grid size for a persistent launch over `tiles` tiles of `threads` threads:
// enough workgroups to fill every CU a few times, never more than there are tiles.
inline unsigned persistent_grid(long long tiles, stitch::jobs::Stage stage, int per_cu = 8) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop{};
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    long long cap = static_cast<long long>(cus) * per_cu * stitch::jobs::workgroups_per_cu(stage) / 8;
    if (cap < cus) cap = cus;
    const long long g = tiles < cap ? tiles : cap;
    return static_cast<unsigned>(g > 0 ? g : 1);
}
