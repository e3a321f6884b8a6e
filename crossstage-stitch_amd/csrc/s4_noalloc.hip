// S4 matching without a caller workspace and without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The workspace s4_match needs (top-2 records, reverse minima, per-segment
// partial keys) lives in one file-scope device array of fixed size; its
// device address is looked up once per device with hipGetSymbolAddress and
// passed to the same kernels the workspace variant launches. The kernels
// write every workspace word they read, so nothing needs clearing between
// calls. Each GPU holds one copy of the array until the process ends, and
// every call on that GPU shares it: calls that overlap on one GPU (two
// streams, two threads) race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cs_stitch.h"
#include "stitch/stages.hpp"

using namespace stitch;

namespace {

constexpr std::size_t kScratchBytes = CS_S4_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "the scratch array keeps the workspace's 256-byte alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The scratch array's device address on the current device.
hipError_t scratch_address(void** out) {
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(g_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace

extern "C" int cs_s4_match_noalloc(const cs_keypoints* a, const cs_descriptors* da, const cs_keypoints* b,
                                   const cs_descriptors* db, const cs_s4_config* config, const cs_matches* output,
                                   void* stream) {
    if (!a || !da || !b || !db || !config || !output) return CS_INVALID_ARGUMENT;
    const s4::Config cfg{config->ratio_threshold, config->cross_check != 0};
    std::size_t needed = 0;
    const Status sized = s4::s4_workspace_bytes(a->capacity, b->capacity, cfg, &needed);
    if (sized != Status::OK) return static_cast<int>(sized);
    if (needed > kScratchBytes) return CS_INSUFFICIENT_WORKSPACE;  // capacities beyond the fixed array
    void* ws = nullptr;
    if (scratch_address(&ws) != hipSuccess) return CS_DEVICE_ERROR;
    const Keypoints ka{a->x, a->y, a->response, a->capacity, a->d_count};
    const Keypoints kb{b->x, b->y, b->response, b->capacity, b->d_count};
    const Descriptors ea{static_cast<DescriptorType>(da->type), da->dim, da->row_stride_bytes, da->data};
    const Descriptors eb{static_cast<DescriptorType>(db->type), db->dim, db->row_stride_bytes, db->data};
    const Matches m{output->idx_a, output->idx_b, output->distance, output->capacity, output->d_count};
    return static_cast<int>(s4::s4_match(ka, ea, kb, eb, cfg, m, ws, kScratchBytes, static_cast<hipStream_t>(stream)));
}
