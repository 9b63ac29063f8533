// spk_noalloc.hip -- spk_trace_noalloc: batched closest-hit / occlusion queries that allocate
// nothing, for callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// spk_trace allocates its traversal scratch lazily per scene: a per-thread overflow area for
// the traversal stack (entries past the LDS slots) and a work word.  Here that scratch is one
// file-scope device array, `scratch`, of kNoAllocScratchBytes per GPU, which every GPU holds
// from the library's load until the process ends.  Its device address is looked up once per
// device with hipGetSymbolAddress and only that address is handed to kernels.  The kernels
// and results are spk_trace's: rays are assigned to waves in fixed 64-ray chunks whatever the
// grid, so a grid shrunk to fit the scratch traces every ray exactly as spk_trace does.
//
// Nothing in the scratch needs clearing between calls: a thread writes a spill slot before it
// reads it, and the work word is not read by the trace kernel.  All calls on one GPU share the
// array, so calls that overlap on one GPU (two streams, two host threads) race on it.
#include <algorithm>
#include <mutex>

#include "spk_internal.h"

namespace spk {
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr size_t kNoAllocWorkBytes = 256;                   // the work word, padded
constexpr int kNoAllocBlock = 128;                          // launch_trace's block size
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct DeviceScratch {
    void* ptr = nullptr;  // device address of `scratch` on this GPU
    int full_grid = 0;    // spk_trace's persistent grid on this GPU
};
std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// Looks up this GPU's copy of `scratch` (and the trace grid) on the first call on a device;
// `device` is current.
int device_scratch(int device, DeviceScratch* out) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("spk_trace_noalloc: device %d beyond the %d devices this library tracks", device, kMaxDevices);
        return SPK_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[device];
    if (!d.ptr) {
        void* p = nullptr;
        SPK_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("spk_trace_noalloc: no device address for the scratch array on device %d", device);
            return SPK_ERR_HIP;
        }
        int cus = 0;
        SPK_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        d.full_grid = cus * std::max(1, trace_blocks_per_cu(kNoAllocBlock));
        d.ptr = p;
    }
    *out = d;
    return SPK_OK;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" int spk_trace_noalloc(spk_scene* s, const float* org, const float* dir, const float* tmax, int64_t n,
                                 int32_t any_hit, int32_t* prim_out, float* t_out, void* stream) {
    clear_error();
    if (!s || n < 0 || (n > 0 && (!org || !dir || !prim_out))) {
        set_error("spk_trace_noalloc: bad arguments");
        return SPK_ERR_INVALID;
    }
    if (n >= (int64_t(1) << 32) - 64) {
        set_error("spk_trace_noalloc: n too large");
        return SPK_ERR_INVALID;
    }
    SceneTraceInfo si;
    scene_trace_info(s, &si);
    // each persistent thread owns spill_cap stack words; a block of them must fit
    const size_t per_block = (size_t)kNoAllocBlock * (size_t)si.spill_cap * 4u;
    const size_t max_blocks = (kNoAllocScratchBytes - kNoAllocWorkBytes) / per_block;
    if (max_blocks < 1) {
        set_error("spk_trace_noalloc: the scene's BVH needs %zu bytes of stack overflow per block, more than the "
                  "%zu-byte scratch holds",
                  per_block, kNoAllocScratchBytes - kNoAllocWorkBytes);
        return SPK_ERR_INVALID;
    }
    if (n == 0) return SPK_OK;
    SPK_CHECK_HIP(hipSetDevice(si.device));
    DeviceScratch ds;
    int rc;
    if ((rc = device_scratch(si.device, &ds))) return rc;
    const int grid = (int)std::min<size_t>((size_t)ds.full_grid, max_blocks);
    TraceBufs b;
    b.org = org;
    b.dir = dir;
    b.tmax = tmax;
    b.prim_out = prim_out;
    b.t_out = t_out;
    b.n = n;
    b.any_hit = any_hit ? 1 : 0;
    uint32_t* work = reinterpret_cast<uint32_t*>(ds.ptr);
    uint32_t* spill = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(ds.ptr) + kNoAllocWorkBytes);
    return launch_trace(*si.dev, b, work, spill, si.spill_cap, grid, (hipStream_t)stream);
}
