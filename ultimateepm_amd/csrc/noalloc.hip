// noalloc.hip — the band solve without device allocation, for callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).
//
// uepm_solve_device allocates its plan's chunk workspace on first use and grows
// it when a larger solve needs more.  uepm_solve_device_noalloc runs the same
// kernels on a workspace carved out of `scratch`, a file-scope device array of
// kNoallocScratchBytes that every GPU the library runs on holds from its first
// uepm_ctx_create until the process ends.  The chunk is the largest number of
// k-points (at most 4096) whose workspace fits the array; one chunk is in flight
// at a time, on the caller's stream.
//
// In host code `scratch` names a host-side stand-in, not the device array: its
// device address is looked up once per device with hipGetSymbolAddress (at
// context creation, with that device current) and only that address reaches
// the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "uepm_internal.hpp"

using namespace uepm_host;

namespace {

constexpr size_t kNoallocScratchBytes = (size_t)UEPM_NOALLOC_SCRATCH_BYTES;
static_assert(kNoallocScratchBytes % 256 == 0, "the scratch array holds 256-byte aligned buffers");

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device (null: not looked up)

}  // namespace

namespace uepm_host {

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(UEPM_ERR_INVALID_ARG, "no-allocation scratch: device ordinal out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[device]) return UEPM_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(UEPM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
    return UEPM_OK;
}

}  // namespace uepm_host

extern "C" int uepm_solve_device_noalloc(uepm_plan* p, const double* d_k, int64_t nk, double* d_evals,
                                         double* d_evecs, double* d_grads, void* stream) {
    const int v = solve_args_ok(p, nk, d_k, d_evals, d_evecs, d_grads, "uepm_solve_device_noalloc");
    if (v != UEPM_OK || nk == 0) return v;
    const int device = p->ctx->device;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (device >= 0 && device < kMaxDevices) base = g_scratch[device];
    }
    if (!base) return fail(UEPM_ERR_INVALID_ARG, "uepm_solve_device_noalloc: the scratch array of the plan's "
                                                 "device was not set up (uepm_ctx_create)");
    // the largest chunk whose workspace fits the array
    const size_t per_k = plan_carve(p, 1, nullptr, nullptr);
    if (per_k > kNoallocScratchBytes)
        return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_solve_device_noalloc: the plan needs " + std::to_string(per_k) +
                                               " bytes of workspace per k-point, more than the " +
                                               std::to_string(kNoallocScratchBytes) + "-byte scratch array");
    int64_t chunk = std::min<int64_t>({nk, 4096, (int64_t)(kNoallocScratchBytes / per_k)});
    while (chunk > 1 && plan_carve(p, chunk, nullptr, nullptr) > kNoallocScratchBytes) --chunk;
    HIP_OK(hipSetDevice(device));
    LaneView view;
    plan_carve(p, chunk, static_cast<char*>(base), &view);
    return solve_chunks(p, d_k, nk, d_evals, d_evecs, d_grads, (hipStream_t)stream, &view, 1, chunk, true);
}
