// copy.hip — device-to-device copy with 16-byte vector loads and stores.
//
// Used by uepm_memcpy_dtod; its known traffic (read B, write B bytes) is also
// the calibration point for the HBM byte counters of the profiler.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "solver.hpp"

namespace uepm {

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16,
                                                   const uint8_t* __restrict__ tsrc, uint8_t* __restrict__ tdst,
                                                   int tail) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail) tdst[threadIdx.x] = tsrc[threadIdx.x];
}

hipError_t launch_copy_bytes(void* dst, const void* src, uint64_t bytes, int num_cus, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    // both pointers 16-byte aligned: vector body plus a byte tail of < 16
    const int64_t n16 = (int64_t)(bytes / 16);
    const int tail = (int)(bytes % 16);
    const int64_t want = (n16 + 255) / 256;
    const int blocks = (int)(want < (int64_t)num_cus * 8 ? (want > 0 ? want : 1) : (int64_t)num_cus * 8);
    copy_kernel<<<blocks, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16,
                                        static_cast<const uint8_t*>(src) + 16 * n16,
                                        static_cast<uint8_t*>(dst) + 16 * n16, tail);
    return hipGetLastError();
}

}  // namespace uepm
