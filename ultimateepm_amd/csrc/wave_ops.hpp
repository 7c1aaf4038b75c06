// wave_ops.hpp — wave-wide sums on gfx950 without the LDS crossbar: DPP row
// operations inside each 16-lane row, then the VALU row swaps
// v_permlane16_swap / v_permlane32_swap across the four rows.  64-bit values
// move as two 32-bit halves.  Every lane receives the total.
#pragma once

#include <hip/hip_runtime.h>

namespace uepm {
namespace wave {

template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// gfx950 row swaps, written out because the two-result builtins are not
// usable here.  permlane16: a' = [a.R0 b.R0 a.R2 b.R2], b' = [a.R1 b.R1 a.R3 b.R3]
// (R = 16-lane row); permlane32: a' = [a.lo b.lo], b' = [a.hi b.hi].  The s_nop
// pairs keep VALU writes and reads of the operands apart from the swap.
__device__ __forceinline__ void swap16(int& a, int& b) {
    asm volatile("s_nop 2\n\tv_permlane16_swap_b32 %0, %1\n\ts_nop 2" : "+v"(a), "+v"(b));
}
__device__ __forceinline__ void swap32(int& a, int& b) {
    asm volatile("s_nop 2\n\tv_permlane32_swap_b32 %0, %1\n\ts_nop 2" : "+v"(a), "+v"(b));
}

// A value as 32-bit words, for the moves above.
template <typename T>
struct Words;
template <>
struct Words<float> {
    int w[1];
    __device__ __forceinline__ explicit Words(float v) { w[0] = __builtin_bit_cast(int, v); }
    __device__ __forceinline__ float value() const { return __builtin_bit_cast(float, w[0]); }
    static constexpr int n = 1;
};
template <>
struct Words<double> {
    int w[2];
    __device__ __forceinline__ explicit Words(double v) {
        const long long b = __builtin_bit_cast(long long, v);
        w[0] = static_cast<int>(b);
        w[1] = static_cast<int>(b >> 32);
    }
    __device__ __forceinline__ double value() const {
        const long long b = (static_cast<long long>(w[1]) << 32) | static_cast<unsigned int>(w[0]);
        return __builtin_bit_cast(double, b);
    }
    static constexpr int n = 2;
};

template <int CTRL, typename T>
__device__ __forceinline__ T dpp_move(T v) {
    Words<T> x(v);
#pragma unroll
    for (int i = 0; i < Words<T>::n; ++i) x.w[i] = dpp<CTRL>(x.w[i]);
    return x.value();
}

// sum over the 16 lanes of each row; every lane of the row gets its row's total
template <typename T>
__device__ __forceinline__ T row_sum16(T v) {
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    v += dpp_move<0x140>(v);  // row_mirror
    return v;
}

// sum over the wave; every lane gets the total
template <typename T>
__device__ __forceinline__ T sum(T v) {
    v = row_sum16(v);
    Words<T> a(v), b(v);
#pragma unroll
    for (int i = 0; i < Words<T>::n; ++i) swap16(a.w[i], b.w[i]);
    const T s = a.value() + b.value();
    Words<T> c(s), d(s);
#pragma unroll
    for (int i = 0; i < Words<T>::n; ++i) swap32(c.w[i], d.w[i]);
    return c.value() + d.value();
}

}  // namespace wave
}  // namespace uepm
