// backtransform.hip — eigenvectors of H(k) from eigenvectors of the
// tridiagonal T: x = Q z with Q = H_0 H_1 ... H_{n-2}, H_j = I - tau_j v_j v_j^H
// (the reflectors of the Householder reduction), plus the Hellmann-Feynman
// gradient of each band (Hamiltonian.cpp:146-194).
//
//   cluster_orth_kernel  one wave per k-point: modified Gram-Schmidt of the
//                        T-eigenvectors inside clusters of close eigenvalues
//                        (inverse iteration leaves them only nearly orthogonal),
//                        then unit norm.
//   apply_q_kernel       one 16-lane DPP row per 1, 2 or 4 bands: lane rs holds
//                        rows rs + 16 q of its bands in registers, so the dot
//                        v_j^H x of every reflector is a 16-lane DPP sum (no LDS,
//                        no barriers) and each reflector row a lane loads serves
//                        all of its bands.  Reflector columns are read from the
//                        reduction's output (L2-resident, shared by the bands of
//                        a k-point) one step ahead of their use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>

#include "epm_physics.hpp"
#include "solver.hpp"

namespace uepm {

namespace {

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// sum over the 16 lanes of a DPP row (every lane gets the total)
__device__ __forceinline__ float row_sum16(float v) {
    v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    return v;
}
__device__ __forceinline__ double row_sum16(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return v;
}

template <typename R>
struct C2 {
    R x, y;
};
__device__ __forceinline__ C2<float> ld_c(const float2* p) {
    const float2 a = *p;
    return {a.x, a.y};
}
__device__ __forceinline__ C2<double> ld_c(const double2* p) {
    const double2 a = *p;
    return {a.x, a.y};
}

__device__ __forceinline__ double wave_sum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Modified Gram-Schmidt inside eigenvalue clusters (gap <= 1e-3 ||T||_inf), then
// unit norm; Z is [k][band][row], n rows, nb bands.
__global__ __launch_bounds__(64) void cluster_orth_kernel(const double* __restrict__ dall,
                                                          const double* __restrict__ eall,
                                                          const double* __restrict__ evals, double* __restrict__ Z,
                                                          int n, int nb) {
    const int64_t kk = blockIdx.x;
    const int lane = threadIdx.x;
    const double* d = dall + kk * n;
    const double* e = eall + kk * n;
    const double* ev = evals + kk * nb;
    double* z = Z + kk * (int64_t)nb * n;
    double tn = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double er = i + 1 < n ? fabs(e[i]) : 0.0;
        tn = fmax(tn, fabs(d[i]) + el + er);
    }
    for (int o = 32; o > 0; o >>= 1) tn = fmax(tn, __shfl_xor(tn, o, 64));
    const double ortol = 1e-3 * tn;
    int cstart = 0;
    for (int b = 0; b < nb; ++b) {
        if (b > 0 && ev[b] - ev[b - 1] > ortol) cstart = b;
        double* zb = z + (size_t)b * n;
        for (int c = cstart; c < b; ++c) {
            const double* zc = z + (size_t)c * n;
            double s = 0.0;
            for (int i = lane; i < n; i += 64) s += zc[i] * zb[i];
            s = wave_sum64(s);
            for (int i = lane; i < n; i += 64) zb[i] -= s * zc[i];
        }
        double nn = 0.0;
        for (int i = lane; i < n; i += 64) nn += zb[i] * zb[i];
        nn = wave_sum64(nn);
        const double inv = 1.0 / sqrt(nn);
        for (int i = lane; i < n; i += 64) zb[i] *= inv;
        __threadfence_block();  // rows of z_b written by other lanes are read for the next band
    }
}

// ---------------------------------------------------------------------------
// x = Q z for BPL bands per 16-lane DPP row (blockDim / 16 rows per workgroup;
// band (blockIdx.y * rows + row) * BPL + u); R is the arithmetic type, VT the
// storage of the reflectors (column j of V holds v_j at rows j+2..n-1;
// v_j[j+1] = 1).  A reflector row loaded by a lane serves its BPL bands.
template <typename R, typename VT, int ROWS, int BPL>
__global__ __launch_bounds__(256) void apply_q_kernel(const VT* __restrict__ V, int n, int ld,
                                                      const VT* __restrict__ tauall, const double* __restrict__ Z,
                                                      int nb, double2* __restrict__ X, double* __restrict__ grads,
                                                      DevParams P, const double* __restrict__ kpts,
                                                      const int32_t* __restrict__ G, const int* __restrict__ kstat,
                                                      unsigned long long* __restrict__ status, int64_t k0) {
    // the last stage of the FP64 form: a k-point the tridiagonal stage flagged
    // counts as failed (its eigenvalues and vectors are NaN)
    if (status && blockIdx.y == 0 && threadIdx.x == 0 && kstat[blockIdx.x]) {
        atomicAdd(&status[0], 1ull);
        atomicMin(&status[1], (unsigned long long)(k0 + blockIdx.x));
    }
    // register budget: prefetch the next reflector when two copies of v fit
    // beside the bands, keep one copy when it fits, else re-read v (an L1 hit)
    constexpr int CW = 2 * (int)sizeof(R) / 4;  // VGPRs per complex
    constexpr bool KEEP = (BPL + 1) * ROWS * CW <= 200;
    constexpr bool PREFETCH = (BPL + 2) * ROWS * CW <= 200;
    const int64_t kk = blockIdx.x;
    const int tid = threadIdx.x, rs = tid & 15;
    const int b0 = (blockIdx.y * (blockDim.x >> 4) + (tid >> 4)) * BPL;
    const VT* Vk = V + kk * (int64_t)ld * n;
    const VT* tau = tauall + kk * n;

    C2<R> x[BPL][ROWS];
#pragma unroll
    for (int u = 0; u < BPL; ++u) {
        const bool live = b0 + u < nb;
        const double* z = Z + (kk * nb + (live ? b0 + u : 0)) * (int64_t)n;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const int i = rs + 16 * q;
            x[u][q].x = (live && i < n) ? (R)z[i] : (R)0;
            x[u][q].y = (R)0;
        }
    }
    // v_j at this lane's row i; zero at rows <= j+1 and >= n (v_j[j+1] = 1 is set by the caller)
    auto vrow = [&](int j, int q) -> C2<R> {
        const int i = rs + 16 * q;
        const C2<R> a = ld_c(Vk + (int64_t)j * ld + (i < n ? i : 0));
        const bool ok = i > j + 1 && i < n;
        return {ok ? a.x : (R)0, ok ? a.y : (R)0};
    };
    C2<R> v[KEEP ? ROWS : 1], vn[PREFETCH ? ROWS : 1];
    if (KEEP && n >= 2) {
#pragma unroll
        for (int q = 0; q < ROWS; ++q) v[KEEP ? q : 0] = vrow(n - 2, q);
    }
    for (int j = n - 2; j >= 0; --j) {
        if (PREFETCH && j > 0) {
#pragma unroll
            for (int q = 0; q < ROWS; ++q) vn[PREFETCH ? q : 0] = vrow(j - 1, q);  // in flight during this step
        }
        if (KEEP && !PREFETCH) {
#pragma unroll
            for (int q = 0; q < ROWS; ++q) v[KEEP ? q : 0] = vrow(j, q);
        }
        const C2<R> t = ld_c(tau + j);
        if (!(t.x == (R)0 && t.y == (R)0)) {  // wave-uniform
            const int q0 = (j + 1) >> 4;  // first block of rows that can be nonzero
            const int r1 = (j + 1) & 15;  // lane holding row j+1 in block q0
            R sr[BPL], si[BPL];
#pragma unroll
            for (int u = 0; u < BPL; ++u) sr[u] = si[u] = (R)0;
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                if (q >= q0) {
                    C2<R> vq = KEEP ? v[KEEP ? q : 0] : vrow(j, q);
                    if (q == q0 && rs == r1) vq = {(R)1, (R)0};
                    // conj(v) x
#pragma unroll
                    for (int u = 0; u < BPL; ++u) {
                        sr[u] = fma(vq.x, x[u][q].x, fma(vq.y, x[u][q].y, sr[u]));
                        si[u] = fma(vq.x, x[u][q].y, fma(-vq.y, x[u][q].x, si[u]));
                    }
                }
            }
            R pr[BPL], pi[BPL];
#pragma unroll
            for (int u = 0; u < BPL; ++u) {
                const R ar = row_sum16(sr[u]), ai = row_sum16(si[u]);
                // s = tau (v^H x); x -= s v
                pr[u] = t.x * ar - t.y * ai;
                pi[u] = t.x * ai + t.y * ar;
            }
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                if (q >= q0) {
                    C2<R> vq = KEEP ? v[KEEP ? q : 0] : vrow(j, q);
                    if (q == q0 && rs == r1) vq = {(R)1, (R)0};
#pragma unroll
                    for (int u = 0; u < BPL; ++u) {
                        x[u][q].x = fma(-pr[u], vq.x, fma(pi[u], vq.y, x[u][q].x));
                        x[u][q].y = fma(-pr[u], vq.y, fma(-pi[u], vq.x, x[u][q].y));
                    }
                }
            }
        }
        if (PREFETCH) {
#pragma unroll
            for (int q = 0; q < ROWS; ++q) v[KEEP ? q : 0] = vn[PREFETCH ? q : 0];
        }
    }
#pragma unroll
    for (int u = 0; u < BPL; ++u) {
        const int b = b0 + u;
        if (b >= nb) break;
        double2* xo = X + (kk * nb + b) * (int64_t)n;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const int i = rs + 16 * q;
            if (i < n) xo[i] = make_double2((double)x[u][q].x, (double)x[u][q].y);
        }
        if (grads) {
            // Hellmann-Feynman on the kinetic term; spinor rows i >= N belong to G_{i-N}
            const double* k = kpts + 3 * kk;
            const int N = P.N;
            double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const int i = rs + 16 * q;
                if (i < n) {
                    const int g = i < N ? i : i - N;
                    const double w = (double)x[u][q].x * (double)x[u][q].x + (double)x[u][q].y * (double)x[u][q].y;
                    sx += (k[0] + (double)G[3 * g]) * w;
                    sy += (k[1] + (double)G[3 * g + 1]) * w;
                    sz += (k[2] + (double)G[3 * g + 2]) * w;
                }
            }
            sx = row_sum16(sx);
            sy = row_sum16(sy);
            sz = row_sum16(sz);
            if (rs == 0) {
                double* gg = grads + 3 * (kk * nb + b);
                gg[0] = sx * P.grad_pref * P.frac_to_phys;
                gg[1] = sy * P.grad_pref * P.frac_to_phys;
                gg[2] = sz * P.grad_pref * P.frac_to_phys;
            }
        }
    }
}

// ------------------------------------------------------------------ launchers
namespace {
// Bands per 16-lane row of apply_q_kernel (1, 2 or 4): as many as keep the
// kernel within its registers at two or more waves per SIMD -- complex64 4 up
// to 192 rows and 2 up to 256, complex128 2 up to 192.
int apply_bands_per_lane(bool f32, int n) {
    if (f32) return n <= 192 ? 4 : n <= 256 ? 2 : 1;
    return n <= 192 ? 2 : 1;
}

template <typename R, typename VT>
hipError_t launch_apply(const VT* dV, int n, int ld, const VT* dtau, const double* dZ, int nb, double2* dX,
                        double* dgrads, const DevParams& P, const double* dk, const int32_t* dG, int64_t nk,
                        const int* kstat, unsigned long long* status, int64_t k0, int bands_per_group,
                        hipStream_t s) {
    // bands_per_group (4, 8 or 16) bands per workgroup, BPL of them per 16-lane row
    const int cap = (bands_per_group == 4 || bands_per_group == 8) ? bands_per_group : 16;
    const int bpw = nb >= cap ? cap : ((nb + 3) / 4) * 4;
    // 1, 2 or 4 bands per row (a divisor of bpw, a multiple of 4) and at
    // least four rows, one full wave, per workgroup
    int bpl = apply_bands_per_lane(sizeof(R) == sizeof(float), n);
    while (bpl > 1 && bpw / bpl < 4) bpl >>= 1;
    const dim3 grid((unsigned)nk, (unsigned)((nb + bpw - 1) / bpw)), block(16 * (bpw / bpl));
#define UEPM_APPLY(ROWS, BPL)                                                                                       \
    hipLaunchKernelGGL((apply_q_kernel<R, VT, ROWS, BPL>), grid, block, 0, s, dV, n, ld, dtau, dZ, nb, dX, dgrads, \
                       P, dk, dG, kstat, status, k0)
#define UEPM_APPLY_ROWS(BPL)                    \
    if (n <= 64) UEPM_APPLY(4, BPL);            \
    else if (n <= 128) UEPM_APPLY(8, BPL);      \
    else if (n <= 192) UEPM_APPLY(12, BPL);     \
    else if (n <= 256) UEPM_APPLY(16, BPL);     \
    else if (n <= 384) UEPM_APPLY(24, BPL);     \
    else if (n <= 576) UEPM_APPLY(36, BPL);     \
    else return hipErrorInvalidValue;
    if (bpl >= 4) {
        UEPM_APPLY_ROWS(4)
    } else if (bpl == 2) {
        UEPM_APPLY_ROWS(2)
    } else {
        UEPM_APPLY_ROWS(1)
    }
#undef UEPM_APPLY_ROWS
#undef UEPM_APPLY
    return hipGetLastError();
}
}  // namespace

hipError_t launch_cluster_orth(const double* dd, const double* de, const double* devals, double* dZ, int n, int nb,
                               int64_t nk, hipStream_t s) {
    hipLaunchKernelGGL(cluster_orth_kernel, dim3((unsigned)nk), dim3(64), 0, s, dd, de, devals, dZ, n, nb);
    return hipGetLastError();
}

hipError_t launch_apply_q_f64(const double2* dV, int n, int ld, const double2* dtau, const double* dZ, int nb,
                              double2* dX, double* dgrads, const DevParams& P, const double* dk, const int32_t* dG,
                              int64_t nk, const int* kstat, unsigned long long* status, int64_t k0,
                              int bands_per_group, hipStream_t s) {
    return launch_apply<double, double2>(dV, n, ld, dtau, dZ, nb, dX, dgrads, P, dk, dG, nk, kstat, status, k0,
                                         bands_per_group, s);
}

hipError_t launch_apply_q_c64(const float2* dV, int n, int ld, const float2* dtau, const double* dZ, int nb,
                              double2* dX, const DevParams& P, const double* dk, const int32_t* dG, int64_t nk,
                              int bands_per_group, hipStream_t s) {
    return launch_apply<float, float2>(dV, n, ld, dtau, dZ, nb, dX, nullptr, P, dk, dG, nk, nullptr, nullptr, 0,
                                       bands_per_group, s);
}

}  // namespace uepm
