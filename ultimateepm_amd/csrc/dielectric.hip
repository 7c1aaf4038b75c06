// Wave-vector and energy dependent dielectric function from the eigenpairs at
// k and k + q (DielectricFunction::compute_dielectric_function,
// src/EPP/DielectricFunction.cpp:83-165):
//
//   S(q, E) = sum_k sum_{c >= 4} sum_{v < 4} |x_c(k+q) . x_v(k)|^2
//             [ (D - E) / ((D - E)^2 + eta^2) + (D + E) / ((D + E)^2 + eta^2) ],
//   D = E_c(k+q) - E_v(k).
//
// The overlap is the reference's col(c).adjoint().dot(col(v)) in Eigen, whose
// dot conjugates its left operand again: the plain bilinear sum
// sum_i x_c,i x_v,i over the plane-wave coefficients.
//
// Two kernels.  dielectric_overlap_kernel: one wave per k-point forms the
// (nb - 4) x 4 overlaps from the band-contiguous eigenvectors of the two
// solves (the valence vectors stay in L1/L2 across the conduction loop) and
// writes weight and energy difference per pair; an optional per-k factor
// scales the weights (the tetrahedron volume share of a mesh vertex in
// BZ_States::compute_dielectric_function, src/BZ_MESH/bz_states.cpp:95-164).
// dielectric_lorentz_kernel: energy groups x slices of the pair list,
// per-thread partial sums in registers, a fixed reduction order, partials per
// slice; the finish kernel adds the slices in order into the running sum of
// the q row.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "solver.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kEnergiesPerGroup = 8;
constexpr int kValence = 4;  // index_first_conduction_band (DielectricFunction.cpp:84)

__device__ inline double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__global__ void __launch_bounds__(kThreads) dielectric_overlap_kernel(
    const double2* __restrict__ Xk, const double2* __restrict__ Xkq, const double* __restrict__ Ek,
    const double* __restrict__ Ekq, const double* __restrict__ kweight, int nb, int dim, int64_t nk,
    double* __restrict__ weight, double* __restrict__ delta) {
    const int lane = threadIdx.x & 63;
    const int64_t k = blockIdx.x * (int64_t)(kThreads / 64) + (threadIdx.x >> 6);
    if (k >= nk) return;  // whole waves leave together; no block barrier follows
    const double2* xk = Xk + k * nb * (int64_t)dim;
    const double2* xq = Xkq + k * nb * (int64_t)dim;
    const int npair = (nb - kValence) * kValence;
    for (int c = kValence; c < nb; ++c) {
        double re[kValence], im[kValence];
#pragma unroll
        for (int v = 0; v < kValence; ++v) re[v] = im[v] = 0.0;
        for (int i = lane; i < dim; i += 64) {
            const double2 a = xq[c * (int64_t)dim + i];
#pragma unroll
            for (int v = 0; v < kValence; ++v) {
                const double2 b = xk[v * (int64_t)dim + i];
                re[v] += a.x * b.x - a.y * b.y;
                im[v] += a.x * b.y + a.y * b.x;
            }
        }
#pragma unroll
        for (int v = 0; v < kValence; ++v) {
            const double sr = wave_sum(re[v]), si = wave_sum(im[v]);
            if (lane == 0) {
                const double mag = hypot(sr, si);  // std::abs of the complex sum
                const int64_t p = k * npair + (c - kValence) * kValence + v;
                weight[p] = kweight ? (mag * mag) * kweight[k] : mag * mag;
                delta[p] = Ekq[k * nb + c] - Ek[k * nb + v];
            }
        }
    }
}

__global__ void __launch_bounds__(kThreads) dielectric_lorentz_kernel(
    const double* __restrict__ weight, const double* __restrict__ delta, int64_t nitems, int64_t per_slice,
    const double* __restrict__ energies, int n_energies, double eta, double* __restrict__ partial) {
    const int g0 = blockIdx.x * kEnergiesPerGroup;
    double E[kEnergiesPerGroup], acc[kEnergiesPerGroup];
#pragma unroll
    for (int j = 0; j < kEnergiesPerGroup; ++j) {
        E[j] = g0 + j < n_energies ? energies[g0 + j] : 0.0;
        acc[j] = 0.0;
    }
    const double eta2 = eta * eta;
    const int64_t t_begin = blockIdx.y * per_slice;
    const int64_t t_end = min(nitems, t_begin + per_slice);
    for (int64_t t = t_begin + threadIdx.x; t < t_end; t += kThreads) {
        const double w = weight[t], d = delta[t];
#pragma unroll
        for (int j = 0; j < kEnergiesPerGroup; ++j) {
            const double f1 = (d - E[j]) / ((d - E[j]) * (d - E[j]) + eta2);
            const double f2 = (d + E[j]) / ((d + E[j]) * (d + E[j]) + eta2);
            acc[j] += w * (f1 + f2);
        }
    }
    __shared__ double part[kThreads / 64][kEnergiesPerGroup];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kEnergiesPerGroup; ++j) {
        const double s = wave_sum(acc[j]);
        if (lane == 0) part[wave][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < kEnergiesPerGroup && g0 + (int)threadIdx.x < n_energies) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += part[w][threadIdx.x];
        partial[blockIdx.y * (int64_t)n_energies + g0 + threadIdx.x] = s;
    }
}

__global__ void dielectric_finish_kernel(const double* __restrict__ partial, int slices, int n_energies,
                                         double* __restrict__ sum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_energies) return;
    double s = 0.0;
    for (int k = 0; k < slices; ++k) s += partial[(int64_t)k * n_energies + i];
    sum[i] = sum[i] + s;
}

constexpr int64_t kItemsPerSlice = 4096;

}  // namespace

namespace uepm {

int dielectric_slices(int64_t nitems) {
    int64_t s = (nitems + kItemsPerSlice - 1) / kItemsPerSlice;
    return static_cast<int>(s < 1 ? 1 : s);
}

hipError_t launch_dielectric_accumulate(const double2* Xk, const double2* Xkq, const double* Ek, const double* Ekq,
                                        const double* kweight, int nb, int dim, int64_t nk, const double* energies,
                                        int n_energies,
                                        double eta, double* weight, double* delta, double* partial, double* sum,
                                        hipStream_t s) {
    if (nk <= 0) return hipSuccess;
    const unsigned ob = static_cast<unsigned>((nk + kThreads / 64 - 1) / (kThreads / 64));
    hipLaunchKernelGGL(dielectric_overlap_kernel, dim3(ob), dim3(kThreads), 0, s, Xk, Xkq, Ek, Ekq, kweight, nb, dim,
                       nk, weight, delta);
    const int64_t nitems = nk * (nb - kValence) * kValence;
    const int slices = dielectric_slices(nitems);
    const unsigned groups = static_cast<unsigned>((n_energies + kEnergiesPerGroup - 1) / kEnergiesPerGroup);
    hipLaunchKernelGGL(dielectric_lorentz_kernel, dim3(groups, slices), dim3(kThreads), 0, s, weight, delta, nitems,
                       kItemsPerSlice, energies, n_energies, eta, partial);
    const unsigned fb = static_cast<unsigned>((n_energies + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(dielectric_finish_kernel, dim3(fb), dim3(kThreads), 0, s, partial, slices, n_energies, sum);
    return hipGetLastError();
}

}  // namespace uepm
