// refine.hip — FP64 Rayleigh-Ritz refinement of approximate eigenpairs (the
// precision form's last stage), one workgroup per k-point, on gfx950.
//
// Input: nbr approximate lowest eigenvectors X0 of H(k) from the complex64
// reduction (accurate to ~1e-6), the FP64 H(k) as assembled.  In FP64:
//   Y = H X0,  S = X0^H X0,  M = X0^H Y,  S = L L^H (Cholesky),
//   M' = L^{-1} M L^{-H} = Z Theta Z^H (cyclic Jacobi),  X1 = X0 L^{-H} Z.
// Theta are the Ritz values of the FP64 H on span(X0): the error of each
// is second order in the eigenvector error (~1e-11 eV for isolated bands) and
// the Ritz step resolves mixing inside near-degenerate clusters, which a
// per-vector Rayleigh quotient would not.  The lowest nb of the nbr Ritz
// pairs are written (eigenvalues ascending, unit eigenvectors, gradients).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "epm_physics.hpp"
#include "solver.hpp"

namespace uepm {

namespace {

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// Orders LDS traffic between lanes of one wave (compiler and memory).
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

constexpr int kRefBS = 256;
constexpr int kNbrMax = 28;
constexpr int kChunk = 64;

}  // namespace

template <int BS>
__global__ __launch_bounds__(BS) void refine_kernel(const double2* __restrict__ H, int n, int ld,
                                                    double2* __restrict__ Xall, double2* __restrict__ Yall, int nbr,
                                                    int nb, double* __restrict__ evals, double2* __restrict__ evecs,
                                                    double* __restrict__ grads, DevParams P,
                                                    const double* __restrict__ kpts, const int32_t* __restrict__ G) {
    __shared__ double2 Xc[kNbrMax * kChunk];        // chunk of X0 rows for Y = H X0
    __shared__ double2 Mm[kNbrMax * kNbrMax];       // M, then M'
    __shared__ double2 Ss[kNbrMax * kNbrMax];       // S, then L
    __shared__ double2 Zm[kNbrMax * kNbrMax];       // Jacobi vectors, then C = L^{-H} Z
    __shared__ double2 Wt[kNbrMax * kNbrMax];
    __shared__ double th[kNbrMax];
    __shared__ int perm[kNbrMax];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t kk = blockIdx.x;
    const double2* A = H + kk * (int64_t)ld * n;
    double2* X = Xall + kk * (int64_t)nbr * n;  // X[b * n + r]
    double2* Y = Yall + kk * (int64_t)nbr * n;

    // ---- Y = H X0 (H column-major full, Hermitian)
    for (int r0 = 0; r0 < n; r0 += BS) {
        const int r = r0 + tid;
        double2 acc[kNbrMax];
#pragma unroll
        for (int b = 0; b < kNbrMax; ++b) acc[b] = make_double2(0.0, 0.0);
        for (int c0 = 0; c0 < n; c0 += kChunk) {
            const int cn = min(kChunk, n - c0);
            __syncthreads();
            for (int i = tid; i < nbr * kChunk; i += BS) {
                const int b = i / kChunk, c = i % kChunk;
                Xc[b * kChunk + c] = c < cn ? X[b * n + c0 + c] : make_double2(0.0, 0.0);
            }
            __syncthreads();
            if (r < n) {
                for (int c = 0; c < cn; ++c) {
                    const double2 h = A[r + (int64_t)(c0 + c) * ld];
#pragma unroll
                    for (int b = 0; b < kNbrMax; ++b)
                        if (b < nbr) {
                            const double2 x = Xc[b * kChunk + c];
                            acc[b].x = fma(h.x, x.x, fma(-h.y, x.y, acc[b].x));
                            acc[b].y = fma(h.x, x.y, fma(h.y, x.x, acc[b].y));
                        }
                }
            }
        }
        if (r < n)
#pragma unroll
            for (int b = 0; b < kNbrMax; ++b)
                if (b < nbr) Y[b * n + r] = acc[b];
    }
    __syncthreads();
    __threadfence_block();
    // ---- S = X0^H X0, M = X0^H Y (one wave per (a, b) pair, lanes over rows)
    for (int pr = wv; pr < nbr * nbr; pr += BS / 64) {
        const int a = pr / nbr, b = pr % nbr;
        double sx = 0, sy = 0, mx = 0, my = 0;
        for (int r = lane; r < n; r += 64) {
            const double2 xa = X[a * n + r];
            const double2 s = zmulc(xa, X[b * n + r]);
            const double2 m = zmulc(xa, Y[b * n + r]);
            sx += s.x;
            sy += s.y;
            mx += m.x;
            my += m.y;
        }
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
            mx += __shfl_xor(mx, o, 64);
            my += __shfl_xor(my, o, 64);
        }
        if (lane == 0) {
            Ss[a * kNbrMax + b] = make_double2(sx, sy);
            Mm[a * kNbrMax + b] = make_double2(mx, my);
        }
    }
    __syncthreads();
    // ---- small dense problem in wave 0
    if (wv == 0) {
        const int i = lane;
        // Hermitian parts: M <- (M + M^H)/2 (M is Hermitian up to rounding)
        // Cholesky S = L L^H, L lower, in Ss
        for (int j = 0; j < nbr; ++j) {
            double djj = Ss[j * kNbrMax + j].x;
            for (int k = 0; k < j; ++k) {
                const double2 l = Ss[j * kNbrMax + k];
                djj -= l.x * l.x + l.y * l.y;
            }
            const double ljj = sqrt(djj);
            if (i > j && i < nbr) {
                // L[i][j] = (S[i][j] - sum_k L[i][k] conj(L[j][k])) / L[j][j]
                double2 s = Ss[i * kNbrMax + j];
                for (int k = 0; k < j; ++k) s = zsub(s, zmul(Ss[i * kNbrMax + k], make_double2(Ss[j * kNbrMax + k].x, -Ss[j * kNbrMax + k].y)));
                Ss[i * kNbrMax + j] = zscale(s, 1.0 / ljj);
            }
            if (i == 0) Ss[j * kNbrMax + j] = make_double2(ljj, 0.0);
            wsync();
        }
        // Wt = L^{-1} M  (lane = column of M)
        if (i < nbr) {
            for (int r = 0; r < nbr; ++r) {
                double2 s = Mm[r * kNbrMax + i];
                for (int k = 0; k < r; ++k) s = zsub(s, zmul(Ss[r * kNbrMax + k], Wt[k * kNbrMax + i]));
                Wt[r * kNbrMax + i] = zscale(s, 1.0 / Ss[r * kNbrMax + r].x);
            }
        }
        wsync();
        // M' = (L^{-1} Wt^H)^H  (lane = column of Wt^H, i.e. row i of Wt); symmetrised
        if (i < nbr) {
            for (int r = 0; r < nbr; ++r) {
                const double2 w = Wt[i * kNbrMax + r];
                double2 s = make_double2(w.x, -w.y);
                for (int k = 0; k < r; ++k) s = zsub(s, zmul(Ss[r * kNbrMax + k], Zm[k * kNbrMax + i]));
                Zm[r * kNbrMax + i] = zscale(s, 1.0 / Ss[r * kNbrMax + r].x);  // (L^{-1} Wt^H)[r][i]
            }
        }
        wsync();
        if (i < nbr)
            for (int c = 0; c < nbr; ++c) {
                const double2 u = Zm[c * kNbrMax + i];  // M'[i][c] = conj(U[c][i])
                Mm[i * kNbrMax + c] = make_double2(u.x, -u.y);
            }
        wsync();
        if (i < nbr)
            for (int c = 0; c < nbr; ++c) Zm[i * kNbrMax + c] = make_double2(i == c ? 1.0 : 0.0, 0.0);
        if (i < nbr) {
            for (int c = i + 1; c < nbr; ++c) {  // symmetrise
                const double2 a = Mm[i * kNbrMax + c], b = Mm[c * kNbrMax + i];
                const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
                Mm[i * kNbrMax + c] = h;
                Mm[c * kNbrMax + i] = make_double2(h.x, -h.y);
            }
            Mm[i * kNbrMax + i].y = 0.0;
        }
        wsync();
        // cyclic complex Jacobi: rotate while the off-diagonal element moves an
        // eigenvalue by more than 1e-14 relative (second order |c|^2 / gap).
        for (int sweep = 0; sweep < 12; ++sweep) {
            int rotated = 0;
            for (int p = 0; p < nbr - 1; ++p)
                for (int q = p + 1; q < nbr; ++q) {
                    const double2 c = Mm[p * kNbrMax + q];
                    const double ac = sqrt(c.x * c.x + c.y * c.y);
                    const double a = Mm[p * kNbrMax + p].x, b = Mm[q * kNbrMax + q].x;
                    const double scale = fabs(a) + fabs(b) + 1e-300;
                    if (ac <= 1e-15 * scale) continue;
                    rotated = 1;
                    // e^{-i phi} with c = |c| e^{i phi}
                    const double2 em = make_double2(c.x / ac, -c.y / ac);
                    const double tau_ = (b - a) / (2.0 * ac);
                    const double t = (tau_ >= 0 ? 1.0 : -1.0) / (fabs(tau_) + sqrt(1.0 + tau_ * tau_));
                    const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                    // G = [[cs, sn], [-sn e^{-i phi}, cs e^{-i phi}]]; M <- G^H M G, Z <- Z G
                    if (i < nbr) {
                        const double2 mp = Mm[i * kNbrMax + p], mq = Mm[i * kNbrMax + q];
                        const double2 emq = zmul(em, mq);
                        Mm[i * kNbrMax + p] = zsub(zscale(mp, cs), zscale(emq, sn));
                        Mm[i * kNbrMax + q] = zadd(zscale(mp, sn), zscale(emq, cs));
                        const double2 zp = Zm[i * kNbrMax + p], zq = Zm[i * kNbrMax + q];
                        const double2 ezq = zmul(em, zq);
                        Zm[i * kNbrMax + p] = zsub(zscale(zp, cs), zscale(ezq, sn));
                        Zm[i * kNbrMax + q] = zadd(zscale(zp, sn), zscale(ezq, cs));
                    }
                    wsync();
                    if (i < nbr) {
                        const double2 rp = Mm[p * kNbrMax + i], rq = Mm[q * kNbrMax + i];
                        const double2 erq = zmul(make_double2(em.x, -em.y), rq);  // conj(e^{-i phi}) rq
                        Mm[p * kNbrMax + i] = zsub(zscale(rp, cs), zscale(erq, sn));
                        Mm[q * kNbrMax + i] = zadd(zscale(rp, sn), zscale(erq, cs));
                    }
                    wsync();
                    if (i == 0) {
                        Mm[p * kNbrMax + q] = make_double2(0.0, 0.0);
                        Mm[q * kNbrMax + p] = make_double2(0.0, 0.0);
                        Mm[p * kNbrMax + p].y = 0.0;
                        Mm[q * kNbrMax + q].y = 0.0;
                    }
                    wsync();
                }
            if (!rotated) break;
        }
        wsync();
        // C = L^{-H} Z (back substitution with L^H, lane = column of Z)
        if (i < nbr) {
            for (int r = nbr - 1; r >= 0; --r) {
                double2 s = Zm[r * kNbrMax + i];
                for (int k = r + 1; k < nbr; ++k) {
                    const double2 l = Ss[k * kNbrMax + r];  // (L^H)[r][k] = conj(L[k][r])
                    s = zsub(s, zmulc(l, Wt[k * kNbrMax + i]));
                }
                Wt[r * kNbrMax + i] = zscale(s, 1.0 / Ss[r * kNbrMax + r].x);
            }
        }
        if (i < nbr) th[i] = Mm[i * kNbrMax + i].x;
        wsync();
        if (i == 0) {
            for (int b = 0; b < nbr; ++b) perm[b] = b;
            for (int b = 0; b < nbr; ++b) {  // selection sort of the Ritz values
                int best = b;
                for (int c = b + 1; c < nbr; ++c)
                    if (th[perm[c]] < th[perm[best]]) best = c;
                const int tmp = perm[b];
                perm[b] = perm[best];
                perm[best] = tmp;
            }
        }
    }
    __syncthreads();
    // ---- X1 = X0 C[:, perm[b]] for the lowest nb; outputs
    if (tid < nb) evals[kk * nb + tid] = th[perm[tid]];
    double2* out = evecs ? evecs + kk * (int64_t)nb * n : nullptr;
    for (int b = wv; b < nb; b += BS / 64) {
        const int cb = perm[b];
        double gx = 0, gy = 0, gz = 0;
        const double* k = kpts + 3 * kk;
        for (int r = lane; r < n; r += 64) {
            double2 acc = make_double2(0.0, 0.0);
            for (int a = 0; a < nbr; ++a) {
                const double2 x = X[a * n + r];
                const double2 c = Wt[a * kNbrMax + cb];
                acc.x = fma(x.x, c.x, fma(-x.y, c.y, acc.x));
                acc.y = fma(x.x, c.y, fma(x.y, c.x, acc.y));
            }
            if (out) out[(int64_t)b * n + r] = acc;
            if (grads) {
                // Hellmann-Feynman, kinetic term (Hamiltonian.cpp:146-194); SOC: spin-down row r
                // of basis vector r - N adds its weight to the same k + G.
                const int i = r < P.N ? r : r - P.N;
                const double wt = acc.x * acc.x + acc.y * acc.y;
                gx += (k[0] + (double)G[3 * i]) * wt;
                gy += (k[1] + (double)G[3 * i + 1]) * wt;
                gz += (k[2] + (double)G[3 * i + 2]) * wt;
            }
        }
        if (grads) {
            for (int o = 32; o > 0; o >>= 1) {
                gx += __shfl_xor(gx, o, 64);
                gy += __shfl_xor(gy, o, 64);
                gz += __shfl_xor(gz, o, 64);
            }
            if (lane == 0) {
                double* g = grads + 3 * (kk * nb + b);
                g[0] = gx * P.grad_pref * P.frac_to_phys;
                g[1] = gy * P.grad_pref * P.frac_to_phys;
                g[2] = gz * P.grad_pref * P.frac_to_phys;
            }
        }
    }
}

hipError_t launch_refine(const double2* dH, int n, int ld, double2* dX, int nbr, int nb, double* devals,
                         double2* devecs, double* dgrads, const DevParams& P, const double* dk, const int32_t* dG,
                         int64_t nk, hipStream_t s, double2* dY) {
    if (nbr > kNbrMax || nb > nbr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(refine_kernel<kRefBS>, dim3((unsigned)nk), dim3(kRefBS), 0, s, dH, n, ld, dX, dY, nbr, nb,
                       devals, devecs, dgrads, P, dk, dG);
    return hipGetLastError();
}

}  // namespace uepm
