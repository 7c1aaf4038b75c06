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

// ---------------------------------------------------------------------------
// Blocked back-transform on MFMA (precision form).  The reflectors are applied
// sixteen at a time as Q_b = H_j0 H_j0+1 ... H_j0+15 = I - V T V^H (compact WY),
// the last block first: X <- X - V (T (V^H X)).  T is never formed: with
// G = V^H V, T^-1 = U, the upper triangle with U_jj = 1 / tau_j and U_ij = G_ij
// above the diagonal (the LAPACK xLARFT forward/columnwise recurrence
// T_{0:j,j} = -tau_j T_{0:j,0:j} G_{0:j,j}, T_jj = tau_j, inverted), so
// W' = T W is a back substitution with U.  One 256-lane workgroup per
// k-point; X (n rows x 16 NT bands, complex64: bands b0 .. b0 + nbp - 1 of the
// nb per k-point, the rest of the tile zero) lives in LDS.
//   1. [G | W] = V^H [V | X] on v_mfma_f32_16x16x4f32 (complex products as
//      four real ones), the rows split over the four waves; the per-wave
//      partial sums are added in a fixed order, so results do not depend on
//      scheduling;
//   2. U W' = W by back substitution, one band per lane (a reflector with
//      tau = 0 is the identity: its row of W' is zero);
//   3. X -= V W' on MFMA, 16-row tiles over the waves.
// Each reflector's column is read twice per block (steps 1 and 4) instead of
// once per band group, and the arithmetic runs on the matrix cores.
template <int NT>
__global__ __launch_bounds__(256, 2) void apply_q_wy_kernel(const float2* __restrict__ V, int n, int ld,
                                                         const float2* __restrict__ tauall,
                                                         const double* __restrict__ Z, int nb, int b0, int nbp,
                                                         double2* __restrict__ X) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int NBT = 16 * NT;
    constexpr int XS = NBT + 1;  // padded row stride of X in LDS (complex elements)
    extern __shared__ __align__(16) unsigned char wy_smem[];
    float2* Xs = reinterpret_cast<float2*>(wy_smem);  // X[row][band] at row * XS + band
    __shared__ f4 PG[4][2][64];      // per-wave partial G (re, im) in the MFMA output layout
    __shared__ f4 PW[4][NT][2][64];  // per-wave partial W
    __shared__ float2 Gs[16][17];    // G = V^H V
    __shared__ float2 Ws[16][XS];    // W = V^H X
    __shared__ float2 Wq[16][XS];    // W' = T W
    __shared__ float2 tb[16];        // tau of the block's reflectors

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int64_t kk = blockIdx.x;
    const float2* Vk = V + kk * (int64_t)ld * n;
    const float2* tau = tauall + kk * n;
    auto mfma = [](float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); };
    // element `row` of v_j: 1 at row j + 1, the stored reflector below, zero above
    // and for j past the last reflector (n - 2)
    auto vload = [&](int row, int j) -> float2 {
        if (j > n - 2 || row >= n || row <= j) return make_float2(0.f, 0.f);
        if (row == j + 1) return make_float2(1.f, 0.f);
        return Vk[row + (int64_t)j * ld];
    };

    // X = Z (real) for bands b0 .. b0 + nbp - 1 of the nb per k-point; the rest of the tile is zero
    for (int e = tid; e < n * NBT; e += 256) {
        const int b = e / n, row = e - b * n;
        const float z = b < nbp ? (float)Z[(kk * nb + b0 + b) * (int64_t)n + row] : 0.f;
        Xs[row * XS + b] = make_float2(z, 0.f);
    }
    __syncthreads();

    const int nblk = (n - 1 + 15) >> 4;  // reflectors 0 .. n - 2
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = 16 * blk, r0 = j0 + 1;
        // ---- 1. partial [G | W] over this wave's 4-row chunks ----
        const f4 z4 = {0.f, 0.f, 0.f, 0.f};
        f4 gre = z4, gim = z4, wre[NT], wim[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) wre[t] = wim[t] = z4;
        const int nch = (n - r0 + 3) >> 2;
        // four chunks per pass, their loads issued before the MFMAs that use them
        for (int ch0 = wv; ch0 < nch; ch0 += 16) {
            float2 vv[4], xx[4][NT];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int row = r0 + 4 * (ch0 + 4 * u) + l4;
                // A[c][k] = conj(v_{j0+c}[row]) and B[k][c'] = v_{j0+c'}[row]: the same element per lane
                vv[u] = vload(row, j0 + l16);
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    xx[u][t] = row < n ? Xs[row * XS + 16 * t + l16] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float2 v = vv[u];
                gre = mfma(v.x, v.x, gre);
                gre = mfma(v.y, v.y, gre);
                gim = mfma(v.x, v.y, gim);
                gim = mfma(-v.y, v.x, gim);
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const float2 x = xx[u][t];
                    wre[t] = mfma(v.x, x.x, wre[t]);
                    wre[t] = mfma(v.y, x.y, wre[t]);
                    wim[t] = mfma(v.x, x.y, wim[t]);
                    wim[t] = mfma(-v.y, x.x, wim[t]);
                }
            }
        }
        PG[wv][0][lane] = gre;
        PG[wv][1][lane] = gim;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            PW[wv][t][0][lane] = wre[t];
            PW[wv][t][1][lane] = wim[t];
        }
        if (tid < 16) tb[tid] = j0 + tid <= n - 2 ? tau[j0 + tid] : make_float2(0.f, 0.f);
        __syncthreads();
        // sum the four waves' partials in a fixed order; output item q of lane L
        // is element (4 (L / 16) + q, L % 16)
        for (int fl = tid; fl < (2 + 2 * NT) * 64; fl += 256) {
            const int f = fl >> 6, L = fl & 63;
            f4 sum;
            if (f < 2) {
                sum = PG[0][f][L];
                sum += PG[1][f][L];
                sum += PG[2][f][L];
                sum += PG[3][f][L];
            } else {
                const int t = (f - 2) >> 1, part = (f - 2) & 1;
                sum = PW[0][t][part][L];
                sum += PW[1][t][part][L];
                sum += PW[2][t][part][L];
                sum += PW[3][t][part][L];
            }
            const int c = L & 15;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = 4 * (L >> 4) + q;
                if (f < 2) {
                    float* g = reinterpret_cast<float*>(&Gs[m][c]);
                    g[f] = sum[q];
                } else {
                    const int t = (f - 2) >> 1, part = (f - 2) & 1;
                    float* w = reinterpret_cast<float*>(&Ws[m][16 * t + c]);
                    w[part] = sum[q];
                }
            }
        }
        __syncthreads();
        // ---- 2. W' = T W: back substitution U W' = W, one band per lane ----
        for (int b = tid; b < NBT; b += 256) {
            float2 wq[16];
#pragma unroll
            for (int i = 15; i >= 0; --i) {
                float sr = Ws[i][b].x, si = Ws[i][b].y;
#pragma unroll
                for (int p = i + 1; p < 16; ++p) {
                    const float2 u = Gs[i][p];
                    sr = fmaf(-u.x, wq[p].x, fmaf(u.y, wq[p].y, sr));
                    si = fmaf(-u.x, wq[p].y, fmaf(-u.y, wq[p].x, si));
                }
                const float2 t = tb[i];  // 1 / U_ii
                wq[i] = make_float2(t.x * sr - t.y * si, t.x * si + t.y * sr);
                Wq[i][b] = wq[i];
            }
        }
        __syncthreads();
        // ---- 3. X -= V W' on the rows below j0, 16-row tiles over the waves ----
        const int ntile = (n - r0 + 15) >> 4;
        for (int tt = wv; tt < ntile; tt += 4) {
            const int R0 = r0 + 16 * tt;
            float2 a[4];  // A[m][k] = v_{j0+k}[R0 + m]: lane (m = l16, k = 4 kc + l4)
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) a[kc] = vload(R0 + l16, j0 + 4 * kc + l4);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                f4 dre, dim;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = R0 + 4 * l4 + q;
                    const float2 x = row < n ? Xs[row * XS + 16 * t + l16] : make_float2(0.f, 0.f);
                    dre[q] = x.x;
                    dim[q] = x.y;
                }
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const float2 b = Wq[4 * kc + l4][16 * t + l16];
                    dre = mfma(-a[kc].x, b.x, dre);
                    dre = mfma(a[kc].y, b.y, dre);
                    dim = mfma(-a[kc].x, b.y, dim);
                    dim = mfma(-a[kc].y, b.x, dim);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = R0 + 4 * l4 + q;
                    if (row < n) Xs[row * XS + 16 * t + l16] = make_float2(dre[q], dim[q]);
                }
            }
        }
        __syncthreads();
    }
    // X0[k][band][row] in FP64
    for (int e = tid; e < n * nbp; e += 256) {
        const int b = e / n, row = e - b * n;
        const float2 x = Xs[row * XS + b];
        X[(kk * nb + b0 + b) * (int64_t)n + row] = make_double2((double)x.x, (double)x.y);
    }
}

// The FP64 form's blocked back-transform on v_mfma_f64_16x16x4f64: the three
// steps of apply_q_wy_kernel in complex128, sixteen bands per pass (bands
// b0 .. b0 + nbp - 1 of the nb per k-point), X in LDS.  The FP64 instruction
// puts output item v of lane l at element (4 v + l / 16, l % 16) (the FP32
// one at (4 (l / 16) + v, l % 16)).  The four waves' partial [G | W] are added
// in two levels, (w0 + w2) + (w1 + w3), a fixed order in half the LDS of a
// four-way table.  X's row stride XS is 17 complex elements (conflict-free
// column walks of the load, store and gradient loops) where that fits the LDS,
// else 16 (those one-off loops then conflict; the per-block MFMA loops walk
// rows and do not).  As the FP64 form's last stage it also counts the k-points
// that kstat flags into status (first pass only) and writes the gradients
// (the reflector kernel's sums, in its order).
template <int XS>
__global__ __launch_bounds__(256) void apply_q_wy_f64_kernel(
    const double2* __restrict__ V, int n, int ld, const double2* __restrict__ tauall, const double* __restrict__ Z,
    int nb, int b0, int nbp, double2* __restrict__ X, double* __restrict__ grads, DevParams P,
    const double* __restrict__ kpts, const int32_t* __restrict__ G, const int* __restrict__ kstat,
    unsigned long long* __restrict__ status, int64_t k0) {
    typedef double d4 __attribute__((ext_vector_type(4)));
    extern __shared__ __align__(16) unsigned char wy64_smem[];
    double2* Xs = reinterpret_cast<double2*>(wy64_smem);  // X[row][band] at row * XS + band
    __shared__ d4 PP[2][4][64];  // partials (G re, G im, W re, W im): waves 2, 3, then w0 + w2, w1 + w3
    __shared__ double2 Gs[16][17];  // G = V^H V
    __shared__ double2 Ws[16][17];  // W = V^H X
    __shared__ double2 Wq[16][17];  // W' = T W
    __shared__ double2 tb[16];      // tau of the block's reflectors

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int64_t kk = blockIdx.x;
    if (status && b0 == 0 && tid == 0 && kstat[kk]) {
        atomicAdd(&status[0], 1ull);
        atomicMin(&status[1], (unsigned long long)(k0 + kk));
    }
    const double2* Vk = V + kk * (int64_t)ld * n;
    const double2* tau = tauall + kk * n;
    auto mfma = [](double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); };
    // element `row` of v_j: 1 at row j + 1, the stored reflector below, zero above
    // and for j past the last reflector (n - 2)
    auto vload = [&](int row, int j) -> double2 {
        if (j > n - 2 || row >= n || row <= j) return make_double2(0.0, 0.0);
        if (row == j + 1) return make_double2(1.0, 0.0);
        return Vk[row + (int64_t)j * ld];
    };

    for (int e = tid; e < n * 16; e += 256) {
        const int b = e / n, row = e - b * n;
        const double z = b < nbp ? Z[(kk * nb + b0 + b) * (int64_t)n + row] : 0.0;
        Xs[row * XS + b] = make_double2(z, 0.0);
    }
    __syncthreads();

    const int nblk = (n - 1 + 15) >> 4;  // reflectors 0 .. n - 2
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = 16 * blk, r0 = j0 + 1;
        // ---- 1. partial [G | W] over this wave's 4-row chunks ----
        const d4 z4 = {0.0, 0.0, 0.0, 0.0};
        d4 gre = z4, gim = z4, wre = z4, wim = z4;
        const int nch = (n - r0 + 3) >> 2;
        for (int ch0 = wv; ch0 < nch; ch0 += 16) {
            double2 vv[4], xx[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int row = r0 + 4 * (ch0 + 4 * u) + l4;
                vv[u] = vload(row, j0 + l16);
                xx[u] = row < n ? Xs[row * XS + l16] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double2 v = vv[u], x = xx[u];
                gre = mfma(v.x, v.x, gre);
                gre = mfma(v.y, v.y, gre);
                gim = mfma(v.x, v.y, gim);
                gim = mfma(-v.y, v.x, gim);
                wre = mfma(v.x, x.x, wre);
                wre = mfma(v.y, x.y, wre);
                wim = mfma(v.x, x.y, wim);
                wim = mfma(-v.y, x.x, wim);
            }
        }
        if (wv >= 2) {
            PP[wv - 2][0][lane] = gre;
            PP[wv - 2][1][lane] = gim;
            PP[wv - 2][2][lane] = wre;
            PP[wv - 2][3][lane] = wim;
        }
        if (tid < 16) tb[tid] = j0 + tid <= n - 2 ? tau[j0 + tid] : make_double2(0.0, 0.0);
        __syncthreads();
        if (wv < 2) {  // each lane reads, then overwrites, its own slots
            PP[wv][0][lane] = gre + PP[wv][0][lane];
            PP[wv][1][lane] = gim + PP[wv][1][lane];
            PP[wv][2][lane] = wre + PP[wv][2][lane];
            PP[wv][3][lane] = wim + PP[wv][3][lane];
        }
        __syncthreads();
        {
            const int f = tid >> 6, L = tid & 63, c = L & 15;
            const d4 sum = PP[0][f][L] + PP[1][f][L];
            double2* dst = f < 2 ? &Gs[0][0] : &Ws[0][0];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = 4 * q + (L >> 4);
                reinterpret_cast<double*>(&dst[m * 17 + c])[f & 1] = sum[q];
            }
        }
        __syncthreads();
        // ---- 2. W' = T W: back substitution U W' = W, one band per lane ----
        if (tid < 16) {
            const int b = tid;
            double2 wq[16];
#pragma unroll
            for (int i = 15; i >= 0; --i) {
                double sr = Ws[i][b].x, si = Ws[i][b].y;
#pragma unroll
                for (int p = i + 1; p < 16; ++p) {
                    const double2 u = Gs[i][p];
                    sr = fma(-u.x, wq[p].x, fma(u.y, wq[p].y, sr));
                    si = fma(-u.x, wq[p].y, fma(-u.y, wq[p].x, si));
                }
                const double2 t = tb[i];  // 1 / U_ii
                wq[i] = make_double2(t.x * sr - t.y * si, t.x * si + t.y * sr);
                Wq[i][b] = wq[i];
            }
        }
        __syncthreads();
        // ---- 3. X -= V W' on the rows below j0, 16-row tiles over the waves ----
        const int ntile = (n - r0 + 15) >> 4;
        for (int tt = wv; tt < ntile; tt += 4) {
            const int R0 = r0 + 16 * tt;
            double2 a[4];  // A[m][k] = v_{j0+k}[R0 + m]: lane (m = l16, k = 4 kc + l4)
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) a[kc] = vload(R0 + l16, j0 + 4 * kc + l4);
            d4 dre, dim;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = R0 + 4 * q + l4;
                const double2 x = row < n ? Xs[row * XS + l16] : make_double2(0.0, 0.0);
                dre[q] = x.x;
                dim[q] = x.y;
            }
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                const double2 b = Wq[4 * kc + l4][l16];
                dre = mfma(-a[kc].x, b.x, dre);
                dre = mfma(a[kc].y, b.y, dre);
                dim = mfma(-a[kc].x, b.y, dim);
                dim = mfma(-a[kc].y, b.x, dim);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = R0 + 4 * q + l4;
                if (row < n) Xs[row * XS + l16] = make_double2(dre[q], dim[q]);
            }
        }
        __syncthreads();
    }
    // eigenvectors X[k][band][row]; a k-point the eigen-stage flagged gets NaN
    // throughout (its z is NaN, but row 0, which no reflector touches, would
    // keep a zero imaginary part)
    const bool failed = kstat && kstat[kk];
    for (int e = tid; e < n * nbp; e += 256) {
        const int b = e / n, row = e - b * n;
        const double nan = __builtin_nan("");
        X[(kk * nb + b0 + b) * (int64_t)n + row] = failed ? make_double2(nan, nan) : Xs[row * XS + b];
    }
    if (grads) {
        // Hellmann-Feynman on the kinetic term, 16 lanes per band (rows rs + 16 q
        // summed in ascending q, then across the 16 lanes); spinor rows i >= N belong to G_{i-N}
        const int b = tid >> 4, rs = tid & 15;
        const double* k = kpts + 3 * kk;
        const int N = P.N;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int i = rs; i < n; i += 16) {
            const double2 x = Xs[i * XS + b];
            const int g = i < N ? i : i - N;
            const double w = x.x * x.x + x.y * x.y;
            sx += (k[0] + (double)G[3 * g]) * w;
            sy += (k[1] + (double)G[3 * g + 1]) * w;
            sz += (k[2] + (double)G[3 * g + 2]) * w;
        }
        sx = row_sum16(sx);
        sy = row_sum16(sy);
        sz = row_sum16(sz);
        if (rs == 0 && b < nbp) {
            double* gg = grads + 3 * (kk * nb + b0 + b);
            gg[0] = sx * P.grad_pref * P.frac_to_phys;
            gg[1] = sy * P.grad_pref * P.frac_to_phys;
            gg[2] = sz * P.grad_pref * P.frac_to_phys;
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

namespace {
// dynamic LDS of apply_q_wy_f64_kernel<XS> (X in LDS) and its static part
size_t wy64_lds_bytes(int n, int xs) { return sizeof(double2) * (size_t)n * xs; }
constexpr size_t kWy64StaticBytes = sizeof(double) * 4 * 2 * 4 * 64 + sizeof(double2) * (3 * 16 * 17 + 16);
// row stride of X: 17 where it fits the LDS, else 16, else 0 (does not fit)
int wy64_stride(int n, int lds_limit) {
    if (n < 2) return 0;
    for (int xs : {17, 16})
        if (wy64_lds_bytes(n, xs) + kWy64StaticBytes <= (size_t)lds_limit) return xs;
    return 0;
}
}  // namespace

bool apply_q_wy_f64_fits(int n, int lds_limit) { return wy64_stride(n, lds_limit) > 0; }

hipError_t launch_apply_q_f64(const double2* dV, int n, int ld, const double2* dtau, const double* dZ, int nb,
                              double2* dX, double* dgrads, const DevParams& P, const double* dk, const int32_t* dG,
                              int64_t nk, const int* kstat, unsigned long long* status, int64_t k0,
                              int bands_per_group, int lds_limit, hipStream_t s) {
    if (bands_per_group == 0) {
        // the blocked MFMA back-transform, sixteen bands per pass
        const int xs = wy64_stride(n, lds_limit);
        if (xs == 0) return hipErrorInvalidValue;
        const size_t shm = wy64_lds_bytes(n, xs);
        auto kern = xs == 17 ? apply_q_wy_f64_kernel<17> : apply_q_wy_f64_kernel<16>;
        hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (err != hipSuccess) return err;
        for (int b0 = 0; b0 < nb; b0 += 16)
            hipLaunchKernelGGL(kern, dim3((unsigned)nk), dim3(256), shm, s, dV, n, ld, dtau, dZ, nb, b0,
                               std::min(16, nb - b0), dX, dgrads, P, dk, dG, kstat, status, k0);
        return hipGetLastError();
    }
    return launch_apply<double, double2>(dV, n, ld, dtau, dZ, nb, dX, dgrads, P, dk, dG, nk, kstat, status, k0,
                                         bands_per_group, s);
}

namespace {
// dynamic LDS of apply_q_wy_kernel<NT> (X in LDS) and its static part
size_t wy_lds_bytes(int n, int nt) { return sizeof(float2) * (size_t)n * (16 * nt + 1); }
size_t wy_static_bytes(int nt) {
    return 16 * 4 * 2 * 64 + 16 * 4 * nt * 2 * 64 + sizeof(float2) * (16 * 17 + 2 * 16 * (16 * nt + 1) + 16);
}
}  // namespace

bool apply_q_wy_fits(int n, int nb, int lds_limit) {
    // 17..32 bands run as one two-tile pass, or as two one-tile passes when the
    // two-tile X does not fit the LDS (each pass reads the reflectors once more)
    if (nb > 32 || n < 2) return false;
    return wy_lds_bytes(n, 1) + wy_static_bytes(1) <= (size_t)lds_limit;
}

hipError_t launch_apply_q_c64(const float2* dV, int n, int ld, const float2* dtau, const double* dZ, int nb,
                              double2* dX, const DevParams& P, const double* dk, const int32_t* dG, int64_t nk,
                              int bands_per_group, int lds_limit, hipStream_t s) {
    if (bands_per_group == 0) {
        // the blocked MFMA back-transform
        if (!apply_q_wy_fits(n, nb, lds_limit)) return hipErrorInvalidValue;
        const bool two = nb > 16 && wy_lds_bytes(n, 2) + wy_static_bytes(2) <= (size_t)lds_limit;
        const int nt = two ? 2 : 1;
        const size_t shm = wy_lds_bytes(n, nt);
        auto kern = nt == 1 ? apply_q_wy_kernel<1> : apply_q_wy_kernel<2>;
        hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (err != hipSuccess) return err;
        for (int b0 = 0; b0 < nb; b0 += 16 * nt)
            hipLaunchKernelGGL(kern, dim3((unsigned)nk), dim3(256), shm, s, dV, n, ld, dtau, dZ, nb, b0,
                               std::min(16 * nt, nb - b0), dX);
        return hipGetLastError();
    }
    return launch_apply<float, float2>(dV, n, ld, dtau, dZ, nb, dX, nullptr, P, dk, dG, nk, nullptr, nullptr, 0,
                                       bands_per_group, s);
}

}  // namespace uepm
