// refine.hip — FP64 Rayleigh-Ritz refinement of approximate eigenpairs (the
// precision form's last stage), one workgroup per k-point, on gfx950.
//
// Input: nbr approximate lowest eigenvectors X0 of H(k) from the complex64
// reduction (accurate to ~1e-6), the FP64 H(k) as assembled.  In FP64:
//   Y = H X0,  S = X0^H X0,  M = X0^H Y,  S = L L^H (Cholesky),
//   M' = L^{-1} M L^{-H} = Z Theta Z^H (Jacobi),  X1 = X0 L^{-H} Z.
// For nbr <= 16, Y, S and M run on FP64 MFMA without storing Y.
// Theta are the Ritz values of the FP64 H on span(X0): the error of each
// is second order in the eigenvector error (~1e-11 eV for isolated bands) and
// the Ritz step resolves mixing inside near-degenerate clusters, which a
// per-vector Rayleigh quotient would not.  The lowest nb of the nbr Ritz
// pairs are written (eigenvalues ascending, unit eigenvectors, gradients).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstdlib>

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
// column quads loaded ahead of their MFMAs in Y = H X (one / two band tiles)
constexpr int kUnroll = 8;
constexpr int kUnroll2 = 4;

}  // namespace

// The small Hermitian problem on nbr <= LDM Ritz vectors, in one wave:
// Cholesky S = L L^H, M' = L^{-1} M L^{-H}, parallel-ordered Jacobi M' = Z Theta Z^H,
// C = L^{-H} Z (in Wt), Ritz values th and the ascending order perm.
// Inputs Ss = S, Mm = M (row-major, stride LDM); Ss, Mm, Zm, Wt are overwritten.
// Triangular solves on one column held in registers (v[r], r < nbr), with L
// lower triangular in Ss (row-major, stride LDM, real diagonal); each row
// subtracts its terms in ascending k, as a column-by-column substitution does.
template <int LDM>
__device__ __forceinline__ void lower_solve(int nbr, const double2* __restrict__ L, double2 (&v)[LDM]) {
#pragma unroll
    for (int r = 0; r < LDM; ++r)
        if (r < nbr) {
            double2 s = v[r];
#pragma unroll
            for (int k = 0; k < r; ++k) s = zsub(s, zmul(L[r * LDM + k], v[k]));
            v[r] = zscale(s, 1.0 / L[r * LDM + r].x);
        }
}
// L^H v' = v (back substitution)
template <int LDM>
__device__ __forceinline__ void upper_solve_h(int nbr, const double2* __restrict__ L, double2 (&v)[LDM]) {
#pragma unroll
    for (int r = LDM - 1; r >= 0; --r)
        if (r < nbr) {
            double2 s = v[r];
#pragma unroll
            for (int k = r + 1; k < LDM; ++k)
                if (k < nbr) s = zsub(s, zmulc(L[k * LDM + r], v[k]));  // (L^H)[r][k] = conj(L[k][r])
            v[r] = zscale(s, 1.0 / L[r * LDM + r].x);
        }
}

// The small Hermitian problem on nbr <= LDM Ritz vectors, in one wave:
// Cholesky S = L L^H, M' = L^{-1} M L^{-H}, parallel-ordered Jacobi M' = Z Theta Z^H,
// C = L^{-H} Z (in Wt), Ritz values th and the ascending order perm.
// Inputs Ss = S, Mm = M (row-major, stride LDM); Ss, Mm, Zm, Wt are overwritten.
// Returns false (the same in every lane) when the Jacobi sweeps end at their
// cap with a pair still above the rotation test.
template <int LDM>
__device__ bool small_rayleigh_ritz(int nbr, double2* __restrict__ Ss, double2* __restrict__ Mm,
                                    double2* __restrict__ Zm, double2* __restrict__ Wt, double* __restrict__ th,
                                    int* __restrict__ perm, double4* __restrict__ rot, int lane) {
    const int i = lane;
    // Cholesky, right-looking: column j is scaled, then the trailing lower
    // triangle takes its rank-1 update, the elements spread over the lanes
    // (each element subtracts its terms in ascending j, as a left-looking
    // factorisation does)
    for (int j = 0; j < nbr; ++j) {
        const double ljj = sqrt(Ss[j * LDM + j].x);
        const double rl = 1.0 / ljj;
        wsync();
        if (i > j && i < nbr) Ss[i * LDM + j] = zscale(Ss[i * LDM + j], rl);
        if (i == 0) Ss[j * LDM + j] = make_double2(ljj, 0.0);
        wsync();
        for (int e = i; e < LDM * LDM; e += 64) {
            const int r = e / LDM, c = e % LDM;
            if (c > j && r >= c && r < nbr) {
                const double2 lc = Ss[c * LDM + j];
                Ss[r * LDM + c] = zsub(Ss[r * LDM + c], zmul(Ss[r * LDM + j], make_double2(lc.x, -lc.y)));
            }
        }
        wsync();
    }
    // Wt = L^{-1} M  (lane = column of M), then M' = (L^{-1} Wt^H)^H (lane =
    // column of Wt^H, i.e. row i of Wt): with U = L^{-1} Wt^H, M'[i][c] =
    // conj(U[c][i]).  Up to 16 bands the columns stay in registers; wider
    // problems substitute through LDS (the same operations in the same order).
    if constexpr (LDM <= 16) {
        if (i < nbr) {
            double2 v[LDM];
#pragma unroll
            for (int r = 0; r < LDM; ++r) v[r] = r < nbr ? Mm[r * LDM + i] : make_double2(0.0, 0.0);
            lower_solve<LDM>(nbr, Ss, v);
#pragma unroll
            for (int r = 0; r < LDM; ++r)
                if (r < nbr) Wt[r * LDM + i] = v[r];
        }
        wsync();
        if (i < nbr) {
            double2 v[LDM];
#pragma unroll
            for (int r = 0; r < LDM; ++r) {
                const double2 w = r < nbr ? Wt[i * LDM + r] : make_double2(0.0, 0.0);
                v[r] = make_double2(w.x, -w.y);
            }
            lower_solve<LDM>(nbr, Ss, v);
#pragma unroll
            for (int c = 0; c < LDM; ++c)
                if (c < nbr) Mm[i * LDM + c] = make_double2(v[c].x, -v[c].y);
        }
        wsync();
    } else {
        if (i < nbr) {
            for (int r = 0; r < nbr; ++r) {
                double2 s = Mm[r * LDM + i];
                for (int k = 0; k < r; ++k) s = zsub(s, zmul(Ss[r * LDM + k], Wt[k * LDM + i]));
                Wt[r * LDM + i] = zscale(s, 1.0 / Ss[r * LDM + r].x);
            }
        }
        wsync();
        if (i < nbr) {
            for (int r = 0; r < nbr; ++r) {
                const double2 w = Wt[i * LDM + r];
                double2 s = make_double2(w.x, -w.y);
                for (int k = 0; k < r; ++k) s = zsub(s, zmul(Ss[r * LDM + k], Zm[k * LDM + i]));
                Zm[r * LDM + i] = zscale(s, 1.0 / Ss[r * LDM + r].x);  // U[r][i]
            }
        }
        wsync();
        if (i < nbr)
            for (int c = 0; c < nbr; ++c) {
                const double2 u = Zm[c * LDM + i];
                Mm[i * LDM + c] = make_double2(u.x, -u.y);
            }
        wsync();
    }
    if (i < nbr)
        for (int c = 0; c < nbr; ++c) Zm[i * LDM + c] = make_double2(i == c ? 1.0 : 0.0, 0.0);
    if (i < nbr) {
        for (int c = i + 1; c < nbr; ++c) {  // symmetrise
            const double2 a = Mm[i * LDM + c], b = Mm[c * LDM + i];
            const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
            Mm[i * LDM + c] = h;
            Mm[c * LDM + i] = make_double2(h.x, -h.y);
        }
        Mm[i * LDM + i].y = 0.0;
    }
    wsync();
    // Jacobi in round-robin order: m/2 disjoint pairs per round, m-1 rounds per
    // sweep (m = nbr rounded up to even; the pad index takes no rotation).  A pair
    // is rotated while its off-diagonal element still moves an eigenvalue by more
    // than ~1e-13 of the local scale (|c|^2 / gap) or is not negligible outright.
    const int m = (nbr + 1) & ~1, half = m >> 1;
    // the rotation test of a pair (p < q)
    auto negligible = [&](int p, int q) {
        const double2 c = Mm[p * LDM + q];
        const double ac = sqrt(c.x * c.x + c.y * c.y);
        const double a = Mm[p * LDM + p].x, b = Mm[q * LDM + q].x;
        const double scale = fabs(a) + fabs(b) + 1e-300;
        return ac <= 1e-15 * scale || ac * ac <= 1e-13 * scale * fabs(b - a);
    };
    // every pair below the rotation test (wave-uniform)
    auto all_negligible = [&]() {
        int big = 0;
        for (int pq = i; pq < LDM * LDM; pq += 64) {
            const int p = pq / LDM, q = pq % LDM;
            if (p < q && q < nbr && !negligible(p, q)) big = 1;
        }
        for (int o = 32; o > 0; o >>= 1) big |= __shfl_xor(big, o, 64);
        return big == 0;
    };
    bool converged = false;
    for (int sweep = 0; sweep < 20; ++sweep) {
        // a sweep that finds every pair negligible at its start rotates nothing
        // (no round changes M): test all pairs at once and stop
        if (all_negligible()) {
            converged = true;
            break;
        }
        int rotated = 0;
        for (int round = 0; round < m - 1; ++round) {
            auto elem = [&](int pos) {
                if (pos == 0) return 0;
                const int v = pos - 1 + round;  // < 2 (m - 1)
                return (v >= m - 1 ? v - (m - 1) : v) + 1;
            };
            // rotation parameters, one lane per pair
            int did = 0;
            if (i < half) {
                int p = elem(i), q = elem(m - 1 - i);
                if (p > q) { const int t = p; p = q; q = t; }
                double4 g = make_double4(1.0, 0.0, 1.0, 0.0);  // cs, sn, e^{-i phi}
                if (q < nbr) {
                    const double2 c = Mm[p * LDM + q];
                    const double ac = sqrt(c.x * c.x + c.y * c.y);
                    const double a = Mm[p * LDM + p].x, b = Mm[q * LDM + q].x;
                    if (!negligible(p, q)) {
                        did = 1;
                        const double emx = c.x / ac, emy = -c.y / ac;  // e^{-i phi}, c = |c| e^{i phi}
                        const double tau_ = (b - a) / (2.0 * ac);
                        const double t = (tau_ >= 0 ? 1.0 : -1.0) / (fabs(tau_) + sqrt(1.0 + tau_ * tau_));
                        const double cs = 1.0 / sqrt(1.0 + t * t);
                        g = make_double4(cs, t * cs, emx, emy);
                    }
                }
                rot[i] = g;
            }
            for (int o = 32; o > 0; o >>= 1) did |= __shfl_xor(did, o, 64);
            wsync();
            if (did) {
                rotated = 1;
                // columns: M <- M G, Z <- Z G with G = [[cs, sn], [-sn e, cs e]] on (p, q)
                for (int task = i; task < half * LDM; task += 64) {
                    const int k = task / LDM, row = task % LDM;
                    if (row >= nbr) continue;
                    int p = elem(k), q = elem(m - 1 - k);
                    if (p > q) { const int t = p; p = q; q = t; }
                    if (q >= nbr) continue;
                    const double4 g = rot[k];
                    const double2 em = make_double2(g.z, g.w);
                    const double2 mp = Mm[row * LDM + p], mq = Mm[row * LDM + q];
                    const double2 emq = zmul(em, mq);
                    Mm[row * LDM + p] = zsub(zscale(mp, g.x), zscale(emq, g.y));
                    Mm[row * LDM + q] = zadd(zscale(mp, g.y), zscale(emq, g.x));
                    const double2 zp = Zm[row * LDM + p], zq = Zm[row * LDM + q];
                    const double2 ezq = zmul(em, zq);
                    Zm[row * LDM + p] = zsub(zscale(zp, g.x), zscale(ezq, g.y));
                    Zm[row * LDM + q] = zadd(zscale(zp, g.y), zscale(ezq, g.x));
                }
                wsync();
                // rows: M <- G^H M
                for (int task = i; task < half * LDM; task += 64) {
                    const int k = task / LDM, col = task % LDM;
                    if (col >= nbr) continue;
                    int p = elem(k), q = elem(m - 1 - k);
                    if (p > q) { const int t = p; p = q; q = t; }
                    if (q >= nbr) continue;
                    const double4 g = rot[k];
                    const double2 emc = make_double2(g.z, -g.w);  // conj(e^{-i phi})
                    const double2 rp = Mm[p * LDM + col], rq = Mm[q * LDM + col];
                    const double2 erq = zmul(emc, rq);
                    Mm[p * LDM + col] = zsub(zscale(rp, g.x), zscale(erq, g.y));
                    Mm[q * LDM + col] = zadd(zscale(rp, g.y), zscale(erq, g.x));
                }
                wsync();
                if (i < half) {
                    int p = elem(i), q = elem(m - 1 - i);
                    if (p > q) { const int t = p; p = q; q = t; }
                    if (q < nbr && rot[i].y != 0.0) {
                        Mm[p * LDM + q] = make_double2(0.0, 0.0);
                        Mm[q * LDM + p] = make_double2(0.0, 0.0);
                        Mm[p * LDM + p].y = 0.0;
                        Mm[q * LDM + q].y = 0.0;
                    }
                }
                wsync();
            }
        }
        if (!rotated) {
            converged = true;
            break;
        }
    }
    if (!converged) converged = all_negligible();
    wsync();
    // C = L^{-H} Z (back substitution with L^H, lane = column of Z)
    if constexpr (LDM <= 16) {
        if (i < nbr) {
            double2 v[LDM];
#pragma unroll
            for (int r = 0; r < LDM; ++r) v[r] = r < nbr ? Zm[r * LDM + i] : make_double2(0.0, 0.0);
            upper_solve_h<LDM>(nbr, Ss, v);
#pragma unroll
            for (int r = 0; r < LDM; ++r)
                if (r < nbr) Wt[r * LDM + i] = v[r];
        }
    } else {
        if (i < nbr) {
            for (int r = nbr - 1; r >= 0; --r) {
                double2 s = Zm[r * LDM + i];
                for (int k = r + 1; k < nbr; ++k) s = zsub(s, zmulc(Ss[k * LDM + r], Wt[k * LDM + i]));
                Wt[r * LDM + i] = zscale(s, 1.0 / Ss[r * LDM + r].x);
            }
        }
    }
    if (i < nbr) {
        th[i] = Mm[i * LDM + i].x;
        perm[i] = i;  // keeps perm a valid index map should a rank below collide (NaN)
    }
    wsync();
    // ascending order: the rank of Ritz value i (ties by index)
    if (i < nbr) {
        const double ti = th[i];
        int rank = 0;
        for (int c = 0; c < nbr; ++c) {
            const double tc = th[c];
            rank += (tc < ti || (tc == ti && c < i)) ? 1 : 0;
        }
        perm[rank] = i;
    }
    wsync();
    return converged;
}

// Failure of the k-point's eigen-solve (one wave, after the small problem):
// the tridiagonal stage's flag, Jacobi sweeps that did not converge, or a
// non-finite Ritz value among the nb returned.  Failures add to status[0] and
// lower status[1] to the k-point's index in the whole solve.
__device__ void report_ritz_status(const int* kstat, unsigned long long* status, int64_t k0, int64_t kk, int nb,
                                   const double* th, const int* perm, bool converged, int lane) {
    if (!status) return;
    int bad = 0;
    for (int b = lane; b < nb; b += 64) bad |= isfinite(th[perm[b]]) ? 0 : 1;
    for (int o = 32; o > 0; o >>= 1) bad |= __shfl_xor(bad, o, 64);
    if (lane == 0 && (bad || !converged || (kstat && kstat[kk]))) {
        atomicAdd(&status[0], 1ull);
        atomicMin(&status[1], (unsigned long long)(k0 + kk));
    }
}

// X1 = X0 C[:, perm[b]] for the lowest nb Ritz pairs: eigenvalues, unit
// eigenvectors and Hellmann-Feynman gradients (Hamiltonian.cpp:146-194).  One
// thread per row and up to eight bands per pass, so a pass reads X0 once;
// red: LDS scratch of 24 (BS / 64) doubles that is free at this point.
template <int BS, int LDM>
__device__ void write_ritz(const double2* __restrict__ X, int n, int nbr, int nb, const double2* Wt,
                           const double* th, const int* perm, int64_t kk, double* evals, double2* evecs,
                           double* grads, const DevParams& P, const double* kpts, const int32_t* G, double* red) {
    constexpr int NW = BS / 64, BC = 8;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid < nb) evals[kk * nb + tid] = th[perm[tid]];
    if (!evecs && !grads) return;
    double2* out = evecs ? evecs + kk * (int64_t)nb * n : nullptr;
    const double* k = kpts + 3 * kk;
    for (int b0 = 0; b0 < nb; b0 += BC) {
        double gx[BC], gy[BC], gz[BC];
#pragma unroll
        for (int u = 0; u < BC; ++u) gx[u] = gy[u] = gz[u] = 0.0;
        for (int r = tid; r < n; r += BS) {
            double2 acc[BC];
#pragma unroll
            for (int u = 0; u < BC; ++u) acc[u] = make_double2(0.0, 0.0);
            for (int a = 0; a < nbr; ++a) {
                const double2 x = X[a * n + r];
#pragma unroll
                for (int u = 0; u < BC; ++u)
                    if (b0 + u < nb) {
                        const double2 c = Wt[a * LDM + perm[b0 + u]];
                        acc[u].x = fma(x.x, c.x, fma(-x.y, c.y, acc[u].x));
                        acc[u].y = fma(x.x, c.y, fma(x.y, c.x, acc[u].y));
                    }
            }
#pragma unroll
            for (int u = 0; u < BC; ++u)
                if (b0 + u < nb) {
                    if (out) out[(int64_t)(b0 + u) * n + r] = acc[u];
                    if (grads) {
                        // Hellmann-Feynman, kinetic term (Hamiltonian.cpp:146-194); SOC: spin-down
                        // row r of basis vector r - N adds its weight to the same k + G.
                        const int i = r < P.N ? r : r - P.N;
                        const double wt = acc[u].x * acc[u].x + acc[u].y * acc[u].y;
                        gx[u] += (k[0] + (double)G[3 * i]) * wt;
                        gy[u] += (k[1] + (double)G[3 * i + 1]) * wt;
                        gz[u] += (k[2] + (double)G[3 * i + 2]) * wt;
                    }
                }
        }
        if (grads) {
            // wave sums, then the waves' partials in LDS, added in wave order
#pragma unroll
            for (int u = 0; u < BC; ++u)
                for (int o = 32; o > 0; o >>= 1) {
                    gx[u] += __shfl_xor(gx[u], o, 64);
                    gy[u] += __shfl_xor(gy[u], o, 64);
                    gz[u] += __shfl_xor(gz[u], o, 64);
                }
            if (lane == 0)
#pragma unroll
                for (int u = 0; u < BC; ++u) {
                    red[(wv * BC + u) * 3 + 0] = gx[u];
                    red[(wv * BC + u) * 3 + 1] = gy[u];
                    red[(wv * BC + u) * 3 + 2] = gz[u];
                }
            __syncthreads();
            if (tid < 3 * BC) {
                const int u = tid / 3, c = tid % 3;
                if (b0 + u < nb) {
                    double sum = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) sum += red[(w * BC + u) * 3 + c];
                    grads[3 * (kk * nb + b0 + u) + c] = sum * P.grad_pref * P.frac_to_phys;
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// nbr <= 16: Y = H X0 on FP64 MFMA (16x16x4), one 16-row tile of Y at a time;
// the accumulator layout of a Y tile (lane l, item v: row 4v + l/16, band l%16)
// is the B-operand layout of the 4-row chunks {4v..4v+3}, so S = X0^H X0 and
// M = X0^H Y accumulate on MFMA straight from the Y accumulators and Y never
// leaves registers.  Waves split the row tiles; their partial S, M meet in LDS.
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma64(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int BS>
__global__ __launch_bounds__(BS, 4) void refine_mfma_kernel(const double2* __restrict__ H, int n, int ld,
                                                         const double2* __restrict__ Xall, int nbr, int nb,
                                                         double* __restrict__ evals, double2* __restrict__ evecs,
                                                         double* __restrict__ grads, DevParams P,
                                                         const double* __restrict__ kpts,
                                                         const int32_t* __restrict__ G, const int* __restrict__ kstat,
                                                         unsigned long long* __restrict__ status, int64_t k0) {
    constexpr int NW = BS / 64;
    // per-wave partial S_re, S_im, M_re, M_im; once S and M are summed (in
    // registers) the same LDS holds the small problem's M, S, Z and W, so a
    // workgroup takes 33 KB and four fit on a CU
    __shared__ __align__(16) double part[NW][4][16][16];
    static_assert(sizeof(part) >= 4 * 16 * 16 * sizeof(double2), "M, S, Z and W fit in the partial slots");
    double2* Mm = reinterpret_cast<double2*>(&part[0][0][0][0]);
    double2* Ss = Mm + 16 * 16;
    double2* Zm = Ss + 16 * 16;
    double2* Wt = Zm + 16 * 16;
    __shared__ double th[16];
    __shared__ int perm[16];
    __shared__ double4 rot[9];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int64_t kk = blockIdx.x;
    const double2* A = H + kk * (int64_t)ld * n;
    const double2* X = Xall + kk * (int64_t)nbr * n;  // X[b * n + r]
    const bool bl = l16 < nbr;
    const double2* Xb = X + (bl ? l16 : 0) * (int64_t)n;  // this lane's band column

    dbl4 sre = {0, 0, 0, 0}, sim = sre, mre = sre, mim = sre;
    const int ntile = (n + 15) >> 4;
    for (int t = wv; t < ntile; t += NW) {
        const int row = 16 * t + l16;  // A-operand row of this lane
        dbl4 yre = {0, 0, 0, 0}, yim = yre;
        // kUnroll column quads per pass: all loads issue before the MFMAs that
        // consume them (the accumulation order is still column-ascending)
        for (int c0 = 0; c0 < n; c0 += 4 * kUnroll) {
            double hr[kUnroll], hi[kUnroll], xr[kUnroll], xi[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                // out-of-range elements load an in-range address and are masked by a
                // 0 / 1 factor (a select would let the compiler sink the load into a branch)
                const int col = c0 + 4 * u + l4;
                const bool hc = row < n && col < n;
                const double2 h = A[(hc ? row : 0) + (int64_t)(hc ? col : 0) * ld];
                const double2 x = Xb[col < n ? col : 0];
                const double mh = hc ? 1.0 : 0.0, mx = (bl && col < n) ? 1.0 : 0.0;
                hr[u] = h.x * mh;
                hi[u] = h.y * mh;
                xr[u] = x.x * mx;
                xi[u] = x.y * mx;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                yre = mfma64(hr[u], xr[u], yre);
                yre = mfma64(-hi[u], xi[u], yre);
                yim = mfma64(hr[u], xi[u], yim);
                yim = mfma64(hi[u], xr[u], yim);
            }
        }
        // S, M over this tile's rows, chunk v = rows 16t + 4v .. 16t + 4v + 3
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * t + 4 * v + l4;
            const double2 x = Xb[r < n ? r : 0];
            const double xr = (bl && r < n) ? x.x : 0.0, xi = (bl && r < n) ? x.y : 0.0;
            sre = mfma64(xr, xr, sre);
            sre = mfma64(xi, xi, sre);
            sim = mfma64(xr, xi, sim);
            sim = mfma64(-xi, xr, sim);
            mre = mfma64(xr, yre[v], mre);
            mre = mfma64(xi, yim[v], mre);
            mim = mfma64(xr, yim[v], mim);
            mim = mfma64(-xi, yre[v], mim);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = 4 * v + l4;  // D row (band a), column l16 (band b)
        part[wv][0][r][l16] = sre[v];
        part[wv][1][r][l16] = sim[v];
        part[wv][2][r][l16] = mre[v];
        part[wv][3][r][l16] = mim[v];
    }
    __syncthreads();
    static_assert(BS == 256, "one (a, b) entry of S and M per thread");
    {
        const int a = tid >> 4, b = tid & 15;
        double s0 = 0, s1 = 0, m0 = 0, m1 = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            s0 += part[w][0][a][b];
            s1 += part[w][1][a][b];
            m0 += part[w][2][a][b];
            m1 += part[w][3][a][b];
        }
        __syncthreads();  // every partial is read before Ss and Mm overwrite them
        if (a < nbr && b < nbr) {
            Ss[a * 16 + b] = make_double2(s0, s1);
            Mm[a * 16 + b] = make_double2(m0, m1);
        }
    }
    __syncthreads();
    if (wv == 0) {
        const bool conv = small_rayleigh_ritz<16>(nbr, Ss, Mm, Zm, Wt, th, perm, rot, lane);
        report_ritz_status(kstat, status, k0, kk, nb, th, perm, conv, lane);
    }
    __syncthreads();
    write_ritz<BS, 16>(X, n, nbr, nb, Wt, th, perm, kk, evals, evecs, grads, P, kpts, G,
                       reinterpret_cast<double*>(Mm));
}

// 16 < nbr <= 32: the same scheme on two 16-band tiles.  A lane's MFMA column
// is band l%16 of tile 0 and band 16 + l%16 of tile 1; the H tile (A operand)
// is shared by both Y tiles, S and M take four 16 x 16 blocks each.  The
// per-wave partials meet in one LDS slot, the waves adding in fixed order,
// so the sums do not depend on scheduling.
template <int BS>
__global__ __launch_bounds__(BS, 2) void refine_mfma2_kernel(const double2* __restrict__ H, int n, int ld,
                                                          const double2* __restrict__ Xall, int nbr, int nb,
                                                          double* __restrict__ evals, double2* __restrict__ evecs,
                                                          double* __restrict__ grads, DevParams P,
                                                          const double* __restrict__ kpts,
                                                          const int32_t* __restrict__ G, const int* __restrict__ kstat,
                                                          unsigned long long* __restrict__ status, int64_t k0) {
    constexpr int NW = BS / 64;
    // One partial slot filled by the waves in turn; once S and M are formed the
    // slot is reused for the Jacobi work arrays Zm and Wt (66 KB of LDS in all,
    // two workgroups per CU).
    __shared__ __align__(16) double part[4][4][16][16];  // block (at, bt), S_re S_im M_re M_im
    static_assert(sizeof(part) >= 2 * 32 * 32 * sizeof(double2), "Zm and Wt fit in the partial slot");
    __shared__ double2 Mm[32 * 32];
    __shared__ double2 Ss[32 * 32];
    double2* Zm = reinterpret_cast<double2*>(&part[0][0][0][0]);
    double2* Wt = Zm + 32 * 32;
    __shared__ double th[32];
    __shared__ int perm[32];
    __shared__ double4 rot[17];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int64_t kk = blockIdx.x;
    const double2* A = H + kk * (int64_t)ld * n;
    const double2* X = Xall + kk * (int64_t)nbr * n;  // X[b * n + r]
    bool bl[2];
    const double2* Xb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        bl[t] = 16 * t + l16 < nbr;
        Xb[t] = X + (bl[t] ? 16 * t + l16 : 0) * (int64_t)n;
    }

    dbl4 acc[4][4];  // [block at*2+bt][S_re, S_im, M_re, M_im]
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] = dbl4{0, 0, 0, 0};
    const int ntile = (n + 15) >> 4;
    for (int t = wv; t < ntile; t += NW) {
        const int row = 16 * t + l16;  // A-operand row of this lane
        dbl4 yre[2], yim[2];
#pragma unroll
        for (int bt = 0; bt < 2; ++bt) yre[bt] = yim[bt] = dbl4{0, 0, 0, 0};
        for (int c0 = 0; c0 < n; c0 += 4 * kUnroll2) {
            double hr[kUnroll2], hi[kUnroll2], xr[kUnroll2][2], xi[kUnroll2][2];
#pragma unroll
            for (int u = 0; u < kUnroll2; ++u) {
                const int col = c0 + 4 * u + l4;
                const bool hc = row < n && col < n;
                const double2 h = A[(hc ? row : 0) + (int64_t)(hc ? col : 0) * ld];
                const double mh = hc ? 1.0 : 0.0;
                hr[u] = h.x * mh;
                hi[u] = h.y * mh;
#pragma unroll
                for (int bt = 0; bt < 2; ++bt) {
                    const double2 x = Xb[bt][col < n ? col : 0];
                    const double mx = (bl[bt] && col < n) ? 1.0 : 0.0;
                    xr[u][bt] = x.x * mx;
                    xi[u][bt] = x.y * mx;
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll2; ++u)
#pragma unroll
                for (int bt = 0; bt < 2; ++bt) {
                    yre[bt] = mfma64(hr[u], xr[u][bt], yre[bt]);
                    yre[bt] = mfma64(-hi[u], xi[u][bt], yre[bt]);
                    yim[bt] = mfma64(hr[u], xi[u][bt], yim[bt]);
                    yim[bt] = mfma64(hi[u], xr[u][bt], yim[bt]);
                }
        }
        // S, M over this tile's rows, chunk v = rows 16t + 4v .. 16t + 4v + 3
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * t + 4 * v + l4;
            double xr[2], xi[2];
#pragma unroll
            for (int bt = 0; bt < 2; ++bt) {
                const double2 x = Xb[bt][r < n ? r : 0];
                xr[bt] = (bl[bt] && r < n) ? x.x : 0.0;
                xi[bt] = (bl[bt] && r < n) ? x.y : 0.0;
            }
#pragma unroll
            for (int at = 0; at < 2; ++at)
#pragma unroll
                for (int bt = 0; bt < 2; ++bt) {
                    dbl4* a4 = acc[2 * at + bt];
                    a4[0] = mfma64(xr[at], xr[bt], a4[0]);
                    a4[0] = mfma64(xi[at], xi[bt], a4[0]);
                    a4[1] = mfma64(xr[at], xi[bt], a4[1]);
                    a4[1] = mfma64(-xi[at], xr[bt], a4[1]);
                    a4[2] = mfma64(xr[at], yre[bt][v], a4[2]);
                    a4[2] = mfma64(xi[at], yim[bt][v], a4[2]);
                    a4[3] = mfma64(xr[at], yim[bt][v], a4[3]);
                    a4[3] = mfma64(-xi[at], yre[bt][v], a4[3]);
                }
        }
    }
    // partials: wave rd writes (rd = 0) or adds its share in round rd
    for (int rd = 0; rd < NW; ++rd) {
        if (wv == rd) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        double& dst = part[q][c][4 * v + l4][l16];
                        dst = rd == 0 ? acc[q][c][v] : dst + acc[q][c][v];
                    }
        }
        __syncthreads();
    }
    for (int e = tid; e < 4 * 256; e += BS) {
        const int q = e >> 8, a = (e >> 4) & 15, b = e & 15;
        const int at = q >> 1, bt = q & 1;
        const int ga = 16 * at + a, gb = 16 * bt + b;
        if (ga < nbr && gb < nbr) {
            Ss[ga * 32 + gb] = make_double2(part[q][0][a][b], part[q][1][a][b]);
            Mm[ga * 32 + gb] = make_double2(part[q][2][a][b], part[q][3][a][b]);
        }
    }
    __syncthreads();  // part is free from here: Zm and Wt take its place
    if (wv == 0) {
        const bool conv = small_rayleigh_ritz<32>(nbr, Ss, Mm, Zm, Wt, th, perm, rot, lane);
        report_ritz_status(kstat, status, k0, kk, nb, th, perm, conv, lane);
    }
    __syncthreads();
    write_ritz<BS, 32>(X, n, nbr, nb, Wt, th, perm, kk, evals, evecs, grads, P, kpts, G,
                       reinterpret_cast<double*>(Mm));
}

const char* refine_kernel_name(int nbr) { return nbr <= 16 ? "refine_mfma_kernel" : "refine_mfma2_kernel"; }

hipError_t launch_refine(const double2* dH, int n, int ld, double2* dX, int nbr, int nb, double* devals,
                         double2* devecs, double* dgrads, const DevParams& P, const double* dk, const int32_t* dG,
                         int64_t nk, const int* kstat, unsigned long long* status, int64_t k0, hipStream_t s) {
    if (nbr > kNbrMax || nb > nbr) return hipErrorInvalidValue;
    if (nbr <= 16)
        hipLaunchKernelGGL(refine_mfma_kernel<kRefBS>, dim3((unsigned)nk), dim3(kRefBS), 0, s, dH, n, ld, dX, nbr,
                           nb, devals, devecs, dgrads, P, dk, dG, kstat, status, k0);
    else
        hipLaunchKernelGGL(refine_mfma2_kernel<kRefBS>, dim3((unsigned)nk), dim3(kRefBS), 0, s, dH, n, ld, dX, nbr,
                           nb, devals, devecs, dgrads, P, dk, dG, kstat, status, k0);
    return hipGetLastError();
}

}  // namespace uepm
