// solver_f64.hip — assembly and tridiagonal eigen-stage of the band solve on gfx950.
//
//   (a) assemble_tiled_kernel / assemble_kernel  H(k) for a batch of k-points
//                           (Hamiltonian::SetMatrix + the 0.5 (H + H^dagger)
//                           step of Diagonalize, src/EPP/Hamiltonian.cpp:50-127)
//   (c) tri_eig_wave_kernel bisection (Sturm counts, multisection) for the lowest
//                           nb eigenvalues of T and inverse iteration for their
//                           vectors, one wave per k
//
// The Householder reduction (b) lives in tridiag_reg.hip / tridiag_blocked.hip
// and the back-transform (d) in backtransform.hip; (b)-(d) replace
// Eigen::SelfAdjointEigenSolver::compute (Hamiltonian.cpp:129): same
// eigenvalues to within the backward-stability envelope of a Householder
// reduction, eigenvectors up to phase (and basis inside degenerate clusters).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstdlib>

#include "epm_physics.hpp"
#include "solver.hpp"

namespace uepm {

// ------------------------------------------------------------ (a) assembly
// h_ij and h_ji of the pair (i, j) before hermitisation and spin-orbit: V(G),
// the kinetic diagonal and the non-local term.  h_ji comes from h_ij where
// that is exact: the phase argument of V(G) negates exactly under i <-> j and
// the device sincos is exactly odd / even there; the non-local phase is
// cos / sin of theta_i - theta_j formed from the per-vector cos / sin, which
// swap to the exact conjugate (products commute, no FMA contraction); the
// square-well radial integrals are symmetric in |K_i|, |K_j| outside the
// near-equal branch of F_l (which reads one side's Bessel values) -- so one
// evaluation gives the reference's two-sided 0.5 (h_ij + conj(h_ji)) bit for
// bit.  Gaussian l = 2 wells and near-equal norms are evaluated both ways.
__device__ __forceinline__ void pair_terms(const DevParams& P, const KState& si, const KState& sj, int gx, int gy,
                                           int gz, bool diag, const double2* vtab, bool swap_exact, double& re_ij,
                                           double& im_ij, double& re_ji, double& im_ji) {
    local_V(P, gx, gy, gz, re_ij, im_ij, vtab);
    re_ji = re_ij;
    im_ji = -im_ij;
    if (diag) {
        const double KG2 = P.kin * (si.kx * si.kx + si.ky * si.ky + si.kz * si.kz);
        re_ij += KG2;
        re_ji += KG2;
    }
    if (P.nl) {
        double r1, i1, r2, i2;
        nonlocal(P, si, sj, r1, i1);
        if (swap_exact && fabs(si.nK - sj.nK) > 1.0e-4) {
            r2 = r1;
            i2 = -i1;
        } else {
            nonlocal(P, sj, si, r2, i2);
        }
        re_ij += r1; im_ij += i1;
        re_ji += r2; im_ji += i2;
    }
}

// One workgroup per k.  Per-basis-vector state in LDS, then every (i, j) pair
// of the lower triangle: the stored value is the reference's
// 0.5 (h_ij + conj(h_ji)) (its Diagonalize hermitisation), mirrored into the
// upper triangle; h_ji is derived from h_ij where that is exact (below).
// Output: column-major dim x dim, leading dimension ld, one matrix per k.
template <int BS>
__global__ __launch_bounds__(BS) void assemble_kernel(DevParams P, const int32_t* __restrict__ G,
                                                       const double* __restrict__ kpts, double2* __restrict__ H,
                                                       int ld) {
    extern __shared__ __align__(16) unsigned char smem[];
    KState* ks = reinterpret_cast<KState*>(smem);
    __shared__ double2 vtab[kVPhaseTab];
    const int kk = blockIdx.x;
    const double* k = kpts + 3 * (size_t)kk;
    const int N = P.N;
    for (int i = threadIdx.x; i < N; i += BS) kstate_init(P, k, G, i, ks[i]);
    for (int t = threadIdx.x; t < kVPhaseTab; t += BS) vtab[t] = local_phase(P, t / 49 - 3, (t / 7) % 7 - 3, t % 7 - 3);
    __syncthreads();
    double2* A = H + (size_t)kk * ld * P.dim;
    const int npairs = N * (N + 1) / 2;
    // h_ji from h_ij where exact (pair_terms)
    const bool swap_exact = !(P.well != 0 && (P.need_l2_a || P.need_l2_c));
    // element t -> (i, j), column-major over the packed lower triangle; the
    // first from the closed form, then advanced by BS per iteration
    int t = threadIdx.x, j = 0, i = 0;
    if (t < npairs) {
        j = (int)((2.0 * N + 1.0 - sqrt((2.0 * N + 1.0) * (2.0 * N + 1.0) - 8.0 * t)) * 0.5);
        while (j > 0 && j * (2 * N - j + 1) / 2 > t) --j;
        while ((j + 1) * (2 * N - j) / 2 <= t) ++j;
        i = j + (t - j * (2 * N - j + 1) / 2);
    }
    for (; t < npairs; t += BS) {
        if (t != (int)threadIdx.x) {
            i += BS;
            while (i >= N && j < N - 1) {  // past the end of column j: continue at the diagonal of column j + 1
                i -= N - (j + 1);
                ++j;
            }
        }
        const KState& si = ks[i];
        const KState& sj = ks[j];
        const int gx = G[3 * i] - G[3 * j], gy = G[3 * i + 1] - G[3 * j + 1], gz = G[3 * i + 2] - G[3 * j + 2];
        double re_ij, im_ij, re_ji, im_ji;
        pair_terms(P, si, sj, gx, gy, gz, i == j, vtab, swap_exact, re_ij, im_ij, re_ji, im_ji);
        if (!P.soc) {
            const double hr = 0.5 * (re_ij + re_ji), hi = 0.5 * (im_ij - im_ji);
            A[i + (size_t)j * ld] = make_double2(hr, hi);
            A[j + (size_t)i * ld] = make_double2(hr, -hi);
        } else {
            double s_ij[8], s_ji[8];
            soc_block(P, si, sj, gx, gy, gz, s_ij);
            soc_block(P, sj, si, -gx, -gy, -gz, s_ji);
            // blocks: [[M+S00, S01],[S10, M+S11]]; element (row, col)
            // (i, j) of block (p, q) pairs with (j, i) of block (q, p).
            const int n0 = N;
            // up-up
            {
                const double ar = re_ij + s_ij[0], ai = im_ij + s_ij[1];
                const double br = re_ji + s_ji[0], bi = im_ji + s_ji[1];
                const double hr = 0.5 * (ar + br), hi = 0.5 * (ai - bi);
                A[i + (size_t)j * ld] = make_double2(hr, hi);
                A[j + (size_t)i * ld] = make_double2(hr, -hi);
            }
            // down-down
            {
                const double ar = re_ij + s_ij[6], ai = im_ij + s_ij[7];
                const double br = re_ji + s_ji[6], bi = im_ji + s_ji[7];
                const double hr = 0.5 * (ar + br), hi = 0.5 * (ai - bi);
                A[(i + n0) + (size_t)(j + n0) * ld] = make_double2(hr, hi);
                A[(j + n0) + (size_t)(i + n0) * ld] = make_double2(hr, -hi);
            }
            // down-up (row i+N, col j) pairs with up-down (row j, col i+N)
            {
                const double hr = 0.5 * (s_ij[4] + s_ji[2]), hi = 0.5 * (s_ij[5] - s_ji[3]);
                A[(i + n0) + (size_t)j * ld] = make_double2(hr, hi);
                A[j + (size_t)(i + n0) * ld] = make_double2(hr, -hi);
            }
            if (i != j) {
                // down-up (row j+N, col i) pairs with up-down (row i, col j+N)
                const double hr = 0.5 * (s_ji[4] + s_ij[2]), hi = 0.5 * (s_ji[5] - s_ij[3]);
                A[(j + n0) + (size_t)i * ld] = make_double2(hr, hi);
                A[i + (size_t)(j + n0) * ld] = make_double2(hr, -hi);
            }
        }
    }
}

// Without spin-orbit: each wave takes 16 x 16 tiles (bi >= bj) of the lower
// triangle, writes the tile and, through its LDS slot, the conjugate
// transpose above the diagonal, so both halves go out in 256-byte column
// runs (the mirrored writes of assemble_kernel are one element per column).
// Element values are assemble_kernel's, bit for bit.
template <int BS>
__global__ __launch_bounds__(BS) void assemble_tiled_kernel(DevParams P, const int32_t* __restrict__ G,
                                                             const double* __restrict__ kpts,
                                                             double2* __restrict__ H, int ld) {
    extern __shared__ __align__(16) unsigned char smem[];
    KState* ks = reinterpret_cast<KState*>(smem);
    __shared__ double2 vtab[kVPhaseTab];
    __shared__ double2 T[BS / 64][16][17];
    const int kk = blockIdx.x;
    const double* k = kpts + 3 * (size_t)kk;
    const int N = P.N;
    for (int i = threadIdx.x; i < N; i += BS) kstate_init(P, k, G, i, ks[i]);
    for (int t = threadIdx.x; t < kVPhaseTab; t += BS) vtab[t] = local_phase(P, t / 49 - 3, (t / 7) % 7 - 3, t % 7 - 3);
    __syncthreads();
    double2* A = H + (size_t)kk * ld * P.dim;
    const bool swap_exact = !(P.well != 0 && (P.need_l2_a || P.need_l2_c));
    const int nt = (N + 15) >> 4, ntiles = nt * (nt + 1) / 2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int il = lane & 15, jq = lane >> 4;
    double2(*Tw)[17] = T[wv];
    for (int tile = wv; tile < ntiles; tile += BS / 64) {
        // tile -> (bi, bj), column-major over the lower triangle of tiles
        int bj = 0, rem = tile;
        while (rem >= nt - bj) {
            rem -= nt - bj;
            ++bj;
        }
        const int bi = bj + rem;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int jl = jq + 4 * q;
            const int i = 16 * bi + il, j = 16 * bj + jl;
            if (i < N && j < N && i >= j) {
                const KState& si = ks[i];
                const KState& sj = ks[j];
                const int gx = G[3 * i] - G[3 * j], gy = G[3 * i + 1] - G[3 * j + 1], gz = G[3 * i + 2] - G[3 * j + 2];
                double re_ij, im_ij, re_ji, im_ji;
                pair_terms(P, si, sj, gx, gy, gz, i == j, vtab, swap_exact, re_ij, im_ij, re_ji, im_ji);
                const double hr = 0.5 * (re_ij + re_ji), hi = 0.5 * (im_ij - im_ji);
                // the diagonal as assemble_kernel's mirrored write leaves it
                A[i + (size_t)j * ld] = make_double2(hr, i == j ? -hi : hi);
                Tw[jl][il] = make_double2(hr, hi);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rl = il, cl = jq + 4 * q;
            const int row = 16 * bj + rl, col = 16 * bi + cl;
            if (row < N && col < N && row < col) {
                const double2 h = Tw[rl][cl];  // h(col, row)
                A[row + (size_t)col * ld] = make_double2(h.x, -h.y);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------- (c) tridiagonal eigenvalues / vectors
// Start vector of inverse iteration: a fixed pseudo-random value per (row, band).
__device__ __forceinline__ double lane_rand(uint32_t i, uint32_t b) {
    uint32_t h = i * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (double)(h >> 8) * (1.0 / 8388608.0) - 1.0;
}

// Wave-per-k variant: the d, e of a k-point are staged in LDS (16 n bytes of
// dynamic LDS) and read from there by every Sturm count and inverse-iteration
// sweep; lanes = bands x S shifts (multisection: one sweep narrows the bracket
// of every band S+1-fold); inverse iteration on the first lane of a band.
__device__ __forceinline__ double rcp_nr64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// 1/x from v_rcp_f64 and one Newton step: the Sturm recurrence is backward
// stable, and an error of a few ulp in a pivot moves only the counts of
// probes within rounding of an eigenvalue
__device__ __forceinline__ double rcp_nr64_1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}

// number of eigenvalues of T below x (Sturm count, pivmin-guarded)
__device__ __forceinline__ int sturm_count_u(const double* __restrict__ d, const double* __restrict__ e, int n,
                                             double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = q < 0.0;
    int i = 1;
    for (; i + 8 <= n; i += 8) {
        double dd[8], ee[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            dd[u] = d[i + u];
            ee[u] = e[i + u - 1];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            q = (dd[u] - x) - (ee[u] * ee[u]) * rcp_nr64_1(q);
            if (fabs(q) < pivmin) q = -pivmin;
            c += q < 0.0;
        }
    }
    for (; i < n; ++i) {
        q = (d[i] - x) - (e[i - 1] * e[i - 1]) * rcp_nr64_1(q);
        if (fabs(q) < pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

// abstol_rel: bisection also stops once a bracket is narrower than
// abstol_rel * ||T||_inf (0: to the rounding level of the eigenvalue itself).
template <int S>
__global__ __launch_bounds__(64, 4) void tri_eig_wave_kernel(const double* __restrict__ dall,
                                                          const double* __restrict__ eall, int n, int nb, int64_t nk,
                                                          double* __restrict__ evals, double* __restrict__ Z,
                                                          double* __restrict__ W, int inv_sweeps,
                                                          double abstol_rel, int* __restrict__ kstat,
                                                          unsigned long long* __restrict__ status, int64_t k0) {
    constexpr int NBW = 64 / S;  // bands per wave
    const int64_t kk = blockIdx.x;
    const int lane = threadIdx.x, sub = lane % S;
    const int b = blockIdx.y * NBW + lane / S;
    const bool live = b < nb;
    const double* d = dall + kk * n;
    const double* e = eall + kk * n;
    // Gershgorin interval and scale (wave-uniform)
    double lo = DBL_MAX, hi = -DBL_MAX, emax2 = 0.0, tnorm = 0.0;
    bool finite = true;
    for (int i = 0; i < n; ++i) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double er = i + 1 < n ? fabs(e[i]) : 0.0;
        finite = finite && isfinite(d[i]) && isfinite(er);
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        emax2 = fmax(emax2, er * er);
        tnorm = fmax(tnorm, fabs(d[i]) + el + er);
    }
    // A non-finite T (a non-finite k-point, or H(k) overflowing) has no
    // eigen-decomposition (the reference's Eigen solver reports NoConvergence,
    // Hamiltonian.cpp:130-133): NaN eigenvalues and vectors, and the k-point is
    // flagged -- in kstat for a later stage to report, or in status directly
    // when this is the last stage.  Both conditions are the same in every wave.
    if (blockIdx.y == 0 && lane == 0) {
        if (kstat) kstat[kk] = finite ? 0 : 1;
        else if (status && !finite) {
            atomicAdd(&status[0], 1ull);
            atomicMin(&status[1], (unsigned long long)(k0 + kk));
        }
    }
    if (!finite) {
        if (live && sub == 0) {
            evals[kk * nb + b] = __builtin_nan("");
            if (inv_sweeps > 0)
                for (int i = 0; i < n; ++i) Z[(size_t)(kk * nb + b) * n + i] = __builtin_nan("");
        }
        return;
    }
    // d and e in LDS: the Sturm counts re-read them every probe sweep, and the
    // inverse iteration every sweep (3-8% off a launch alone, against reading
    // them from global memory; the same values, so the same results)
    extern __shared__ double tri_lds[];
    double* const sd = tri_lds;
    double* const se = tri_lds + n;
    for (int i = lane; i < n; i += 64) {
        sd[i] = d[i];
        se[i] = e[i];
    }
    __syncthreads();
    const double pivmin = DBL_MIN * fmax(1.0, emax2);
    const double eps = DBL_EPSILON;
    lo -= 2.0 * eps * tnorm + pivmin;
    hi += 2.0 * eps * tnorm + pivmin;
    const double abstol = abstol_rel * tnorm;
    // two shared sweeps first: the wave's 64 lanes probe one interval that holds
    // every band of the wave, and each band takes the probes around its index
    // (a count c at x puts eigenvalues 0..c-1 below x and c.. at or above it)
    {
        const int bw0 = blockIdx.y * NBW;
        const int bw1 = min(nb, bw0 + NBW) - 1;  // last band of this wave
        double L = lo, U = hi;
        for (int sw = 0; sw < 2; ++sw) {
            const double x = L + (double)(lane + 1) * ((U - L) / 65.0);
            const int cnt = sturm_count_u(sd, se, n, x, pivmin);
            double nlo = lo, nhi = hi;
            for (int u = 0; u < 64; ++u) {
                const int cu = __shfl(cnt, u, 64);
                const double xu = __shfl(x, u, 64);
                if (cu > b) nhi = fmin(nhi, xu);
                else nlo = fmax(nlo, xu);
            }
            lo = nlo;
            hi = nhi;
            L = __shfl(lo, 0, 64);  // band bw0 sits on lane 0, band bw1 on lane (bw1 - bw0) S
            U = __shfl(hi, (bw1 - bw0) * S, 64);
        }
    }
    // multisection: lane `sub` of a band probes lo + (sub+1) (hi-lo)/(S+1)
    for (int it = 0; it < 200; ++it) {
        const double w = hi - lo;
        const bool done = !live || w <= 2.0 * eps * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin || w <= abstol;
        if (__all(done)) break;
        const double x = lo + (double)(sub + 1) * (w / (double)(S + 1));
        const int cnt = done ? 0 : sturm_count_u(sd, se, n, x, pivmin);
        // gather the S probes of this band
        double nlo = lo, nhi = hi;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int cu = __shfl(cnt, lane - sub + u, 64);
            const double xu = __shfl(x, lane - sub + u, 64);
            if (cu > b) nhi = fmin(nhi, xu);
            else nlo = fmax(nlo, xu);
        }
        if (!done) {
            if (nlo >= nhi || nlo <= lo && nhi >= hi) {  // no progress possible (rounding): stop here
                lo = hi = 0.5 * (lo + hi);
            } else {
                lo = nlo;
                hi = nhi;
            }
        }
    }
    const double lam = 0.5 * (lo + hi);
    if (live && sub == 0) evals[kk * nb + b] = lam;
    if (inv_sweeps <= 0 || !live || sub != 0) return;

    // Inverse iteration with an LU factorisation of T - lam I with partial
    // pivoting (rows k, k+1 interchanged when |e_k| > |pivot|), inv_sweeps
    // sweeps from a fixed pseudo-random start.  Each forward sweep recomputes the
    // multipliers and interchanges (the same operations, so the same values)
    // instead of reading them back, and each iterate's 1/max scaling is
    // applied when the next sweep reads it: the interleaved work arrays (over
    // the nk * nb (k, band) pairs) hold only the U factor and the iterate.
    const int64_t T = nk * nb;
    const int64_t t = kk * nb + b;
    double* __restrict__ U0 = W;                      // reciprocal pivots
    double* __restrict__ U1 = W + (size_t)n * T;      // first super-diagonal
    double* __restrict__ U2 = W + (size_t)2 * n * T;  // second super-diagonal
    double* __restrict__ X = W + (size_t)3 * n * T;   // forward result, then iterate (unscaled)
#define IDX(i) ((size_t)(i) * T + t)
    // The loads of PB rows are issued together ahead of their use (each row is
    // read before this sweep overwrites it), so a sweep waits on memory once
    // per PB rows instead of once per row.
    constexpr int PB = 4;
    const double tiny = eps * tnorm;
    double sc = 1.0;  // scaling of the iterate in X
    for (int iter = 0; iter < inv_sweeps; ++iter) {
        const bool first = iter == 0;
        // forward: factor row k, apply the interchange and the multiplier
        double a = sd[0] - lam;
        double c = n > 1 ? se[0] : 0.0;
        double yk = first ? lane_rand(0u, (uint32_t)b) : X[IDX(0)] * sc;
        for (int k0 = 0; k0 + 1 < n; k0 += PB) {
            double xn[PB];
            if (!first) {
#pragma unroll
                for (int u = 0; u < PB; ++u) xn[u] = (k0 + 1 + u < n) ? X[IDX(k0 + 1 + u)] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const int k = k0 + u;
                if (k + 1 >= n) break;
                const double sub_ = se[k];
                const double dn = sd[k + 1] - lam;
                const double en = (k + 2 < n) ? se[k + 1] : 0.0;
                double yk1 = first ? lane_rand((uint32_t)(k + 1), (uint32_t)b) : xn[u] * sc;
                double mult;
                if (fabs(a) >= fabs(sub_)) {
                    if (a == 0.0) a = tiny;
                    const double ra = rcp_nr64(a);
                    mult = sub_ * ra;
                    if (first) {
                        U0[IDX(k)] = ra;
                        U1[IDX(k)] = c;
                        U2[IDX(k)] = 0.0;
                    }
                    a = dn - mult * c;
                    c = en;
                } else {
                    const double rs = rcp_nr64(sub_);
                    mult = a * rs;
                    if (first) {
                        U0[IDX(k)] = rs;
                        U1[IDX(k)] = dn;
                        U2[IDX(k)] = en;
                    }
                    a = c - mult * dn;
                    c = -mult * en;
                    const double tmp = yk;
                    yk = yk1;
                    yk1 = tmp;
                }
                yk1 -= mult * yk;
                X[IDX(k)] = yk;
                yk = yk1;
            }
        }
        // the last pivot is the same every sweep
        if (a == 0.0) a = tiny;
        const double ra_last = rcp_nr64(a);
        if (first) U0[IDX(n - 1)] = ra_last;
        // back substitution
        double x2 = 0.0, x1 = yk * ra_last;
        X[IDX(n - 1)] = x1;
        double amax = fabs(x1);
        for (int k1 = n - 2; k1 >= 0; k1 -= PB) {
            double xv[PB], u0[PB], u1[PB], u2[PB];
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const int k = k1 - u;
                if (k < 0) break;
                xv[u] = X[IDX(k)];
                u0[u] = U0[IDX(k)];
                u1[u] = U1[IDX(k)];
                u2[u] = U2[IDX(k)];
            }
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const int k = k1 - u;
                if (k < 0) break;
                const double xk = (xv[u] - u1[u] * x1 - u2[u] * x2) * u0[u];
                X[IDX(k)] = xk;
                x2 = x1;
                x1 = xk;
                amax = fmax(amax, fabs(xk));
            }
        }
        sc = 1.0 / amax;
    }
    double nrm = 0.0;
    for (int i0 = 0; i0 < n; i0 += PB) {
        double xv[PB];
#pragma unroll
        for (int u = 0; u < PB; ++u) xv[u] = (i0 + u < n) ? X[IDX(i0 + u)] : 0.0;
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            if (i0 + u >= n) break;
            const double xi = xv[u] * sc;
            nrm += xi * xi;
        }
    }
    const double sn = 1.0 / sqrt(nrm);
    double* z = Z + (size_t)t * n;
    for (int i0 = 0; i0 < n; i0 += PB) {
        double xv[PB];
#pragma unroll
        for (int u = 0; u < PB; ++u) xv[u] = (i0 + u < n) ? X[IDX(i0 + u)] : 0.0;
#pragma unroll
        for (int u = 0; u < PB; ++u)
            if (i0 + u < n) z[i0 + u] = (xv[u] * sc) * sn;
    }
#undef IDX
}

// ------------------------------------------------------------ launchers
constexpr int kAsmBS = 256;

hipError_t launch_assemble(const DevParams& P, const int32_t* dG, const double* dk, int64_t nk, double2* dH, int ld,
                           hipStream_t s) {
    // one KState per basis vector in LDS beside the phase table (and the tiled
    // kernel's transpose slots): up to 766 vectors (860 with spin-orbit) in
    // the 160 KB of a CU
    const size_t shm = sizeof(KState) * (size_t)P.N;
    const bool tiled = !P.soc;
    const size_t fixed = sizeof(double2) * kVPhaseTab + (tiled ? sizeof(double2) * (kAsmBS / 64) * 16 * 17 : 0);
    if (shm + fixed > 163840) return hipErrorInvalidValue;
    auto kern = tiled ? assemble_tiled_kernel<kAsmBS> : assemble_kernel<kAsmBS>;
    hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(kern, dim3((unsigned)nk), dim3(kAsmBS), shm, s, P, dG, dk, dH, ld);
    return hipGetLastError();
}

hipError_t launch_tri_eig(const double* dd, const double* de, int n, int nb, int64_t nk, double* devals, double* dZ,
                          double* dW, int inv_sweeps, double abstol_rel, int* kstat, unsigned long long* status,
                          int64_t k0, hipStream_t s) {
    const size_t shm = sizeof(double) * 2 * (size_t)n;  // d and e of the k-point
    if (nb <= 16)
        hipLaunchKernelGGL(tri_eig_wave_kernel<4>, dim3((unsigned)nk, 1), dim3(64), shm, s, dd, de, n, nb, nk, devals,
                           dZ, dW, inv_sweeps, abstol_rel, kstat, status, k0);
    else if (nb <= 32)
        hipLaunchKernelGGL(tri_eig_wave_kernel<2>, dim3((unsigned)nk, 1), dim3(64), shm, s, dd, de, n, nb, nk, devals,
                           dZ, dW, inv_sweeps, abstol_rel, kstat, status, k0);
    else
        hipLaunchKernelGGL(tri_eig_wave_kernel<1>, dim3((unsigned)nk, (unsigned)((nb + 63) / 64)), dim3(64), shm, s, dd,
                           de, n, nb, nk, devals, dZ, dW, inv_sweeps, abstol_rel, kstat, status, k0);
    return hipGetLastError();
}

// status[0] = 0 failed k-points, status[1] = no first index (a kernel node, so
// that a solve captured into a graph clears it on every replay)
__global__ void status_reset_kernel(unsigned long long* status) {
    status[0] = 0ull;
    status[1] = ~0ull;
}

hipError_t launch_status_reset(unsigned long long* status, hipStream_t s) {
    hipLaunchKernelGGL(status_reset_kernel, dim3(1), dim3(1), 0, s, status);
    return hipGetLastError();
}

}  // namespace uepm
