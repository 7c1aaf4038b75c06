// solver_f64.hip — FP64 (complex128) band solve stages on gfx950.
//
//   (a) assemble_kernel     H(k) for a batch of k-points (Hamiltonian::SetMatrix +
//                           the 0.5 (H + H^dagger) step of Diagonalize,
//                           src/EPP/Hamiltonian.cpp:50-127)
//   (b) tridiag_kernel      Householder reduction H = Q T Q^H, T real symmetric
//                           tridiagonal, one workgroup per k
//   (c) tri_eig_wave_kernel bisection (Sturm counts, multisection) for the lowest
//                           nb eigenvalues of T and inverse iteration for their
//                           vectors, one wave per k
//   (d) backtransform_kernel  cluster re-orthogonalisation, x = Q z, unit norm,
//                           Hellmann-Feynman gradients (Hamiltonian.cpp:146-194)
//
// (b)-(d) replace Eigen::SelfAdjointEigenSolver::compute (Hamiltonian.cpp:129):
// same eigenvalues to within the backward-stability envelope of a Householder
// reduction, eigenvectors up to phase (and basis inside degenerate clusters).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstdlib>

#include "epm_physics.hpp"
#include "solver.hpp"

namespace uepm {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup; every thread gets the result. `scratch` holds
// BS/64 doubles; a __syncthreads() precedes reuse of `scratch`.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) s += scratch[i];
    return s;
}

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

// ---------------------------------------------------- (b) tridiagonalisation
// Unblocked Householder on the full Hermitian matrix held column-major in
// device memory.  Step j: reflector H_j = I - tau v v^H annihilating
// A[j+2:, j]; p = tau A22 v; w = p - (tau/2)(p^H v) v; A22 -= v w^H + w v^H.
// v (v[0] = 1 implied) overwrites A[j+2:, j]; outputs d, e (real), tau.
template <int BS>
__global__ __launch_bounds__(BS) void tridiag_kernel(double2* __restrict__ H, int n, int ld, double* __restrict__ dout,
                                                     double* __restrict__ eout, double2* __restrict__ tout) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* v = reinterpret_cast<double2*>(smem);
    double2* w = v + n;
    double* red = reinterpret_cast<double*>(w + n);
    __shared__ double2 s_tau;
    const int kk = blockIdx.x;
    double2* A = H + (size_t)kk * ld * n;
    double* d = dout + (size_t)kk * n;
    double* e = eout + (size_t)kk * n;
    double2* tau = tout + (size_t)kk * n;
    const int tid = threadIdx.x;

    for (int j = 0; j + 1 < n; ++j) {
        double2* colj = A + (size_t)j * ld;
        const double2 alpha = colj[j + 1];  // read before the barriers in block_sum
        double part = 0.0;
        for (int r = j + 2 + tid; r < n; r += BS) {
            const double2 x = colj[r];
            part += x.x * x.x + x.y * x.y;
        }
        const double xn2 = block_sum<BS>(part, red);
        double2 tj = make_double2(0.0, 0.0);
        double beta = alpha.x;
        double2 sc = make_double2(1.0, 0.0);
        if (!(xn2 == 0.0 && alpha.y == 0.0)) {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            beta = alpha.x >= 0 ? -nrm : nrm;
            tj = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            // sc = 1 / (alpha - beta)
            const double ar = alpha.x - beta, ai = alpha.y;
            const double den = ar * ar + ai * ai;
            sc = make_double2(ar / den, -ai / den);
        }
        for (int r = j + 2 + tid; r < n; r += BS) {
            const double2 x = cmul(colj[r], sc);
            colj[r] = x;
            v[r] = x;
        }
        if (tid == 0) {
            v[j + 1] = make_double2(1.0, 0.0);
            colj[j + 1] = make_double2(beta, 0.0);
            e[j] = beta;
            tau[j] = tj;
            d[j] = A[j + (size_t)j * ld].x;
        }
        __syncthreads();
        if (tj.x == 0.0 && tj.y == 0.0) continue;
        // p = tau * A22 v
        for (int r = j + 1 + tid; r < n; r += BS) {
            double2 acc = make_double2(0.0, 0.0);
            for (int c = j + 1; c < n; ++c) {
                const double2 a = A[r + (size_t)c * ld];
                const double2 vc = v[c];
                acc.x += a.x * vc.x - a.y * vc.y;
                acc.y += a.x * vc.y + a.y * vc.x;
            }
            w[r] = cmul(tj, acc);
        }
        __syncthreads();
        double pr = 0.0, pi = 0.0;
        for (int r = j + 1 + tid; r < n; r += BS) {
            const double2 q = cmulc(w[r], v[r]);
            pr += q.x;
            pi += q.y;
        }
        pr = block_sum<BS>(pr, red);
        __syncthreads();
        pi = block_sum<BS>(pi, red);
        // al = -0.5 * tau * (p^H v)
        const double2 al = cmul(make_double2(-0.5 * tj.x, -0.5 * tj.y), make_double2(pr, pi));
        for (int r = j + 1 + tid; r < n; r += BS) w[r] = cadd(w[r], cmul(al, v[r]));
        __syncthreads();
        // A22 -= v w^H + w v^H
        for (int r = j + 1 + tid; r < n; r += BS) {
            const double2 vr = v[r], wr = w[r];
            for (int c = j + 1; c < n; ++c) {
                const double2 vc = v[c], wc = w[c];
                double2 a = A[r + (size_t)c * ld];
                // v_r conj(w_c) + w_r conj(v_c)
                a.x -= (vr.x * wc.x + vr.y * wc.y) + (wr.x * vc.x + wr.y * vc.y);
                a.y -= (vr.y * wc.x - vr.x * wc.y) + (wr.y * vc.x - wr.x * vc.y);
                A[r + (size_t)c * ld] = a;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[n - 1] = A[(n - 1) + (size_t)(n - 1) * ld].x;
        e[n - 1] = 0.0;
        tau[n - 1] = make_double2(0.0, 0.0);
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

// Wave-per-k variant: the d, e of a k-point are wave-uniform (scalar loads,
// SGPR operands); lanes = bands x S shifts (multisection: one sweep narrows the
// bracket of every band S+1-fold); inverse iteration on the first lane of a band.
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
                                                          double abstol_rel) {
    constexpr int NBW = 64 / S;  // bands per wave
    const int64_t kk = blockIdx.x;
    const int lane = threadIdx.x, sub = lane % S;
    const int b = blockIdx.y * NBW + lane / S;
    const bool live = b < nb;
    const double* d = dall + kk * n;
    const double* e = eall + kk * n;
    // Gershgorin interval and scale (wave-uniform)
    double lo = DBL_MAX, hi = -DBL_MAX, emax2 = 0.0, tnorm = 0.0;
    for (int i = 0; i < n; ++i) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double er = i + 1 < n ? fabs(e[i]) : 0.0;
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        emax2 = fmax(emax2, er * er);
        tnorm = fmax(tnorm, fabs(d[i]) + el + er);
    }
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
            const int cnt = sturm_count_u(d, e, n, x, pivmin);
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
        const int cnt = done ? 0 : sturm_count_u(d, e, n, x, pivmin);
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
        double a = d[0] - lam;
        double c = n > 1 ? e[0] : 0.0;
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
                const double sub_ = e[k];
                const double dn = d[k + 1] - lam;
                const double en = (k + 2 < n) ? e[k + 1] : 0.0;
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

// ------------------------------------------- (d) back-transform + gradients
// One workgroup per k, 4 waves; wave w owns bands b = w, w+4, ...
// Z: nk x nb x n real eigenvectors of T (unit norm).  X (output, complex):
// nk x nb x dim.  Bands within ortol of each other are re-orthogonalised
// (modified Gram-Schmidt, ascending) before x = Q z.
__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2((double)a.x, (double)a.y); }

// RT: reflector storage type (double2 for the FP64 path, float2 for the
// complex64 reduction of the mixed-precision path); arithmetic is FP64.
template <int BS, typename RT>
__global__ __launch_bounds__(BS) void backtransform_kernel(const RT* __restrict__ H, int n, int ld,
                                                           const RT* __restrict__ tauall,
                                                           const double* __restrict__ dall,
                                                           const double* __restrict__ eall,
                                                           const double* __restrict__ evals, const double* Z,
                                                           double2* __restrict__ X, double* __restrict__ grads,
                                                           DevParams P, const double* __restrict__ kpts,
                                                           const int32_t* __restrict__ G, int nb) {
    const int kk = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    constexpr int NW = BS / 64;
    const RT* A = H + (size_t)kk * ld * n;
    const RT* tau = tauall + (size_t)kk * n;
    const double* z = Z + (size_t)kk * nb * n;
    double2* x = X + (size_t)kk * nb * n;
    const double* ev = evals + (size_t)kk * nb;
    // ||T||_1 for the cluster tolerance
    double tn = 0.0;
    {
        const double* d = dall + (size_t)kk * n;
        const double* e = eall + (size_t)kk * n;
        for (int i = lane; i < n; i += 64) {
            const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
            const double er = i + 1 < n ? fabs(e[i]) : 0.0;
            tn = fmax(tn, fabs(d[i]) + el + er);
        }
        for (int o = 32; o > 0; o >>= 1) tn = fmax(tn, __shfl_xor(tn, o, 64));
    }
    const double ortol = 1e-3 * tn;
    // Copy z -> x (real), MGS within clusters.  Serial over bands in wave 0.
    if (wv == 0) {
        int cstart = 0;
        for (int b = 0; b < nb; ++b) {
            if (b > 0 && ev[b] - ev[b - 1] > ortol) cstart = b;
            double2* xb = x + (size_t)b * n;
            for (int i = lane; i < n; i += 64) xb[i] = make_double2(z[(size_t)b * n + i], 0.0);
            for (int c = cstart; c < b; ++c) {
                const double2* xc = x + (size_t)c * n;
                double s = 0.0;
                for (int i = lane; i < n; i += 64) s += xc[i].x * xb[i].x;
                s = wave_sum(s);
                for (int i = lane; i < n; i += 64) xb[i].x -= s * xc[i].x;
            }
            double nn = 0.0;
            for (int i = lane; i < n; i += 64) nn += xb[i].x * xb[i].x;
            nn = wave_sum(nn);
            const double inv = 1.0 / sqrt(nn);
            for (int i = lane; i < n; i += 64) xb[i].x *= inv;
        }
    }
    __syncthreads();
    // x = H_0 H_1 ... H_{n-2} z: apply H_{n-2} first.
    for (int b = wv; b < nb; b += NW) {
        double2* xb = x + (size_t)b * n;
        for (int j = n - 2; j >= 0; --j) {
            const double2 tj = to_d2(tau[j]);
            if (tj.x == 0.0 && tj.y == 0.0) continue;
            const RT* vj = A + (size_t)j * ld;  // v at rows j+1..n-1, v[j+1] = 1
            const int r0 = lane + (j + 1 > lane ? ((j + 1 - lane + 63) >> 6) << 6 : 0);
            double sr = 0.0, si = 0.0;
            for (int r = r0; r < n; r += 64) {
                const double2 vr = (r == j + 1) ? make_double2(1.0, 0.0) : to_d2(vj[r]);
                const double2 q = cmulc(vr, xb[r]);
                sr += q.x;
                si += q.y;
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            const double2 s = cmul(tj, make_double2(sr, si));
            for (int r = r0; r < n; r += 64) {
                const double2 vr = (r == j + 1) ? make_double2(1.0, 0.0) : to_d2(vj[r]);
                xb[r] = csub(xb[r], cmul(s, vr));
            }
        }
        __threadfence_block();  // rows i + N were written by other lanes
        if (grads) {
            // Hellmann-Feynman on the kinetic term (Hamiltonian.cpp:146-194)
            const double* k = kpts + 3 * (size_t)kk;
            const int N = P.N;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int i = lane; i < N; i += 64) {
                double wt = xb[i].x * xb[i].x + xb[i].y * xb[i].y;
                if (P.soc) wt += xb[i + N].x * xb[i + N].x + xb[i + N].y * xb[i + N].y;
                sx += (k[0] + (double)G[3 * i]) * wt;
                sy += (k[1] + (double)G[3 * i + 1]) * wt;
                sz += (k[2] + (double)G[3 * i + 2]) * wt;
            }
            sx = wave_sum(sx);
            sy = wave_sum(sy);
            sz = wave_sum(sz);
            if (lane == 0) {
                double* g = grads + 3 * ((size_t)kk * nb + b);
                g[0] = sx * P.grad_pref * P.frac_to_phys;
                g[1] = sy * P.grad_pref * P.frac_to_phys;
                g[2] = sz * P.grad_pref * P.frac_to_phys;
            }
        }
    }
}

// ------------------------------------------------------------ launchers
constexpr int kAsmBS = 256;
// UEPM_ASM_V1=1 keeps the one-element-per-thread assembly for every plan (A/B)
static bool assemble_v1() {
    static const bool on = [] {
        const char* v = std::getenv("UEPM_ASM_V1");
        return v && v[0] == '1';
    }();
    return on;
}
constexpr int kTriBS = 256;
constexpr int kBtBS = 256;

hipError_t launch_assemble(const DevParams& P, const int32_t* dG, const double* dk, int64_t nk, double2* dH, int ld,
                           hipStream_t s) {
    // one KState per basis vector in LDS beside the phase table (and the tiled
    // kernel's transpose slots): up to 766 vectors (860 with spin-orbit) in
    // the 160 KB of a CU
    const size_t shm = sizeof(KState) * (size_t)P.N;
    const bool tiled = !P.soc && !assemble_v1();
    const size_t fixed = sizeof(double2) * kVPhaseTab + (tiled ? sizeof(double2) * (kAsmBS / 64) * 16 * 17 : 0);
    if (shm + fixed > 163840) return hipErrorInvalidValue;
    auto kern = tiled ? assemble_tiled_kernel<kAsmBS> : assemble_kernel<kAsmBS>;
    hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(kern, dim3((unsigned)nk), dim3(kAsmBS), shm, s, P, dG, dk, dH, ld);
    return hipGetLastError();
}

hipError_t launch_tridiag(double2* dH, int n, int ld, int64_t nk, double* dd, double* de, double2* dtau,
                          hipStream_t s) {
    const size_t shm = sizeof(double2) * 2 * (size_t)n + sizeof(double) * (kTriBS / 64);
    hipLaunchKernelGGL(tridiag_kernel<kTriBS>, dim3((unsigned)nk), dim3(kTriBS), shm, s, dH, n, ld, dd, de, dtau);
    return hipGetLastError();
}

hipError_t launch_tri_eig(const double* dd, const double* de, int n, int nb, int64_t nk, double* devals, double* dZ,
                          double* dW, int inv_sweeps, double abstol_rel, hipStream_t s) {
    if (nb <= 16)
        hipLaunchKernelGGL(tri_eig_wave_kernel<4>, dim3((unsigned)nk, 1), dim3(64), 0, s, dd, de, n, nb, nk, devals,
                           dZ, dW, inv_sweeps, abstol_rel);
    else if (nb <= 32)
        hipLaunchKernelGGL(tri_eig_wave_kernel<2>, dim3((unsigned)nk, 1), dim3(64), 0, s, dd, de, n, nb, nk, devals,
                           dZ, dW, inv_sweeps, abstol_rel);
    else
        hipLaunchKernelGGL(tri_eig_wave_kernel<1>, dim3((unsigned)nk, (unsigned)((nb + 63) / 64)), dim3(64), 0, s, dd,
                           de, n, nb, nk, devals, dZ, dW, inv_sweeps, abstol_rel);
    return hipGetLastError();
}

hipError_t launch_backtransform(const double2* dH, int n, int ld, const double2* dtau, const double* dd,
                                const double* de, const double* devals, const double* dZ, double2* dX, double* dgrads,
                                const DevParams& P, const double* dk, const int32_t* dG, int nb, int64_t nk,
                                hipStream_t s) {
    hipLaunchKernelGGL((backtransform_kernel<kBtBS, double2>), dim3((unsigned)nk), dim3(kBtBS), 0, s, dH, n, ld, dtau,
                       dd, de, devals, dZ, dX, dgrads, P, dk, dG, nb);
    return hipGetLastError();
}

hipError_t launch_backtransform_c64(const float2* dV, int n, int ld, const float2* dtau, const double* dd,
                                    const double* de, const double* devals, const double* dZ, double2* dX,
                                    const DevParams& P, const double* dk, const int32_t* dG, int nb, int64_t nk,
                                    hipStream_t s) {
    hipLaunchKernelGGL((backtransform_kernel<kBtBS, float2>), dim3((unsigned)nk), dim3(kBtBS), 0, s, dV, n, ld, dtau,
                       dd, de, devals, dZ, dX, nullptr, P, dk, dG, nb);
    return hipGetLastError();
}

}  // namespace uepm
