// tridiag_blocked.hip — batched blocked Householder reduction of Hermitian H(k)
// to real symmetric tridiagonal form, one workgroup per k-point, on gfx950.
//
// Storage.  The lower triangle of H is packed by columns.  Column c holds rows
// c..n-1 and is padded to a length congruent to 2 modulo the number of
// elements per 128-byte LDS bank row, so that both walks of the symmetric
// matrix-vector product are bank-conflict free: lanes = consecutive rows r
// reading A(r, c) at a fixed c (row walk), and lanes = consecutive rows r
// reading A(c, r) at a fixed c (column walk; the start offset of column r
// advances by 2 per column, the in-column offset c - r falls by 1).
// The trailing columns [c0, n) live in LDS; the leading columns [0, c0) that
// do not fit live in a per-workgroup global scratch region (hundreds of KB at
// the large dims, beyond a workgroup's share of the L2, so each step streams
// its live part from memory) and are touched only during the first c0 steps.
//
// Algorithm (LAPACK xHETRD/xLATRD lower, restated).  Columns are reduced in
// panels of NBP.  Inside a panel the trailing matrix is read-only: step j
// applies the panel's pending rank-2 updates to column j only, forms the
// reflector H_j = I - tau v v^H, and computes
//   w = tau (A v - V (W^H v) - W (V^H v)),  w -= (tau/2)(w^H v) v
// with A the matrix as of the panel start.  After the panel, the trailing
// lower triangle takes the rank-2*NBP update A -= V W^H + W V^H once.
// Results: d (diagonal), e (sub-diagonal, real), tau, and the reflectors
// v_j (v_j[j+1] = 1 implied) written below the sub-diagonal of column j of Vout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "solver.hpp"
#include "uepm.h"
#include "wave_ops.hpp"

namespace uepm {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};
template <typename T>
__device__ __forceinline__ cx<T> cxm(T a, T b) {
    return {a, b};
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
    return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) {
    return {a.x, -a.y};
}
// acc += a * b
template <typename T>
__device__ __forceinline__ void cfma(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// acc += conj(a) * b
template <typename T>
__device__ __forceinline__ void cfmac(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
}
// acc -= a * conj(b)
template <typename T>
__device__ __forceinline__ void cfms_c(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.x = fma(-a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(-a.y, b.x, fma(a.x, b.y, acc.y));
}

// wave-wide sum, every lane gets it (DPP and row swaps: no LDS crossbar trips)
template <typename T>
__device__ __forceinline__ T wsum(T v) {
    return wave::sum(v);
}

// 1/x from the hardware reciprocal and two Newton steps (full double precision).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// Column padding: length congruent to 2 modulo MOD (elements per bank row).
__host__ __device__ inline int padded_len(int L, int MOD) { return L + (((2 - L) % MOD) + MOD) % MOD; }

// panel width: 8 in complex64, 4 in complex128 (whose panel columns V, W at
// width 8 would not leave the LDS room at the largest dims)
template <typename T>
constexpr int tri_nbp() { return sizeof(T) == 4 ? 8 : 4; }
constexpr int kTriParts = 4;
constexpr int kTriMaxRowsPerLane = 9;  // rows per lane in the symmetric product: n <= 576
static_assert(64 * kTriMaxRowsPerLane >= UEPM_MAX_DIM, "every plan dimension fits the symmetric product");

template <typename T>
struct TriSmem {
    // bytes of the fixed part of LDS (everything but the matrix region)
    static __host__ __device__ int fixed_bytes(int n, int BS, int NBP) {
        const int E = (int)sizeof(cx<T>);
        int b = ((n * 4 + 15) / 16) * 16;  // colofs
        b += 2 * NBP * n * E;              // panel V and W columns
        b += 2 * n * E;                    // y (row sums / w), ycol (column sums of global rows)
        b += kTriParts * n * E;            // partial row sums of the symv (waves w, w + 4 share one)
        b += 2 * NBP * E;                  // dots a_q, b_q
        b += NBP * E;                      // alpha of the panel's steps
        b += 16 * 8;                       // scalars
        b += 4 * (BS / 64) * 8;            // block reductions
        return ((b + 15) / 16) * 16;
    }
};

// MFMA 16x16x4: A 16x4 (lane l: row l%16, k l/16), B 4x16 (lane l: k l/16,
// col l%16).  C/D 16x16, lane l item v at column l%16 and row
//   f32: 4*(l/16) + v        f64: 4*v + l/16
// (both measured on gfx950 with a known-product probe).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lq, int v) { return 4 * v + lq; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lq, int v) { return 4 * lq + v; }
};

template <typename T, int BS, int NBP>
__global__ __launch_bounds__(BS, 1) void tridiag_blocked_kernel(const double2* __restrict__ Hin, int n, int ld,
                                                                 int c0, cx<T>* __restrict__ scratch,
                                                                 int64_t scratch_stride, cx<T>* __restrict__ Vout,
                                                                 double* __restrict__ dout, double* __restrict__ eout,
                                                                 cx<T>* __restrict__ tauout, int s_end) {
    constexpr int MOD = 128 / (int)sizeof(cx<T>);
    constexpr int NW = BS / 64;
    constexpr int U = NBP / 2;  // K = 4 NBP real = 4 U per lane group
    static_assert(NBP % 2 == 0, "NBP must be even");
    using M = Mfma<T>;
    using v4 = typename M::v4;
    extern __shared__ __align__(16) unsigned char smem[];
    int* colofs = reinterpret_cast<int*>(smem);
    unsigned char* p = smem + ((n * 4 + 15) / 16) * 16;
    cx<T>* Vst = reinterpret_cast<cx<T>*>(p);  // Vst[q * n + r] = v_{p0+q}[r]
    p += NBP * n * sizeof(cx<T>);
    cx<T>* Wl = reinterpret_cast<cx<T>*>(p);   // Wl[q * n + r]
    p += NBP * n * sizeof(cx<T>);
    cx<T>* ys = reinterpret_cast<cx<T>*>(p);
    p += n * sizeof(cx<T>);
    cx<T>* ycol = reinterpret_cast<cx<T>*>(p);
    p += n * sizeof(cx<T>);
    cx<T>* Yp = reinterpret_cast<cx<T>*>(p);  // Yp[part * n + r]
    p += kTriParts * n * sizeof(cx<T>);
    cx<T>* dots = reinterpret_cast<cx<T>*>(p);
    p += 2 * NBP * sizeof(cx<T>);
    cx<T>* alph = reinterpret_cast<cx<T>*>(p);
    p += NBP * sizeof(cx<T>);
    double* scal = reinterpret_cast<double*>(p);
    p += 16 * 8;
    double* red = reinterpret_cast<double*>(p);
    cx<T>* lmat = reinterpret_cast<cx<T>*>(smem + TriSmem<T>::fixed_bytes(n, BS, NBP));

    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int64_t kk = blockIdx.x;
    cx<T>* gmat = scratch + kk * scratch_stride;
    if (tid == 0) {
        int off = 0;
        for (int c = 0; c < n; ++c) {
            if (c == c0) off = 0;
            colofs[c] = off;
            off += padded_len(n - c, MOD);
        }
    }
    __syncthreads();

    // element (r, c), r >= c: global scratch for c < c0, LDS otherwise
    auto gcol = [&](int c) -> cx<T>* { return gmat + (colofs[c] - c); };
    auto lcol = [&](int c) -> cx<T>* { return lmat + (colofs[c] - c); };
    auto load = [&](int r, int c) -> cx<T> { return c < c0 ? gcol(c)[r] : lcol(c)[r]; };
    auto store = [&](int r, int c, cx<T> a) {
        if (c < c0) gcol(c)[r] = a;
        else lcol(c)[r] = a;
    };
    const cx<T> zero = cxm<T>(0, 0);

    const double2* Hk = Hin + kk * (int64_t)ld * n;
    for (int c = 0; c < n; ++c) {
        if (c < c0) {
            cx<T>* g = gcol(c);
            for (int r = c + tid; r < n; r += BS) {
                const double2 h = Hk[r + (int64_t)c * ld];
                g[r] = cxm<T>((T)h.x, (T)h.y);
            }
        } else {
            cx<T>* l = lcol(c);
            for (int r = c + tid; r < n; r += BS) {
                const double2 h = Hk[r + (int64_t)c * ld];
                l[r] = cxm<T>((T)h.x, (T)h.y);
            }
        }
    }
    double* d = dout + kk * n;
    double* e = eout + kk * n;
    cx<T>* tau = tauout + kk * n;
    __syncthreads();

    // steps s < p_stop run here; s_end < n - 1 (a multiple of NBP) stops at a panel
    // boundary, where the trailing matrix is fully updated, and hands it over
    const int p_stop = min(n - 1, s_end);
    for (int p0 = 0; p0 < p_stop; p0 += NBP) {
        const int nbp = min(NBP, p_stop - p0);
        double tpr = 0.0, tpi = 0.0;  // tau of the previous step in this panel
        cx<T> nxt = zero;             // this thread's row of the next column (raw), prefetched
        bool have_nxt = false;
        for (int jj = 0; jj < nbp; ++jj) {
            const int j = p0 + jj;
            const int m = n - j - 1;
            cx<T>* vj = Vst + jj * n;
            // alpha of the previous step: alpha = -(tau/2)(w^H v), w = Wl + alpha v (lazy)
            cx<T> acur = zero;
            if (jj > 0) {
                double sr2 = 0.0, si2 = 0.0;
#pragma unroll
                for (int i = 0; i < NW; ++i) {
                    sr2 += red[2 * NW + i];
                    si2 += red[3 * NW + i];
                }
                acur = cxm<T>((T)(-0.5 * (tpr * sr2 - tpi * si2)), (T)(-0.5 * (tpr * si2 + tpi * sr2)));
                if (tid == 0) alph[jj - 1] = acur;
            }
            // S1: column j with the panel's pending updates, one row per thread -> ys
            for (int r = j + tid, it = 0; r < n; r += BS, ++it) {
                cx<T> a = (it == 0 && have_nxt) ? nxt : load(r, j);
                for (int q = 0; q < jj; ++q) {
                    const cx<T> aq = (q == jj - 1) ? acur : alph[q];
                    const cx<T> wr = Wl[q * n + r] + aq * Vst[q * n + r];
                    const cx<T> wjq = Wl[q * n + j] + aq * Vst[q * n + j];
                    cfms_c(a, Vst[q * n + r], wjq);
                    cfms_c(a, wr, Vst[q * n + j]);
                }
                ys[r] = a;
            }
            __syncthreads();
            // S2, redundantly in every wave: norm, reflector H_j = I - tau v v^H
            // (v[j+1] = 1); every wave writes the same v.
            double part = 0.0;
            for (int r = j + 2 + lane; r < n; r += 64) {
                const cx<T> a = ys[r];
                part += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
            }
            const double xn2 = wsum(part);
            const double djj = (double)ys[j].x;
            const double ar = (double)ys[j + 1].x, ai = (double)ys[j + 1].y;
            double beta = ar, tr = 0.0, ti = 0.0, sr = 1.0, si = 0.0;
            if (!(xn2 == 0.0 && ai == 0.0)) {
                const double nrm = sqrt(ar * ar + ai * ai + xn2);
                beta = ar >= 0 ? -nrm : nrm;
                const double ib = rcp_nr(beta);
                tr = (beta - ar) * ib;
                ti = -ai * ib;
                const double br = ar - beta, iden = rcp_nr(br * br + ai * ai);
                sr = br * iden;
                si = -ai * iden;
            }
            const cx<T> sc = cxm<T>((T)sr, (T)si);
            const cx<T> tj = cxm<T>((T)tr, (T)ti);
            for (int r = lane; r < n; r += 64) {
                cx<T> x = zero;
                if (r == j + 1) x = cxm<T>(1, 0);
                else if (r > j + 1) x = ys[r] * sc;
                vj[r] = x;
            }
            if (tid == 0) {
                d[j] = djj;
                e[j] = beta;
                tau[j] = tj;
            }
            // S3: y = A v.  Every wave keeps partial row sums for rows r = 64 q + lane
            // in yq[q] (a lane's rows are fixed, whatever the column).
            //  G. Global columns c in [j+1, c0), one wave per column, round robin: the
            //     coalesced column A(r, c), r >= c, feeds the row part y_r += A(r, c) v_c,
            //     and its column part y_c = sum_{r > c} conj(A(r, c)) v_r is one wave
            //     reduction (every element right of the diagonal in row c sits in
            //     column c, so ycol[c] is complete).  Loads run PF - 1 columns ahead
            //     (three in complex64, one in complex128, as registers allow).
            //  L. Rows r >= rl0 = max(c0, j+1) against the LDS columns [rl0, n), both
            //     triangles (A(r, c), or conj(A(c, r)) above the diagonal), in items of
            //     one 64-row block and one column part; nq * TL items, a multiple of NW.
            //  The wave partials go to Yp: waves w < NP write Yp[w], then waves w >= NP
            //  add theirs to Yp[w - NP] (a fixed order, so the sums are reproducible).
            {
                constexpr int NQ = kTriMaxRowsPerLane;
                constexpr int PF = sizeof(T) == 4 ? 4 : 2;
                const int cg1 = min(c0, n);
                const int rl0 = max(c0, j + 1);
                // v at this lane's rows: kept in registers in complex64, re-read from LDS
                // in complex128 (whose load buffers take the registers)
                constexpr bool VREG = sizeof(T) == 4;
                cx<T> yq[NQ], vq[VREG ? NQ : 1];
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    yq[q] = zero;
                    const int r = 64 * q + lane;
                    if constexpr (VREG) vq[q] = (64 * q < n && r < n) ? vj[r] : zero;
                }
                // ---- G ----
                auto gload = [&](int c, cx<T>(&b)[NQ]) {
                    const cx<T>* col = gcol(c);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int r = 64 * q + lane;
                        b[q] = zero;
                        if (64 * q + 63 >= c && 64 * q < n)  // wave-uniform
                            if (r >= c && r < n) b[q] = col[r];
                    }
                };
                auto gdo = [&](int c, const cx<T>(&b)[NQ]) {
                    const cx<T> vc = vj[c];
                    cx<T> cs = zero;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        if (64 * q + 63 >= c && 64 * q < n) {
                            const int r = 64 * q + lane;
                            cfma(yq[q], b[q], vc);  // b is zero outside rows [c, n)
                            cfmac(cs, r > c ? b[q] : zero, VREG ? vq[VREG ? q : 0] : vj[min(r, n - 1)]);
                        }
                    }
                    cs.x = wsum(cs.x);
                    cs.y = wsum(cs.y);
                    if (lane == 0) ycol[c] = cs;
                };
                {
                    cx<T> bb[PF][NQ];
#pragma unroll
                    for (int p = 0; p + 1 < PF; ++p) {
                        const int cc = j + 1 + wv + p * NW;
                        if (cc < cg1) gload(cc, bb[p]);
                    }
                    for (int c = j + 1 + wv; c < cg1; c += PF * NW) {
#pragma unroll
                        for (int p = 0; p < PF; ++p) {
                            const int cc = c + p * NW;
                            const int cn = cc + (PF - 1) * NW;  // into the slot freed last
                            if (cn < cg1) gload(cn, bb[(p + PF - 1) % PF]);
                            if (cc < cg1) gdo(cc, bb[p]);
                        }
                    }
                }
                // ---- L ----
                if (rl0 < n) {
                    const int qa = rl0 >> 6, nq = ((n + 63) >> 6) - qa;
                    int TL = NW;
                    while (TL > 1 && (nq * (TL >> 1)) % NW == 0) TL >>= 1;
                    const int m = n - rl0;
                    for (int it = wv; it < nq * TL; it += NW) {
                        const int q = qa + it % nq, pp = it / nq;
                        const int rr = 64 * q + lane;
                        const bool live = rr >= rl0 && rr < n;
                        const int r = min(max(rr, rl0), n - 1);
                        const int cs = rl0 + (pp * m) / TL, ce = rl0 + ((pp + 1) * m) / TL;
                        cx<T> acc0 = zero, acc1 = zero;
                        if (cs < ce) {
                            int off = colofs[cs] - cs;
                            int P = padded_len(n - cs, MOD);
                            const int bcol = colofs[r] - r;
                            auto fetch = [&](int c, int o) -> cx<T> {
                                const bool in = c <= r;
                                cx<T> a = lmat[in ? o + r : bcol + c];
                                a.y = in ? a.y : -a.y;
                                return a;
                            };
                            int c = cs;
                            if (c + 3 < ce) {
                                int o0 = off, o1, o2, o3;
                                o1 = o0 + P - 1; P = padded_len(n - c - 1, MOD);
                                o2 = o1 + P - 1; P = padded_len(n - c - 2, MOD);
                                o3 = o2 + P - 1; P = padded_len(n - c - 3, MOD);
                                off = o3 + P - 1; P = padded_len(n - c - 4, MOD);
                                cx<T> a0 = fetch(c, o0), a1 = fetch(c + 1, o1), a2 = fetch(c + 2, o2), a3 = fetch(c + 3, o3);
                                cx<T> w0 = vj[c], w1 = vj[c + 1], w2 = vj[c + 2], w3 = vj[c + 3];
                                for (; c + 7 < ce; c += 4) {
                                    const int q0 = off;
                                    const int q1 = q0 + P - 1; P = padded_len(n - c - 5, MOD);
                                    const int q2 = q1 + P - 1; P = padded_len(n - c - 6, MOD);
                                    const int q3 = q2 + P - 1; P = padded_len(n - c - 7, MOD);
                                    off = q3 + P - 1; P = padded_len(n - c - 8, MOD);
                                    const cx<T> b0 = fetch(c + 4, q0), b1 = fetch(c + 5, q1), b2 = fetch(c + 6, q2),
                                                b3 = fetch(c + 7, q3);
                                    const cx<T> u0 = vj[c + 4], u1 = vj[c + 5], u2 = vj[c + 6], u3 = vj[c + 7];
                                    cfma(acc0, a0, w0);
                                    cfma(acc1, a1, w1);
                                    cfma(acc0, a2, w2);
                                    cfma(acc1, a3, w3);
                                    a0 = b0; a1 = b1; a2 = b2; a3 = b3;
                                    w0 = u0; w1 = u1; w2 = u2; w3 = u3;
                                }
                                cfma(acc0, a0, w0);
                                cfma(acc1, a1, w1);
                                cfma(acc0, a2, w2);
                                cfma(acc1, a3, w3);
                                c += 4;
                            }
                            for (; c < ce; ++c) {
                                cfma(acc0, fetch(c, off), vj[c]);
                                off += P - 1;
                                P = padded_len(n - c - 1, MOD);
                            }
                        }
                        const cx<T> acc = live ? acc0 + acc1 : zero;
#pragma unroll
                        for (int qq = 0; qq < NQ; ++qq)
                            if (qq == q) yq[qq] = yq[qq] + acc;  // wave-uniform
                    }
                }
                // ---- wave partials to Yp ----
                constexpr int NP = NW < kTriParts ? NW : kTriParts;
                static_assert(NW <= 2 * kTriParts, "two rounds of partial sums");
                if (wv < NP) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int r = 64 * q + lane;
                        if (64 * q < n && r < n) Yp[wv * n + r] = yq[q];
                    }
                }
                if constexpr (NW > NP) {
                    __syncthreads();
                    if (wv >= NP) {
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            const int r = 64 * q + lane;
                            if (64 * q < n && r < n) Yp[(wv - NP) * n + r] = Yp[(wv - NP) * n + r] + yq[q];
                        }
                    }
                }
            }
            // S3c: a_q = sum_r conj(W(r,q)) v_r, b_q = sum_r conj(V(r,q)) v_r, r in [j+1, n)
            for (int q = wv; q < 2 * jj; q += NW) {
                const int qq = q < jj ? q : q - jj;
                const cx<T>* u = (q < jj ? Wl : Vst) + qq * n;
                cx<T> acc = zero;
                for (int r = j + 1 + lane; r < n; r += 64) cfmac(acc, u[r], vj[r]);
                acc.x = wsum(acc.x);
                acc.y = wsum(acc.y);
                if (lane == 0) dots[q < jj ? qq : NBP + qq] = acc;
            }
            __syncthreads();
            // Phase C: v into the matrix storage (all waves are past their reads of
            // column j); y = A v - V a - W b with W = Wl + V diag(alpha);
            // Wl[:, jj] = tau y; partial sums of w^H v for the next step's alpha.
            for (int r = j + 2 + tid; r < n; r += BS) store(r, j, vj[r]);
            // next column of this panel: its storage holds the panel-start value until its own step
            have_nxt = jj + 1 < nbp;
            if (have_nxt && j + 1 + tid < n) nxt = load(j + 1 + tid, j + 1);
            double pvr = 0.0, pvi = 0.0;
            for (int r = tid; r < n; r += BS) {
                cx<T> w = zero;
                if (r > j) {
                    cx<T> y = Yp[r];
#pragma unroll
                    for (int pp = 1; pp < (NW < kTriParts ? NW : kTriParts); ++pp) y = y + Yp[pp * n + r];
                    if (r < c0) y = y + ycol[r];
                    for (int q = 0; q < jj; ++q) {
                        const cx<T> aq = alph[q], bq = dots[NBP + q];
                        const cx<T> cv = dots[q] + cxm<T>(2 * aq.x, 0) * bq;
                        y = y - Vst[q * n + r] * cv;
                        y = y - Wl[q * n + r] * bq;
                    }
                    w = tj * y;
                    const cx<T> vr = vj[r];
                    pvr += (double)w.x * (double)vr.x + (double)w.y * (double)vr.y;
                    pvi += (double)w.x * (double)vr.y - (double)w.y * (double)vr.x;
                }
                Wl[jj * n + r] = w;
            }
            pvr = wsum(pvr);
            pvi = wsum(pvi);
            if (lane == 0) {
                red[2 * NW + wv] = pvr;
                red[3 * NW + wv] = pvi;
            }
            tpr = tr;
            tpi = ti;
            __syncthreads();
        }
        // alpha of the panel's last step, then W = Wl + V diag(alpha) explicitly
        {
            double sr2 = 0.0, si2 = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                sr2 += red[2 * NW + i];
                si2 += red[3 * NW + i];
            }
            const cx<T> al_last = cxm<T>((T)(-0.5 * (tpr * sr2 - tpi * si2)), (T)(-0.5 * (tpr * si2 + tpi * sr2)));
            for (int r = tid; r < n; r += BS)
                for (int q = 0; q < nbp; ++q) {
                    const cx<T> aq = (q == nbp - 1) ? al_last : alph[q];
                    Wl[q * n + r] = Wl[q * n + r] + aq * Vst[q * n + r];
                }
        }
        __syncthreads();
        // Trailing update of the lower triangle of A[p1:, p1:] on MFMA:
        //   A -= L R^H, L = [V | W], R = [W | V]  (2 NBP complex = 4 NBP real K)
        //   Re: [Lr Li][Rr Ri]^T,  Im: [Li -Lr][Rr Ri]^T
        // Lane group lq covers complex K indices q = lq + 4u, u < U.
        const int p1 = p0 + nbp;
        const int m2 = n - p1;
        if (m2 > 0) {
            const int nblk = (m2 + 15) >> 4;
            const int li = lane & 15, lq = lane >> 4;
            auto Lval = [&](int r, int q) -> cx<T> {
                if (r >= n) return zero;
                if (q < NBP) return q < nbp ? Vst[q * n + r] : zero;
                return (q - NBP) < nbp ? Wl[(q - NBP) * n + r] : zero;
            };
            auto Rval = [&](int c, int q) -> cx<T> {
                if (c >= n) return zero;
                if (q < NBP) return q < nbp ? Wl[q * n + c] : zero;
                return (q - NBP) < nbp ? Vst[(q - NBP) * n + c] : zero;
            };
            for (int t = wv; t < nblk; t += NW) {
                const int bi = nblk - 1 - t;
                const int r = p1 + 16 * bi + li;
                T xre[2 * U], xim[2 * U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const cx<T> L = Lval(r, lq + 4 * u);
                    xre[u] = -L.x;
                    xim[u] = -L.y;
                    xre[u + U] = -L.y;
                    xim[u + U] = L.x;
                }
                // tile (bi, bj): accumulators are loaded ahead, so the global (or LDS)
                // latency of the next tiles overlaps this tile's MFMAs
                const int rr0 = p1 + 16 * bi;
                auto cload = [&](int bj, v4& cre, v4& cim) {
                    const int cc = p1 + 16 * bj + li;
                    cx<T>* cbase = cc < n ? (cc < c0 ? gcol(cc) : lcol(cc)) : nullptr;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int rr = rr0 + M::row(lq, v);
                        cx<T> a = zero;
                        if (cc < n && rr < n && rr >= cc) a = cbase[rr];
                        cre[v] = a.x;
                        cim[v] = a.y;
                    }
                };
                auto tile = [&](int bj, v4 cre, v4 cim) {
                    const int cc = p1 + 16 * bj + li;
                    T y[2 * U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const cx<T> R = Rval(cc, lq + 4 * u);
                        y[u] = R.x;
                        y[u + U] = R.y;
                    }
#pragma unroll
                    for (int q = 0; q < 2 * U; ++q) {
                        cre = M::op(xre[q], y[q], cre);
                        cim = M::op(xim[q], y[q], cim);
                    }
                    cx<T>* cbase = cc < n ? (cc < c0 ? gcol(cc) : lcol(cc)) : nullptr;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int rr = rr0 + M::row(lq, v);
                        if (cc < n && rr < n && rr >= cc) cbase[rr] = cxm<T>(cre[v], rr == cc ? (T)0 : cim[v]);
                    }
                };
                if constexpr (sizeof(T) == 4) {
                    // two tiles in flight: slot 0 holds tile bj, slot 1 tile bj + 1
                    v4 re0, im0, re1, im1;
                    cload(0, re0, im0);
                    if (bi >= 1) cload(1, re1, im1);
                    for (int bj = 0; bj <= bi; bj += 2) {
                        {
                            const v4 cre = re0, cim = im0;
                            if (bj + 2 <= bi) cload(bj + 2, re0, im0);
                            tile(bj, cre, cim);
                        }
                        if (bj + 1 <= bi) {
                            const v4 cre = re1, cim = im1;
                            if (bj + 3 <= bi) cload(bj + 3, re1, im1);
                            tile(bj + 1, cre, cim);
                        }
                    }
                } else {
                    // complex128: one tile ahead (a second slot would spill)
                    v4 nre, nim;
                    cload(0, nre, nim);
                    for (int bj = 0; bj <= bi; ++bj) {
                        const v4 cre = nre, cim = nim;
                        if (bj < bi) cload(bj + 1, nre, nim);
                        tile(bj, cre, cim);
                    }
                }
            }
        }
        __syncthreads();
    }
    cx<T>* Vk = Vout + kk * (int64_t)ld * n;
    if (p_stop < n - 1) {
        // the updated trailing block A[p_stop:, p_stop:] (lower triangle) in place,
        // at its own rows and columns of Vout, for the next reduction to continue
        for (int c = p_stop; c < n; ++c)
            for (int r = c + tid; r < n; r += BS) Vk[r + (int64_t)c * ld] = load(r, c);
    } else if (tid == 0) {
        d[n - 1] = (double)load(n - 1, n - 1).x;
        e[n - 1] = 0.0;
        tau[n - 1] = zero;
    }
    // reflectors out: column j rows j+2..n-1, j < p_stop
    for (int c = 0; c + 2 < n && c < p_stop; ++c)
        for (int r = c + 2 + tid; r < n; r += BS) Vk[r + (int64_t)c * ld] = load(r, c);
}

// ------------------------------------------------------------------ host side

TriPlan tri_plan(int n, int elem_bytes, int lds_limit, int bs) {
    TriPlan tp{};
    tp.n = n;
    tp.bs = bs;
    const int MOD = 128 / elem_bytes;
    const int fixed = elem_bytes == 16 ? TriSmem<double>::fixed_bytes(n, bs, tri_nbp<double>())
                                       : TriSmem<float>::fixed_bytes(n, bs, tri_nbp<float>());
    const int avail = (lds_limit - fixed) / elem_bytes;
    int c0 = n, used = 0;
    for (int c = n - 1; c >= 0; --c) {
        const int L = padded_len(n - c, MOD);
        if (used + L > avail) break;
        used += L;
        c0 = c;
    }
    int g = 0;
    for (int c = 0; c < c0; ++c) g += padded_len(n - c, MOD);
    tp.c0 = c0;
    tp.lds_bytes = fixed + used * elem_bytes;
    tp.scratch_elems = (int64_t)std::max(g, 1);
    return tp;
}

int tridiag_blocked_panel(int elem_bytes) { return elem_bytes == 16 ? tri_nbp<double>() : tri_nbp<float>(); }

template <typename T, int BS>
static hipError_t launch_tri_blocked_bs(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                        void* Vout, double* dd, double* de, void* dtau, int s_end, hipStream_t s) {
    auto kern = tridiag_blocked_kernel<T, BS, tri_nbp<T>()>;
    hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, tp.lds_bytes);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(kern, dim3((unsigned)nk), dim3(BS), tp.lds_bytes, s, dH, tp.n, ld, tp.c0,
                       static_cast<cx<T>*>(scratch), tp.scratch_elems, static_cast<cx<T>*>(Vout), dd, de,
                       static_cast<cx<T>*>(dtau), s_end);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_tri_blocked_t(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                       void* Vout, double* dd, double* de, void* dtau, int s_end, hipStream_t s) {
    if (s_end < tp.n - 1 && (s_end < 0 || s_end % tri_nbp<T>() != 0)) return hipErrorInvalidValue;
    if (tp.n > 64 * kTriMaxRowsPerLane) return hipErrorInvalidValue;  // rows per lane in the product
    if (tp.bs >= 512) return launch_tri_blocked_bs<T, 512>(tp, dH, ld, nk, scratch, Vout, dd, de, dtau, s_end, s);
    return launch_tri_blocked_bs<T, 256>(tp, dH, ld, nk, scratch, Vout, dd, de, dtau, s_end, s);
}

hipError_t launch_tridiag_blocked_f64(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                      double2* Vout, double* dd, double* de, double2* dtau, hipStream_t s, int s_end) {
    return launch_tri_blocked_t<double>(tp, dH, ld, nk, scratch, Vout, dd, de, dtau, s_end, s);
}

hipError_t launch_tridiag_blocked_c64(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                      float2* Vout, double* dd, double* de, float2* dtau, hipStream_t s, int s_end) {
    return launch_tri_blocked_t<float>(tp, dH, ld, nk, scratch, Vout, dd, de, dtau, s_end, s);
}

}  // namespace uepm
