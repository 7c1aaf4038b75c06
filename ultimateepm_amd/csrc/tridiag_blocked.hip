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
// do not fit live in a per-workgroup global scratch region (L2-resident while
// the workgroup runs) and are only touched during the first c0 steps.
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

constexpr int kTriNBP = 4;
constexpr int kTriParts = 4;
constexpr int kTriMaxRowsPerLane = 9;  // n <= 576  // max column parts per row in the symv

template <typename T>
struct TriSmem {
    // bytes of the fixed part of LDS (everything but the matrix region)
    static __host__ __device__ int fixed_bytes(int n, int BS, int NBP) {
        const int E = (int)sizeof(cx<T>);
        int b = ((n * 4 + 15) / 16) * 16;  // colofs
        b += 2 * NBP * n * E;              // panel V and W columns
        b += 2 * n * E;                    // y (row sums / w), ycol (column sums of global rows)
        b += kTriParts * n * E;            // per-part partial row sums of the symv
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
            // S3a: y = A v over row blocks (lane = row).  Blocks are split at c0 so
            // a wave is either all-LDS-rows or all-global-rows; LDS row blocks are
            // further split into T_ column parts so that idle waves take a share.
            // Walks are software-pipelined four elements deep.
            int T_ = 1;
            {
                const int slow0 = j + 1, slow1 = max(j + 1, min(c0, n));  // global rows [slow0, slow1)
                const int ns = (slow1 - slow0 + 63) >> 6;
                const int nf = (n - slow1 + 63) >> 6;
                T_ = nf > 0 ? max(1, min(kTriParts, (NW - ns) / nf)) : 1;
                for (int blk = wv; blk < ns + nf * T_; blk += NW) {
                    if (blk >= ns) {
                        // rows r >= c0, column part pp of [j+1, n)
                        const int fb = (blk - ns) % nf, pp = (blk - ns) / nf;
                        const int rmin = slow1 + 64 * fb;
                        const int r = min(rmin + lane, n - 1);
                        const bool live = rmin + lane < n;
                        const int cs = j + 1 + (pp * m) / T_, ce = j + 1 + ((pp + 1) * m) / T_;
                        cx<T> acc0 = zero, acc1 = zero;
                        // global columns [cs, min(ce, c0)): all c < r, coalesced row walk
                        if (cs < c0) {
                            const int cg = min(ce, c0);
                            int off = colofs[cs] - cs;
                            int P = padded_len(n - cs, MOD);
                            int c = cs;
                            for (; c + 1 < cg; c += 2) {
                                const int off1 = off + P - 1;
                                const int P1 = padded_len(n - c - 1, MOD);
                                const cx<T> a0 = gmat[off + r], a1 = gmat[off1 + r];
                                cfma(acc0, a0, vj[c]);
                                cfma(acc1, a1, vj[c + 1]);
                                off = off1 + P1 - 1;
                                P = padded_len(n - c - 2, MOD);
                            }
                            if (c < cg) cfma(acc0, gmat[off + r], vj[c]);
                        }
                        // LDS columns [max(cs, c0), ce): row part A(r, c) for c <= r,
                        // column part conj(A(c, r)) for c > r
                        const int cl = max(cs, c0);
                        if (cl < ce) {
                            int off = colofs[cl] - cl;
                            int P = padded_len(n - cl, MOD);
                            const int bcol = colofs[r] - r;
                            auto fetch = [&](int c, int o) -> cx<T> {
                                const bool in = c <= r;
                                cx<T> a = lmat[in ? o + r : bcol + c];
                                a.y = in ? a.y : -a.y;
                                return a;
                            };
                            int c = cl;
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
                        if (live) Yp[pp * n + r] = acc0 + acc1;
                    } else {
                        // rows r in [j+1, c0): row part only (global columns c in [j+1, r]);
                        // the column part comes from S3b.
                        const int rmin = slow0 + 64 * blk;
                        const int rr = rmin + lane;
                        const bool live = rr < slow1;
                        const int r = min(rr, slow1 - 1);
                        const int cend = min(slow1, rmin + 64);
                        cx<T> acc0 = zero;
                        int off = colofs[j + 1] - (j + 1);
                        int P = padded_len(n - j - 1, MOD);
                        for (int c = j + 1; c < cend; ++c) {
                            cx<T> a = gmat[max(off + r, 0)];  // above-diagonal reads are discarded
                            if (c > r) a = zero;
                            cfma(acc0, a, vj[c]);
                            off += P - 1;
                            P = padded_len(n - c - 1, MOD);
                        }
                        if (live) Yp[r] = acc0;
                    }
                }
            }
            // S3b: column part of rows r in [j+1, c0): wave dots over global column r,
            // eight rows per wave at a time with all loads issued before the reductions.
            for (int r0 = j + 1 + wv; r0 < min(c0, n); r0 += 8 * NW) {
                cx<T> accr[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int r = r0 + i * NW;
                    cx<T> acc = zero;
                    if (r < min(c0, n)) {
                        const cx<T>* colr = gcol(r);
                        for (int c = r + 1 + lane; c < n; c += 64) cfmac(acc, colr[c], vj[c]);
                    }
                    accr[i] = acc;
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int r = r0 + i * NW;
                    const T sx = wsum(accr[i].x), sy = wsum(accr[i].y);
                    if (lane == 0 && r < min(c0, n)) ycol[r] = cxm<T>(sx, sy);
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
                    if (r < c0) y = y + ycol[r];
                    else
                        for (int pp = 1; pp < T_; ++pp) y = y + Yp[pp * n + r];
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
                for (int bj = 0; bj <= bi; ++bj) {
                    const int cc = p1 + 16 * bj + li;
                    T y[2 * U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const cx<T> R = Rval(cc, lq + 4 * u);
                        y[u] = R.x;
                        y[u + U] = R.y;
                    }
                    const int rr0 = p1 + 16 * bi;
                    v4 cre, cim;
                    cx<T>* cbase = nullptr;
                    if (cc < n) cbase = cc < c0 ? gcol(cc) : lcol(cc);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int rr = rr0 + M::row(lq, v);
                        cx<T> a = zero;
                        if (cc < n && rr < n && rr >= cc) a = cbase[rr];
                        cre[v] = a.x;
                        cim[v] = a.y;
                    }
#pragma unroll
                    for (int q = 0; q < 2 * U; ++q) {
                        cre = M::op(xre[q], y[q], cre);
                        cim = M::op(xim[q], y[q], cim);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int rr = rr0 + M::row(lq, v);
                        if (cc < n && rr < n && rr >= cc) cbase[rr] = cxm<T>(cre[v], rr == cc ? (T)0 : cim[v]);
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
    const int fixed = elem_bytes == 16 ? TriSmem<double>::fixed_bytes(n, bs, kTriNBP)
                                       : TriSmem<float>::fixed_bytes(n, bs, kTriNBP);
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

int tridiag_blocked_panel() { return kTriNBP; }

template <typename T, int BS>
static hipError_t launch_tri_blocked_bs(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                        void* Vout, double* dd, double* de, void* dtau, int s_end, hipStream_t s) {
    auto kern = tridiag_blocked_kernel<T, BS, kTriNBP>;
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
    if (s_end < tp.n - 1 && (s_end < 0 || s_end % kTriNBP != 0)) return hipErrorInvalidValue;
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
