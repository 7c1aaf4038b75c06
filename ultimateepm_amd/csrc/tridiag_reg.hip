// tridiag_reg.hip — register-resident Householder reduction of a Hermitian
// H(k) to real symmetric tridiagonal form, one 512-lane workgroup per
// k-point (gfx950): complex64 arithmetic for the precision form (dim <= 288),
// complex128 for the FP64 form (dim <= 192), and a chain of shrinking grids
// (below).
//
// Layout.  The 512 lanes form a PR x PC = 32 x 16 grid: lane (r, c) with
// c = tid % 16, r = tid / 16 (a wave holds four r values and all sixteen c).
// Lane (r, c) keeps in VGPRs the lower-triangle elements A(i, j) with
// i = r + 32 l, j = c + 16 m (cyclic in both directions, so every lane keeps
// a share of the shrinking trailing matrix until the end).  Only blocks
// (l, m) that can hold i >= j are stored: 42 complex values per lane at
// dim <= 192.  Blocks below the 32 x 32 diagonal super-blocks hold A(i, j);
// the two blocks of a row block that lie in its diagonal super-block hold
// A(i, j) / 2 for every element, above the diagonal too, so that each pair
// (i, j), (j, i) of a super-block is stored twice at half weight.  The
// symmetric product then treats every stored element alike: A(i, j) feeds
// row i (A v_j) and column j (conj(A) v_i).  A grid whose blocks
// do not all fit its register budget keeps its last NL blocks in LDS instead,
// in lane-private slots (slot-major, so a wave's accesses are consecutive).
//
// Step s (LAPACK xHETD2, lower, restated; H_s = I - tau v v^H, v[s+1] = 1):
//   a. the lanes of column s put x = A(s+1:, s) in LDS              | barrier
//   b. every wave forms beta, tau, v from x (redundantly)
//   c. y = A v: row sums across the 16 lanes of a row group (DPP), column
//      sums across the 4 row groups of a wave (row swaps), one
//      partial per wave in LDS; each lane also sums its share of v^H A v,
//      so p^H v = conj(tau) v^H A v needs no separate dot product      | barrier
//   d. y = sum of the partials, one row per lane                       | barrier
//   e. w = tau y - (1/2)|tau|^2 (v^H A v) v; A -= v w^H + w v^H on the live blocks
// Blocks whose columns are all <= s are skipped (one wave-uniform test per
// block column: a stored block's rows are live while its columns are).
// Outputs match the LDS-blocked reduction: d, e (real), tau, reflectors below
// the sub-diagonal of column s of Vout (v[s+1] = 1 implied).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>

#include "solver.hpp"
#include "wave_ops.hpp"

namespace uepm {

namespace {

// complex in the arithmetic type R (float for the precision form, double for FP64)
template <typename R>
struct c2 {
    R x, y;
};
template <typename R>
__device__ __forceinline__ c2<R> mk(R a, R b) {
    return {a, b};
}
template <typename R>
__device__ __forceinline__ c2<R> add(c2<R> a, c2<R> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ c2<R> mul(c2<R> a, c2<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// c ? a : b per component (a select of the whole struct would go through memory)
template <typename R>
__device__ __forceinline__ c2<R> sel(bool c, c2<R> a, c2<R> b) {
    return {c ? a.x : b.x, c ? a.y : b.y};
}
// acc += a * b
template <typename R>
__device__ __forceinline__ void fma_ab(c2<R>& acc, c2<R> a, c2<R> b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// acc += conj(a) * b
template <typename R>
__device__ __forceinline__ void fma_cab(c2<R>& acc, c2<R> a, c2<R> b) {
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
}

// Reductions: DPP inside a 16-lane row, VALU row swaps across the rows of a
// wave (wave_ops.hpp; 64-bit values move as two 32-bit halves).
template <typename R>
__device__ __forceinline__ void swap16(R& a, R& b) {
    wave::Words<R> wa(a), wb(b);
#pragma unroll
    for (int i = 0; i < wave::Words<R>::n; ++i) wave::swap16(wa.w[i], wb.w[i]);
    a = wa.value();
    b = wb.value();
}
template <typename R>
__device__ __forceinline__ void swap32(R& a, R& b) {
    wave::Words<R> wa(a), wb(b);
#pragma unroll
    for (int i = 0; i < wave::Words<R>::n; ++i) wave::swap32(wa.w[i], wb.w[i]);
    a = wa.value();
    b = wb.value();
}
// Reduce-scatter of two complex column partials over the four rows of a wave:
// returns, per lane, row 0: Re p0, row 1: Im p0, row 2: Re p1, row 3: Im p1
// (each summed over the four rows).
template <typename R>
__device__ __forceinline__ R cross_rows4_pair(R p0x, R p0y, R p1x, R p1y) {
    swap16(p0x, p0y);  // rows 0,2: Re p0 parts; rows 1,3: Im p0 parts
    p0x += p0y;
    swap16(p1x, p1y);
    p1x += p1y;
    swap32(p0x, p1x);  // lanes 0-31: p0 (Re | Im), lanes 32-63: p1 (Re | Im)
    return p0x + p1x;
}

// 1/x: hardware reciprocal and Newton steps to full precision
__device__ __forceinline__ float rcp_full(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.f), r, r);
}
__device__ __forceinline__ double rcp_full(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

template <typename VT, typename R>
__device__ __forceinline__ VT store_c(R x, R y) {
    VT v;
    v.x = x;
    v.y = y;
    return v;
}


// Stored blocks of the cyclic lower triangle.  With PR = 2 PC and LC = 2 LR
// (the only shape used) row block l stores block columns 0..2l+1, so the
// offsets have closed forms: everything below folds to constants once the
// loops over l and m are unrolled, which keeps the slot array in registers.
template <int PR, int PC, int LR, int LC>
struct RegLayout {
    static_assert(PR == 2 * PC && LC == 2 * LR, "closed forms below assume a 2:1 grid with square coverage");
    static constexpr int mmax(int l) { return 2 * l + 1; }  // = min((PR l + PR - 1) / PC, LC - 1)
    static constexpr int base(int l) { return l * (l + 1); }  // sum of mmax(q) + 1 over q < l
    static constexpr int count() { return base(LR); }
    static constexpr int idx(int l, int m) { return base(l) + m; }
    // first row block l with block (l, m) stored: the smallest l with 2l + 1 >= m
    static constexpr int lmin(int m) { return m / 2; }
    // blocks (l, 2l) and (l, 2l + 1) lie in the diagonal 32 x 32 super-block of
    // rows 32l..32l+31: i and j there are in the same super-block, and so are
    // the mirror pair (j, i)
    static constexpr bool straddles(int l, int m) { return m >= 2 * l; }
};

}  // namespace

// Stage arguments: steps s < s_end run here (s_end >= n - 1: the whole
// reduction); when s_end < n - 1 the updated trailing block A[s_end:, s_end:]
// (lower triangle, full diagonal) goes to Bout row by row (element (i, j) at
// j + i ldb, ldb * ldb per k) for a smaller grid to continue.  IN says how
// the lower triangle of the input is read: 0, column-major (element (i, j) at
// i + j ld); 1, as the conjugate of the upper triangle of a full Hermitian
// matrix (element (j, i) at j + i ld); 2, row by row as a hand-over block.  Input (leading
// dimension ld), reflectors (ldv) and d, e, tau of k-point kk start at
// kk * h_kstride, kk * v_kstride and kk * de_kstride.
template <int PR, int PC, int LR, int LC, int MINWG, typename R, typename VT, typename THin, bool LOW = false,
          int NL = 0, int IN = 0>
__global__ __launch_bounds__(PR* PC, MINWG) void tridiag_reg_kernel(const THin* __restrict__ Hin, int64_t h_kstride,
                                                                    int n, int ld, VT* __restrict__ Vout,
                                                                    int ldv, int64_t v_kstride,
                                                                    double* __restrict__ dout,
                                                                    double* __restrict__ eout,
                                                                    VT* __restrict__ tauout, int64_t de_kstride,
                                                                    int s_end, VT* __restrict__ Bout, int ldb) {
    constexpr int BS = PR * PC;
    constexpr int NW = BS / 64;
    using Lay = RegLayout<PR, PC, LR, LC>;
    using f2 = c2<R>;
    constexpr int NS = Lay::count();
    constexpr int NX = PR * LR;  // rows covered by the grid
    constexpr int NXR = (NX + 63) / 64;
    static_assert(PC == 16 && BS % 64 == 0, "row groups are the 16-lane DPP rows of a wave");
    static_assert(PC * LC == NX && NX <= BS && LC % 2 == 0, "square coverage, one y entry per lane");

    // column s: zero at rows <= s and >= n; two buffers by step parity, since the
    // LOW variant still reads the column of step s in its update phase while a
    // wave ahead may write the column of step s + 1 (no barrier in between)
    __shared__ f2 Xbuf[2][NX];
    __shared__ f2 XD[NX];      // column s as stored (its diagonal entry is A(s, s) / 2)
    __shared__ f2 RP[NX];      // row sums of the symmetric product
    __shared__ f2 CP[NW][NX];  // column sums, one row per wave
    __shared__ f2 Y[NX];       // y = A v, zero at rows <= s and >= n
    __shared__ R Q[NW];    // per-wave partial v^H A v
    // the last NL stored blocks of every lane live in LDS (lane-private slots,
    // slot-major so that a wave's accesses are consecutive), the others in VGPRs
    static_assert(NL >= 0 && NL < NS, "LDS slots");
    constexpr int NR = NS - NL;
    __shared__ f2 AL[NL > 0 ? NL : 1][NL > 0 ? BS : 1];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int c = tid % PC, r = tid / PC;
    const int64_t kk = blockIdx.x;
    const THin* H = Hin + kk * h_kstride;
    VT* Vk = Vout + kk * v_kstride;
    double* d = dout + kk * de_kstride;
    double* e = eout + kk * de_kstride;
    VT* tau_o = tauout + kk * de_kstride;
    const int s_stop = s_end < n - 1 ? s_end : n - 1;

    // ---- load: below the diagonal super-blocks A(i, j), inside them A(i, j) / 2 ----
    f2 A[NR];
    // slot i of this lane (i is a compile-time constant wherever these are used)
    auto aget = [&](int i) -> f2 {
        if constexpr (NL > 0)
            if (i >= NR) return AL[i - NR][tid];
        return A[i];
    };
    auto aset = [&](int i, f2 v) {
        if constexpr (NL > 0)
            if (i >= NR) {
                AL[i - NR][tid] = v;
                return;
            }
        A[i] = v;
    };
#pragma unroll
    for (int l = 0; l < LR; ++l) {
#pragma unroll
        for (int m = 0; m <= Lay::mmax(l); ++m) {
            const int i = r + PR * l, j = c + PC * m;
            f2 a = mk<R>(0, 0);
            if (i < n && j < n && (i >= j || Lay::straddles(l, m))) {
                // the lower element (p, q), p >= q: (i, j), or (j, i) conjugated above the
                // diagonal.  A wave's 16 columns j are consecutive: IN 1 and 2 read the
                // lower elements as contiguous runs (row i of the stored triangle)
                const bool up = i < j;
                const int p = up ? j : i, q = up ? i : j;
                const THin h = H[IN == 0 ? p + (int64_t)q * ld : q + (int64_t)p * ld];
                const R f = Lay::straddles(l, m) ? R(0.5) : R(1);
                const R hy = (IN == 1) != up ? -(R)h.y : (R)h.y;
                a = mk<R>((R)h.x * f, hy * f);
            }
            aset(Lay::idx(l, m), a);
        }
    }
    const f2 one = mk<R>(1, 0), zero = mk<R>(0, 0);

    for (int s = 0; s < s_stop; ++s) {
        f2* X = Xbuf[s & 1];
        const int cs = s % PC, ms = s / PC;
        const int l0 = (s + 1) / PR, m0 = (s + 1) / PC;  // first live row / column blocks
        // ---- a. column s to LDS (every row of the grid) ----
        const bool own = c == cs;
#pragma unroll
        for (int m = 0; m < LC; ++m) {
            if (m == ms) {  // wave-uniform
                if (own) {
#pragma unroll
                    for (int l = 0; l < LR; ++l) {
                        const int i = r + PR * l;
                        if (m <= Lay::mmax(l)) {
                            const f2 a = aget(Lay::idx(l, m));
                            const R f = Lay::straddles(l, m) ? R(2) : R(1);
                            X[i] = sel(i > s, mk(f * a.x, f * a.y), zero);
                            XD[i] = a;
                        } else {
                            X[i] = zero;  // block above the diagonal: rows < s
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- b. reflector, redundantly in every wave ----
        R part = 0;
#pragma unroll
        for (int q = 0; q < NXR; ++q) {
            const int i = s + 2 + lane + 64 * q;
            const f2 a = X[i < NX ? i : NX - 1];
            const R t = fma(a.x, a.x, a.y * a.y);
            part += i < NX ? t : R(0);
        }
        const R xn2 = wave::sum(part);
        const f2 al = X[s + 1];
        R beta = al.x, tr = 0, ti = 0, sr = 1, si = 0;
        if (!(xn2 == R(0) && al.y == R(0))) {
            const R nrm = sqrt(fma(al.x, al.x, fma(al.y, al.y, xn2)));
            beta = al.x >= R(0) ? -nrm : nrm;
            const R ib = rcp_full(beta);
            tr = (beta - al.x) * ib;
            ti = -al.y * ib;
            const R br = al.x - beta, iden = rcp_full(fma(br, br, al.y * al.y));
            sr = br * iden;
            si = -al.y * iden;
        }
        const f2 tau = mk(tr, ti);
        const f2 sc = mk(sr, si);
        if (tid == 0) {
            d[s] = 2.0 * (double)XD[s].x;
            e[s] = (double)beta;
            tau_o[s] = store_c<VT>(tau.x, tau.y);
        }
        {
            const int i = s + 2 + tid;
            if (i < n) {
                const f2 v = mul(X[i], sc);
                Vk[i + (int64_t)s * ldv] = store_c<VT>(v.x, v.y);
            }
        }
        // v at this lane's rows and columns (X is zero outside s+1..n-1; loads
        // unconditional so they issue together)
        // LOW (fewer registers, for two workgroups per CU): the column values of v and
        // w are formed from LDS where they are used instead of being kept, and each
        // pair of block columns is reduced as soon as its partial sums are complete
        // (the same operations in the same order)
        auto vcol = [&](int m) -> f2 {
            const int j = c + PC * m;
            return sel(j == s + 1, one, mul(X[j], sc));
        };
        f2 vr[LR], vc[LOW ? 1 : LC];
#pragma unroll
        for (int l = 0; l < LR; ++l) {
            const int i = r + PR * l;
            const f2 v = mul(X[i], sc);
            vr[l] = sel(i == s + 1, one, v);
        }
        if constexpr (!LOW) {
#pragma unroll
            for (int m = 0; m < LC; ++m) vc[m] = vcol(m);
        }
        // ---- c. y = A v on the live blocks; local share of v^H A v ----
        f2 rp[LR], cp[LOW ? 2 : LC];
#pragma unroll
        for (int l = 0; l < LR; ++l) rp[l] = zero;
#pragma unroll
        for (int m = 0; m < (LOW ? 2 : LC); ++m) cp[m] = zero;
        const int q4 = lane >> 4;
        R* cpw = &CP[wv][0].x;
#pragma unroll
        for (int m = LC - 1; m >= 0; --m) {
            const int mc = LOW ? (m & 1) : m;  // cp slot of block column m
            if (m >= m0) {  // a stored block is live while its block column is
                const f2 vcm = LOW ? vcol(m) : vc[LOW ? 0 : m];
#pragma unroll
                for (int l = Lay::lmin(m); l < LR; ++l) {
                    const f2 a = aget(Lay::idx(l, m));
                    fma_ab(rp[l], a, vcm);
                    fma_cab(cp[mc], a, vr[l]);
                }
            }
            if constexpr (LOW) {
                // column sums of the pair (m, m + 1), m even: lane row q holds (Re, Im)
                // of block column m (q < 2) or m + 1 (q >= 2)
                if ((m & 1) == 0) {
                    if (m + 1 >= m0) {
                        const R t = cross_rows4_pair(cp[0].x, cp[0].y, cp[1].x, cp[1].y);
                        cpw[2 * (c + PC * (m + (q4 >> 1))) + (q4 & 1)] = t;
                    }
                    cp[0] = zero;
                    cp[1] = zero;
                }
            }
        }
        // sum over stored elements of conj(v_i) A_ij v_j; v^H A v = 2 Re of its total
        R qv = 0;
#pragma unroll
        for (int l = 0; l < LR; ++l)
            if (l >= l0) qv = fma(vr[l].x, rp[l].x, fma(vr[l].y, rp[l].y, qv));
#pragma unroll
        for (int l = 0; l < LR; ++l) {
            if (l >= l0) {
                const R sx = wave::row_sum16(rp[l].x), sy = wave::row_sum16(rp[l].y);
                if (c == 0) RP[r + PR * l] = mk(sx, sy);
            }
        }
        // column sums, two block columns per reduce-scatter: lane row q holds
        // (Re, Im) of block column m (q < 2) or m + 1 (q >= 2)
        if constexpr (!LOW) {
#pragma unroll
            for (int m = 0; m < LC; m += 2) {
                if (m + 1 >= m0) {
                    const R t = cross_rows4_pair(cp[m].x, cp[m].y, cp[LOW ? 0 : m + 1].x, cp[LOW ? 0 : m + 1].y);
                    cpw[2 * (c + PC * (m + (q4 >> 1))) + (q4 & 1)] = t;
                }
            }
        }
        qv = wave::sum(qv);
        if (lane == 0) Q[wv] = qv;
        __syncthreads();
        // ---- d. y = row sums + column sums of the waves, one row per lane; then
        // w = tau y + alpha v, alpha = -(tau/2) p^H v = -(1/2)|tau|^2 v^H A v, and v
        // itself, published for the update: w over y in Y, v over the row sums in RP
        // (the next writes of both come after a barrier that every wave passes only
        // once it has finished this step) ----
        if (tid < NX) {
            f2 y = zero;
            if (tid > s && tid < n) {
                y = RP[tid];
#pragma unroll
                for (int w = 0; w < NW; ++w) y = add(y, CP[w][tid]);
            }
            R vav = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) vav += Q[w];
            const R alpha = -(tau.x * tau.x + tau.y * tau.y) * vav;  // v^H A v = 2 vav
            const f2 v = sel(tid == s + 1, one, mul(X[tid], sc));
            f2 wt = mul(tau, y);
            wt.x = fma(alpha, v.x, wt.x);
            wt.y = fma(alpha, v.y, wt.y);
            Y[tid] = wt;
            RP[tid] = v;
        }
        __syncthreads();
        // ---- e. rank-2 update A -= v w^H + w v^H, at half weight in the diagonal super-blocks ----
        f2 wr[LR], vh[LR], wh[LR];
#pragma unroll
        for (int l = 0; l < LR; ++l) {
            wr[l] = Y[r + PR * l];
            vh[l] = mk(R(0.5) * vr[l].x, R(0.5) * vr[l].y);
            wh[l] = mk(R(0.5) * wr[l].x, R(0.5) * wr[l].y);
        }
#pragma unroll
        for (int m = LC - 1; m >= 0; --m) {
            if (m >= m0) {
                const f2 vcm = LOW ? RP[c + PC * m] : vc[LOW ? 0 : m];
                const f2 wcm = Y[c + PC * m];
#pragma unroll
                for (int l = Lay::lmin(m); l < LR; ++l) {
                    // a -= v_i conj(w_j) + w_i conj(v_j)
                    const bool half = Lay::straddles(l, m);
                    const f2 vi = half ? vh[l] : vr[l], wi = half ? wh[l] : wr[l];
                    f2 a = aget(Lay::idx(l, m));
                    a.x = fma(-vi.x, wcm.x, a.x);
                    a.x = fma(-vi.y, wcm.y, a.x);
                    a.x = fma(-wi.x, vcm.x, a.x);
                    a.x = fma(-wi.y, vcm.y, a.x);
                    a.y = fma(-vi.y, wcm.x, a.y);
                    a.y = fma(vi.x, wcm.y, a.y);
                    a.y = fma(-wi.y, vcm.x, a.y);
                    a.y = fma(wi.x, vcm.y, a.y);
                    aset(Lay::idx(l, m), a);
                }
            }
        }
    }
    if (s_stop < n - 1) {
        // hand the trailing block A[s_stop:, s_stop:] over (diagonal stored halved)
        VT* B = Bout + kk * (int64_t)ldb * ldb;
#pragma unroll
        for (int l = 0; l < LR; ++l)
#pragma unroll
            for (int m = 0; m <= Lay::mmax(l); ++m) {
                const int i = r + PR * l, j = c + PC * m;
                if (i < n && j >= s_stop && i >= j) {
                    const f2 a = aget(Lay::idx(l, m));
                    const R f = Lay::straddles(l, m) ? R(2) : R(1);
                    B[(j - s_stop) + (int64_t)(i - s_stop) * ldb] = store_c<VT>(a.x * f, a.y * f);
                }
            }
        return;
    }
    // last diagonal element
    {
        const int s = n - 1;
        const int cs = s % PC, ms = s / PC;
#pragma unroll
        for (int m = 0; m < LC; ++m)
            if (m == ms)
#pragma unroll
                for (int l = 0; l < LR; ++l)
                    if (m <= Lay::mmax(l) && c == cs && r + PR * l == s) d[s] = 2.0 * (double)aget(Lay::idx(l, m)).x;
        if (tid == 0) {
            e[s] = 0.0;
            tau_o[s] = store_c<VT>(R(0), R(0));
        }
    }
}

// One wave per k-point for a block of size n <= 64 (complex64; the chain's
// last stage, or the whole reduction of a small matrix).  Lane i keeps row i
// of the Hermitian block in VGPRs, both triangles, in shifted columns: at step
// s, R[t] = A(i, s + t).  Column s is then R[0] on every lane, the product
// y = A v is a local dot product of the row with v (read from LDS, where each
// lane publishes its v_i and w_i), and the rank-2 update writes R[t - 1] from
// R[t], which moves the columns along at no cost.  No barriers: the wave is
// the workgroup.  Columns are processed in blocks of 8, skipped once they are
// past the live part (those registers are never read again).  Same outputs
// as the grid kernels (d, e, tau, reflectors below the sub-diagonal of column
// s of Vout).  The block is read as the grids read it (IN: 0 lower triangle
// by columns, 1 conjugate of the upper triangle of a full Hermitian matrix,
// 2 lower triangle by rows).
template <int IN, typename THin>
__global__ __launch_bounds__(64, 2) void tridiag_wave64_kernel(const THin* __restrict__ Hin, int64_t h_kstride, int n,
                                                               int ld, float2* __restrict__ Vout, int ldv,
                                                               int64_t v_kstride, double* __restrict__ dout,
                                                               double* __restrict__ eout, float2* __restrict__ tauout,
                                                               int64_t de_kstride) {
    using f2 = c2<float>;
    constexpr int W = 64;
    // (v, w) of lane j at VW[j]; entries W..2W-1 stay zero (columns past the block)
    __shared__ float4 VW[2 * W];
    const int i = threadIdx.x;
    const int64_t kk = blockIdx.x;
    const THin* H = Hin + kk * h_kstride;
    float2* Vk = Vout + kk * v_kstride;
    double* d = dout + kk * de_kstride;
    double* e = eout + kk * de_kstride;
    float2* tau_o = tauout + kk * de_kstride;
    const f2 zero = mk<float>(0, 0), one = mk<float>(1, 0);
    VW[W + i] = make_float4(0.f, 0.f, 0.f, 0.f);

    // ---- row i of the block (zero outside it) ----
    f2 R[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        f2 a = zero;
        if (i < n && j < n) {
            // lower element A(r, c), r >= c, as the grids read it
            const int r = j <= i ? i : j, c = j <= i ? j : i;
            const THin h = H[IN == 0 ? r + (int64_t)c * ld : c + (int64_t)r * ld];
            const float hx = (float)h.x, hy = IN == 1 ? -(float)h.y : (float)h.y;
            a = j <= i ? mk(hx, hy) : mk(hx, -hy);  // A(i, j) = conj(A(j, i)) above the diagonal
        }
        R[j] = a;
    }

    for (int s = 0; s < n - 1; ++s) {
        const int L = n - 1 - s;  // live columns s + 1 .. n - 1 sit in R[1 .. L]
        // ---- reflector from x = A(s+1:, s), redundantly in every lane ----
        const f2 x = R[0];
        const bool below = i > s + 1 && i < n;
        const float xn2 = wave::sum(below ? fmaf(x.x, x.x, x.y * x.y) : 0.f);
        const float alx = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x.x), s + 1));
        const float aly = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x.y), s + 1));
        const float dss = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x.x), s));
        float beta = alx, tr = 0, ti = 0, sr = 1, si = 0;
        if (!(xn2 == 0.f && aly == 0.f)) {
            const float nrm = sqrtf(fmaf(alx, alx, fmaf(aly, aly, xn2)));
            beta = alx >= 0.f ? -nrm : nrm;
            const float ib = rcp_full(beta);
            tr = (beta - alx) * ib;
            ti = -aly * ib;
            const float br = alx - beta, iden = rcp_full(fmaf(br, br, aly * aly));
            sr = br * iden;
            si = -aly * iden;
        }
        const f2 tau = mk(tr, ti), sc = mk(sr, si);
        if (i == 0) {
            d[s] = (double)dss;
            e[s] = (double)beta;
            tau_o[s] = make_float2(tau.x, tau.y);
        }
        const f2 v = i == s + 1 ? one : (below ? mul(x, sc) : zero);
        if (below) Vk[i + (int64_t)s * ldv] = make_float2(v.x, v.y);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        VW[i] = make_float4(v.x, v.y, 0.f, 0.f);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // ---- y = A v over the live columns (v_j = 0 for j <= s) ----
        f2 y = zero;
#pragma unroll
        for (int b = 0; b < W / 8; ++b) {
            if (8 * b + 1 <= L) {  // wave-uniform
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int t = 8 * b + u + 1;
                    if (t < W) {
                        const float4 q = VW[s + t];
                        fma_ab(y, R[t], mk(q.x, q.y));
                    }
                }
            }
        }
        // ---- w = tau y + alpha v, alpha = -(1/2) |tau|^2 v^H A v ----
        const float vav = wave::sum(fmaf(v.x, y.x, v.y * y.y));
        const float alpha = -0.5f * (tau.x * tau.x + tau.y * tau.y) * vav;
        f2 w = mul(tau, y);
        w.x = fmaf(alpha, v.x, w.x);
        w.y = fmaf(alpha, v.y, w.y);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        VW[i] = make_float4(v.x, v.y, w.x, w.y);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // ---- A -= v w^H + w v^H on the live columns, shifted down by one ----
#pragma unroll
        for (int b = 0; b < W / 8; ++b) {
            if (8 * b + 1 <= L) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int t = 8 * b + u + 1;
                    if (t < W) {
                        const float4 q = VW[s + t];  // (v_j, w_j), j = s + t
                        f2 a = R[t];
                        a.x = fmaf(-v.x, q.z, a.x);
                        a.x = fmaf(-v.y, q.w, a.x);
                        a.x = fmaf(-w.x, q.x, a.x);
                        a.x = fmaf(-w.y, q.y, a.x);
                        a.y = fmaf(-v.y, q.z, a.y);
                        a.y = fmaf(v.x, q.w, a.y);
                        a.y = fmaf(-w.y, q.x, a.y);
                        a.y = fmaf(w.x, q.y, a.y);
                        R[t - 1] = a;
                    }
                }
            }
        }
    }
    // last diagonal element
    {
        const int s = n - 1;
        const float dss = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, R[0].x), s));
        if (i == 0) {
            d[s] = (double)dss;
            e[s] = 0.0;
            tau_o[s] = make_float2(0.f, 0.f);
        }
    }
}

// n <= 64: 2 x 4 blocks; n <= 128: 4 x 8; n <= 192: 6 x 12; n <= 256: 8 x 16;
// n <= 288: 9 x 18 (rows x columns of the 32 x 16 grid)
int tridiag_reg_max_n() { return 288; }

// Hand-over between grids.  The late steps of a large grid carry its full
// per-step overhead for a small trailing matrix, so each grid stops where the
// next smaller one can take the trailing block: 288-grid (9 x 18 blocks) ->
// 256 (8 x 16) -> 224 (7 x 14) -> 192 (6 x 12) -> 160 (5 x 10) -> 128 (4 x 8)
// -> 96 (3 x 6) -> 64 (2 x 4).  Register budgets (workgroups per CU): 288
// low-register, 256 VGPRs with 16 of its 90 blocks in LDS (1); 256 and 224
// low-register, 256 and 223 VGPRs (1); 192 low-register, 128 VGPRs with 14 of
// its 42 blocks in LDS (2); 160 low-register, 128 VGPRs (2); 128 at 128 VGPRs
// (2); 96 low-register, 80 (3); 64 low-register, 63 (4).
namespace {
constexpr int kNumGrids = 8;
constexpr int kGrid[kNumGrids] = {288, 256, 224, 192, 160, 128, 96, 64};
// index of the smallest grid that holds n
int grid_of(int n) {
    int g = 0;
    while (g < kNumGrids - 1 && n <= kGrid[g + 1]) ++g;
    return g;
}
}  // namespace

int tridiag_reg_split(int n) {
    const int g = grid_of(n);
    return g < kNumGrids - 1 ? kGrid[g + 1] : 0;
}

size_t tridiag_reg_work_bytes(int n) {
    // two ping-pong hand-over blocks: the first hand-over's size and the next one's
    const int t1 = tridiag_reg_split(n);
    if (!t1) return 0;
    const int t2 = tridiag_reg_split(t1);
    return sizeof(float2) * ((size_t)t1 * t1 + (size_t)t2 * t2);
}

namespace {
template <int IN, typename THin>
void wave64_stage(const THin* H, int64_t h_kstride, int n, int ld, int64_t nk, float2* V, int ldv, int64_t v_kstride,
                  double* dd, double* de, float2* dtau, int64_t de_kstride, hipStream_t s) {
    hipLaunchKernelGGL((tridiag_wave64_kernel<IN, THin>), dim3((unsigned)nk), dim3(64), 0, s, H, h_kstride, n, ld, V, ldv,
                       v_kstride, dd, de, dtau, de_kstride);
}

template <typename R, typename VT, int LR, int LC, int MINWG, int IN, typename THin, bool LOW = false, int NL = 0>
void reg_stage(const THin* H, int64_t h_kstride, int n, int ld, int64_t nk, VT* V, int ldv, int64_t v_kstride,
               double* dd, double* de, VT* dtau, int64_t de_kstride, int s_end, VT* B, int ldb, hipStream_t s) {
    hipLaunchKernelGGL((tridiag_reg_kernel<32, 16, LR, LC, MINWG, R, VT, THin, LOW, NL, IN>), dim3((unsigned)nk), dim3(512), 0,
                       s, H, h_kstride, n, ld, V, ldv, v_kstride, dd, de, dtau, de_kstride, s_end, B, ldb);
}

constexpr int NL288 = 16, NL192 = 14;  // blocks held in LDS by the 288- and 192-wide grids

template <typename THin, int IN>
void grid_stage(int g, const THin* H, int64_t hk, int m, int ld, int64_t nk, float2* V, int ldv, int64_t vk,
                double* dd, double* de, float2* dtau, int64_t dk, int s_end, float2* B, int ldb, hipStream_t s) {
    switch (g) {
        case 0: reg_stage<float, float2, 9, 18, 2, IN, THin, true, NL288>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 1: reg_stage<float, float2, 8, 16, 2, IN, THin, true>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 2: reg_stage<float, float2, 7, 14, 2, IN, THin, true>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 3: reg_stage<float, float2, 6, 12, 4, IN, THin, true, NL192>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 4: reg_stage<float, float2, 5, 10, 4, IN, THin, true>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 5: reg_stage<float, float2, 4, 8, 4, IN, THin>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 6: reg_stage<float, float2, 3, 6, 6, IN, THin, true>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        default: wave64_stage<IN, THin>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s); break;  // last: s_end >= m - 1
    }
}
}  // namespace

namespace {
// The complex64 chain on a block of size m whose element (0, 0) sits at H
// (leading dimension ld, per-k stride hk; read as IN says, IN = 0 or 1);
// outputs at the block's offset of the full matrix's V (ldv, stride vk) and
// d, e, tau (stride dk).
template <int IN, typename THin>
hipError_t c64_chain(const THin* H, int64_t hk, int m, int ld, int64_t nk, float2* V, int ldv, int64_t vk, double* dd,
                     double* de, float2* dtau, int64_t dk, float2* work, hipStream_t s) {
    const int t1 = tridiag_reg_split(m);
    if (!t1) {
        grid_stage<THin, IN>(grid_of(m), H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, m, (float2*)nullptr, 0, s);
        return hipGetLastError();
    }
    if (!work) return hipErrorInvalidValue;
    // first grid on the block, then each smaller grid on the trailing block of
    // its size, in ping-pong blocks (the block of size t lives at
    // buf[i] + k * t * t; a stage reads all of its block before it writes the next)
    float2* buf[2] = {work, work + nk * (int64_t)t1 * t1};
    grid_stage<THin, IN>(grid_of(m), H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, m - t1, buf[0], t1, s);
    int r = t1, cur = 0;
    while (true) {
        const int off = m - r;
        const int t = tridiag_reg_split(r);
        grid_stage<float2, 2>(grid_of(r), (const float2*)buf[cur], (int64_t)r * r, r, r, nk, V + off + (int64_t)off * ldv, ldv,
                   vk, dd + off, de + off, dtau + off, dk, t ? r - t : r, t ? buf[cur ^ 1] : (float2*)nullptr, t,
                   s);
        if (!t) break;
        r = t;
        cur ^= 1;
    }
    return hipGetLastError();
}
}  // namespace

hipError_t launch_tridiag_reg_c64(const double2* dH, int n, int ld, int64_t nk, float2* Vout, double* dd, double* de,
                                  float2* dtau, float2* work, hipStream_t s) {
    if (n < 1 || n > kGrid[0] || nk <= 0) return hipErrorInvalidValue;
    const int64_t hk = (int64_t)ld * n;
    return c64_chain<1>(dH, hk, n, ld, nk, Vout, ld, hk, dd, de, dtau, n, work, s);
}

int tridiag_reg_c64_tail(int n) {
    if (n < 2) return 0;
    if (n <= kGrid[0]) return n;
    const int P = tridiag_blocked_panel((int)sizeof(float2));
    const int s0 = ((n - kGrid[0] + P - 1) / P) * P;
    return n - s0;
}

hipError_t launch_tridiag_reg_c64_block(const float2* H, int64_t hk, int m, int ld, int64_t nk, float2* V, int ldv,
                                        int64_t vk, double* dd, double* de, float2* dtau, int64_t dk, float2* work,
                                        hipStream_t s) {
    if (m < 1 || m > kGrid[0] || nk <= 0) return hipErrorInvalidValue;
    return c64_chain<0>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, work, s);
}

// ---- FP64: grids 192 (6 x 12, low-register, 256 VGPRs with a small spill), 160
// (5 x 10, low-register) and 128 (4 x 8), one workgroup per CU; 96 (3 x 6,
// low-register, 128 VGPRs) and 64 (2 x 4), two.  A larger FP64 matrix runs the
// blocked reduction first and hands over its trailing block at a panel boundary.
namespace {
constexpr int kNumGrids64 = 5;
constexpr int kGrid64[kNumGrids64] = {192, 160, 128, 96, 64};
int grid64_of(int n) {
    int g = 0;
    while (g < kNumGrids64 - 1 && n <= kGrid64[g + 1]) ++g;
    return g;
}
int split64(int n) {
    const int g = grid64_of(n);
    return g < kNumGrids64 - 1 ? kGrid64[g + 1] : 0;
}
constexpr int NL64_192 = 10, NL64_96 = 2;  // blocks held in LDS by the 192- and 96-wide FP64 grids
template <int IN>
void grid64_stage(int g, const double2* H, int64_t hk, int m, int ld, int64_t nk, double2* V, int ldv, int64_t vk,
                  double* dd, double* de, double2* dtau, int64_t dk, int s_end, double2* B, int ldb, hipStream_t s) {
    switch (g) {
        case 0: reg_stage<double, double2, 6, 12, 1, IN, double2, true, NL64_192>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 1: reg_stage<double, double2, 5, 10, 1, IN, double2, true>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 2: reg_stage<double, double2, 4, 8, 1, IN, double2>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        case 3: reg_stage<double, double2, 3, 6, 4, IN, double2, true, NL64_96>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
        default: reg_stage<double, double2, 2, 4, 4, IN, double2>(H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s); break;
    }
}
}  // namespace

int tridiag_reg_c128_tail(int n) {
    if (n < 2) return 0;
    if (n <= kGrid64[0]) return n;
    // the blocked reduction stops at a panel boundary s0 >= n - 192
    const int P = tridiag_blocked_panel((int)sizeof(double2));
    const int s0 = ((n - kGrid64[0] + P - 1) / P) * P;
    return n - s0;
}

size_t tridiag_reg_c128_work_bytes(int m) {
    const int t1 = split64(m);
    if (!t1) return 0;
    const int t2 = split64(t1);
    return sizeof(double2) * ((size_t)t1 * t1 + (size_t)t2 * t2);
}

hipError_t launch_tridiag_reg_c128(const double2* H, int64_t hk, int m, int ld, bool hermitian_full, int64_t nk,
                                   double2* V, int ldv, int64_t vk, double* dd, double* de, double2* dtau, int64_t dk,
                                   double2* work, hipStream_t s) {
    if (m < 1 || m > kGrid64[0] || nk <= 0) return hipErrorInvalidValue;
    // the first grid reads the block's rows (the conjugate upper triangle) when
    // the whole Hermitian matrix is there, its lower columns otherwise
    auto first = [&](int s_end, double2* B, int ldb) {
        if (hermitian_full)
            grid64_stage<1>(grid64_of(m), H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s);
        else
            grid64_stage<0>(grid64_of(m), H, hk, m, ld, nk, V, ldv, vk, dd, de, dtau, dk, s_end, B, ldb, s);
    };
    const int t1 = split64(m);
    if (!t1) {
        first(m, nullptr, 0);
        return hipGetLastError();
    }
    if (!work) return hipErrorInvalidValue;
    double2* buf[2] = {work, work + nk * (int64_t)t1 * t1};
    first(m - t1, buf[0], t1);
    int r = t1, cur = 0;
    while (true) {
        const int off = m - r;
        const int t = split64(r);
        grid64_stage<2>(grid64_of(r), buf[cur], (int64_t)r * r, r, r, nk, V + off + (int64_t)off * ldv, ldv, vk, dd + off,
                     de + off, dtau + off, dk, t ? r - t : r, t ? buf[cur ^ 1] : (double2*)nullptr, t, s);
        if (!t) break;
        r = t;
        cur ^= 1;
    }
    return hipGetLastError();
}

}  // namespace uepm
