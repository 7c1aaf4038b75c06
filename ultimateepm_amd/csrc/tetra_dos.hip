// Linear-tetrahedron density of states on a Brillouin-zone mesh.
//
// DOS(E) for one band = sum over tetrahedra of the iso-surface polygon area
// times the mean of 1/|grad E| at its corners, over (2 pi)^3
// (Tetra::compute_tetra_dos_energy_band, src/BZ_MESH/mesh_tetra.cpp:477-525,
// summed by MeshBZ::compute_dos_at_energy_and_band, src/BZ_MESH/bz_mesh.cpp:844-862,
// on the energy grid of compute_dos_band_at_band_auto, :889-916).
//
// Work split: one energy group of kEnergiesPerGroup samples x one slice of the
// tetrahedra per workgroup.  A per-band prepass stores each tetrahedron's
// sample-index window (8 bytes), so the scan that decides which tetrahedra
// cut an energy group reads 8 B per tetrahedron from L2; only the cutting
// ones load their corners, energies and gradients.  Partial sums per
// (slice, energy) go to a workspace and a second kernel adds the slices in
// slice order: the result does not depend on scheduling.
//
// Every expression follows the reference's operation order with contraction
// off, so a tetrahedron's contribution rounds as the reference's does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "solver.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kEnergiesPerGroup = 8;

struct V3 {
    double x, y, z;
};
__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 scale(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline V3 divide(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline double norm(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ inline double triple(V3 a, V3 b, V3 c) { return dot(a, cross(b, c)); }
__device__ inline V3 load3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// (1/6) * (e1 . (e2 x e3)) scaled back by 6, as Tetra::compute_signed_volume
// (mesh_tetra.cpp:128-132) and compute_barycentric_coordinates (:173-191) do.
__global__ void tetra_det_kernel(const double* __restrict__ pos, const int32_t* __restrict__ tet, int64_t nt,
                                 double* __restrict__ det) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const V3 p0 = load3(pos, tet[4 * t]);
    const V3 e0 = sub(load3(pos, tet[4 * t + 1]), p0);
    const V3 e1 = sub(load3(pos, tet[4 * t + 2]), p0);
    const V3 e2 = sub(load3(pos, tet[4 * t + 3]), p0);
    const double signed_volume = (1.0 / 6.0) * triple(e0, e1, e2);
    det[t] = 6.0 * signed_volume;
}

// Vertex-major (nv x ld) energies and (nv x ld x 3) gradients -> band-major.
__global__ void bands_to_band_major_kernel(const double* __restrict__ E, const double* __restrict__ G, int64_t nv,
                                           int ld, int first, int nb, double* __restrict__ Eb,
                                           double* __restrict__ Gb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nv * nb) return;
    const int64_t v = i / nb;
    const int b = static_cast<int>(i - v * nb);
    Eb[b * nv + v] = E[v * ld + first + b];
    if (G) {
        const double* g = G + 3 * (v * ld + first + b);
        double* o = Gb + 3 * (b * nv + v);
        o[0] = g[0];
        o[1] = g[1];
        o[2] = g[2];
    }
}

// Sample-index window [lo, hi] of every tetrahedron for one band.  It is a
// filter only (one sample of slack on each side); the exact test
// E < min || E > max is repeated per sample in the DOS kernel.
__global__ void tetra_window_kernel(const int32_t* __restrict__ tet, const uint8_t* __restrict__ in_wedge, int64_t nt,
                                    const double* __restrict__ Eb, double e0, double step, int n_points,
                                    int2* __restrict__ window) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int2 w = make_int2(1, 0);  // empty
    if (!in_wedge || in_wedge[t]) {
        const double a = Eb[tet[4 * t]], b = Eb[tet[4 * t + 1]], c = Eb[tet[4 * t + 2]], d = Eb[tet[4 * t + 3]];
        const double mn = fmin(fmin(a, b), fmin(c, d)), mx = fmax(fmax(a, b), fmax(c, d));
        if (mn <= mx) {
            const double flo = floor((mn - e0) / step) - 1.0, fhi = ceil((mx - e0) / step) + 1.0;
            if (fhi >= 0.0 && flo <= static_cast<double>(n_points - 1)) {
                w.x = flo < 0.0 ? 0 : static_cast<int>(flo);
                w.y = fhi > static_cast<double>(n_points - 1) ? n_points - 1 : static_cast<int>(fhi);
            }
        }
    }
    window[t] = w;
}

struct TetraBand {
    V3 p[4];      // corner positions (1/m)
    V3 g[4];      // corner gradients of the band (eV m)
    V3 edge[3];   // p1-p0, p2-p0, p3-p0
    double e[4];  // corner energies (eV)
    double det;   // 6 x signed volume
    double emin, emax;
    double es[4];  // corner energies, ascending
    V3 ps[4];      // corners in that order
};

__device__ __forceinline__ void load_tetra(TetraBand& T, const double* pos, const int32_t* tet, const double* det,
                                  const double* Eb, const double* Gb, int64_t t) {
    int e_idx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        e_idx[j] = tet[4 * t + j];
        T.p[j] = load3(pos, e_idx[j]);
        T.g[j] = load3(Gb, e_idx[j]);
        T.e[j] = Eb[e_idx[j]];
    }
    T.edge[0] = sub(T.p[1], T.p[0]);
    T.edge[1] = sub(T.p[2], T.p[0]);
    T.edge[2] = sub(T.p[3], T.p[0]);
    T.det = det[t];
    T.emin = T.e[0];
    T.emax = T.e[0];
#pragma unroll
    for (int j = 1; j < 4; ++j) {
        if (T.e[j] < T.emin) T.emin = T.e[j];
        if (T.emax < T.e[j]) T.emax = T.e[j];
    }
    // five-comparator network (Tetra::pre_compute_sorted_slots_per_band,
    // mesh_tetra.cpp:287-315); the corners move with their energies, so the
    // sorted data needs no dynamic indexing.
    double v[4] = {T.e[0], T.e[1], T.e[2], T.e[3]};
    V3 q[4] = {T.p[0], T.p[1], T.p[2], T.p[3]};
#define UEPM_CSWAP(a, b)                \
    if (v[a] > v[b]) {                  \
        const double tv = v[a];         \
        v[a] = v[b];                    \
        v[b] = tv;                      \
        const V3 tq = q[a];             \
        q[a] = q[b];                    \
        q[b] = tq;                      \
    }
    UEPM_CSWAP(0, 1)
    UEPM_CSWAP(2, 3)
    UEPM_CSWAP(0, 2)
    UEPM_CSWAP(1, 3)
    UEPM_CSWAP(1, 2)
#undef UEPM_CSWAP
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T.es[j] = v[j];
        T.ps[j] = q[j];
    }
}

// Contribution of one tetrahedron at one energy (states / (eV m^3), before the
// mesh-level factors).
__device__ __forceinline__ double tetra_dos_at(const TetraBand& T, double energy) {
    if (energy < T.emin || energy > T.emax) return 0.0;
    const double e_0 = T.es[0], e_1 = T.es[1], e_2 = T.es[2], e_3 = T.es[3];
    const V3 P0 = T.ps[0], P1 = T.ps[1], P2 = T.ps[2], P3 = T.ps[3];
    // iso-energy polygon (Tetra::compute_band_iso_energy_surface, mesh_tetra.cpp:339-395)
    V3 iso[4];
    int n;
    iso[3] = {0.0, 0.0, 0.0};
    if (energy < e_1 && energy >= e_0) {
        const double lU = (energy - e_0) / (e_1 - e_0);
        const double lV = (energy - e_0) / (e_2 - e_0);
        const double lW = (energy - e_0) / (e_3 - e_0);
        iso[0] = add(scale(1.0 - lU, P0), scale(lU, P1));
        iso[1] = add(scale(1.0 - lV, P0), scale(lV, P2));
        iso[2] = add(scale(1.0 - lW, P0), scale(lW, P3));
        n = 3;
    } else if (energy < e_2 && energy >= e_1) {
        const double lU = (energy - e_0) / (e_2 - e_0);
        const double lV = (energy - e_0) / (e_3 - e_0);
        const double lW = (e_2 - energy) / (e_2 - e_1);
        const double lX = (energy - e_1) / (e_3 - e_1);
        iso[0] = add(scale(1.0 - lU, P0), scale(lU, P2));
        iso[1] = add(scale(1.0 - lV, P0), scale(lV, P3));
        iso[2] = add(scale(lW, P1), scale(1.0 - lW, P2));
        iso[3] = add(scale(1.0 - lX, P1), scale(lX, P3));
        n = 4;
    } else {
        const double lU = (e_3 - energy) / (e_3 - e_2);
        const double lV = (e_3 - energy) / (e_3 - e_1);
        const double lW = (e_3 - energy) / (e_3 - e_0);
        iso[0] = add(scale(lU, P2), scale(1.0 - lU, P3));
        iso[1] = add(scale(lV, P1), scale(1.0 - lV, P3));
        iso[2] = add(scale(lW, P0), scale(1.0 - lW, P3));
        n = 3;
    }
    // cyclic order around the centroid (order_cyclic, mesh_tetra.cpp:402-450).
    // Loops run to 4 with a guard so every array index is static (registers,
    // no scratch); the stable sort by angle becomes a stable rank.
    V3 c = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < n) c = add(c, iso[i]);
    c = divide(c, static_cast<double>(n));
    V3 nrm = cross(sub(iso[1], iso[0]), sub(iso[2], iso[0]));
    const double len = norm(nrm);
    if (len > 0.0) nrm = divide(nrm, len);
    V3 u = sub(iso[0], c);
    const double ulen = norm(u);
    if (ulen > 0.0) {
        u = divide(u, ulen);
    } else {
        u = {1.0, 0.0, 0.0};
    }
    const V3 v = cross(nrm, u);
    double ang[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const V3 d = sub(iso[i], c);
        ang[i] = i < n ? atan2(dot(d, v), dot(d, u)) : 0.0;
    }
    int rank[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int r = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < n && j != i && (ang[j] < ang[i] || (j < i && !(ang[i] < ang[j])))) ++r;
        rank[i] = r;
    }
    V3 q[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        q[p] = iso[0];
#pragma unroll
        for (int i = 1; i < 4; ++i)
            if (i < n && rank[i] == p) q[p] = iso[i];
    }
    // polygon area (polygon_area, mesh_tetra.cpp:462-474)
    V3 sum = {0.0, 0.0, 0.0};
    sum = add(sum, cross(q[0], q[1]));
    sum = add(sum, cross(q[1], q[2]));
    if (n == 4) {
        sum = add(sum, cross(q[2], q[3]));
        sum = add(sum, cross(q[3], q[0]));
    } else {
        sum = add(sum, cross(q[2], q[0]));
    }
    const double A = 0.5 * norm(sum);
    if (!(A > 0.0)) return 0.0;
    // mean of 1/|grad E| at the polygon corners, gradients interpolated
    // barycentrically (interpolate_gradient_energy_at_band, mesh_tetra.cpp:214-230)
    double inv_grad_sum = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= n) break;
        const V3 v1 = sub(iso[i], T.p[0]);
        const double l2 = triple(v1, T.edge[1], T.edge[2]) / T.det;
        const double l3 = triple(v1, T.edge[2], T.edge[0]) / T.det;
        const double l4 = triple(v1, T.edge[0], T.edge[1]) / T.det;
        const double l1 = 1.0 - l2 - l3 - l4;
        V3 g = {0.0, 0.0, 0.0};
        g = add(g, scale(l1, T.g[0]));
        g = add(g, scale(l2, T.g[1]));
        g = add(g, scale(l3, T.g[2]));
        g = add(g, scale(l4, T.g[3]));
        const double gn = norm(g);
        const double g_min = 1e-12;
        inv_grad_sum += 1.0 / (gn < g_min ? g_min : gn);
    }
    const double inv_grad_avg = inv_grad_sum / static_cast<double>(n);
    const double pref = 1.0 / (8.0 * M_PI * M_PI * M_PI);
    return pref * (A * inv_grad_avg);
}

__device__ inline double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__global__ void __launch_bounds__(kThreads) tetra_dos_kernel(
    const double* __restrict__ pos, const int32_t* __restrict__ tet, const double* __restrict__ det,
    const int2* __restrict__ window, int64_t nt, int64_t per_slice, const double* __restrict__ Eb,
    const double* __restrict__ Gb, double e0, double step, int n_points, double* __restrict__ partial) {
    const int g0 = blockIdx.x * kEnergiesPerGroup;
    const int g1 = min(g0 + kEnergiesPerGroup, n_points) - 1;
    const int64_t t_begin = blockIdx.y * per_slice;
    const int64_t t_end = min(nt, t_begin + per_slice);
    double acc[kEnergiesPerGroup];
#pragma unroll
    for (int j = 0; j < kEnergiesPerGroup; ++j) acc[j] = 0.0;
    for (int64_t t = t_begin + threadIdx.x; t < t_end; t += kThreads) {
        const int2 w = window[t];
        const int lo = max(w.x, g0), hi = min(w.y, g1);
        if (lo > hi) continue;
        TetraBand T;
        load_tetra(T, pos, tet, det, Eb, Gb, t);
        for (int i = lo; i <= hi; ++i) {
            const double energy = e0 + static_cast<double>(i) * step;
            const double c = tetra_dos_at(T, energy);
#pragma unroll
            for (int j = 0; j < kEnergiesPerGroup; ++j)
                if (j == i - g0) acc[j] += c;
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
    if (threadIdx.x < kEnergiesPerGroup && g0 + (int)threadIdx.x < n_points) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += part[w][threadIdx.x];
        partial[blockIdx.y * (int64_t)n_points + g0 + threadIdx.x] = s;
    }
}

// Adds the slices in order and applies the mesh factors in the reference's
// order (bz_mesh.cpp:855-860).
__global__ void tetra_dos_finish_kernel(const double* __restrict__ partial, int slices, int n_points, double e0,
                                        double step, double reduce_bz_factor, double spin_degeneracy,
                                        double wedge_factor, double* __restrict__ energies, double* __restrict__ dos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_points) return;
    double s = 0.0;
    for (int k = 0; k < slices; ++k) s += partial[(int64_t)k * n_points + i];
    s *= reduce_bz_factor;
    s *= spin_degeneracy;
    s *= wedge_factor;
    if (energies) energies[i] = e0 + static_cast<double>(i) * step;
    dos[i] = s;
}

}  // namespace

namespace uepm {

hipError_t launch_tetra_det(const double* pos, const int32_t* tet, int64_t nt, double* det, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((nt + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(tetra_det_kernel, dim3(blocks), dim3(kThreads), 0, s, pos, tet, nt, det);
    return hipGetLastError();
}

hipError_t launch_bands_to_band_major(const double* E, const double* G, int64_t nv, int ld, int first, int nb,
                                      double* Eb, double* Gb, hipStream_t s) {
    const int64_t n = nv * nb;
    if (n <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(bands_to_band_major_kernel, dim3(blocks), dim3(kThreads), 0, s, E, G, nv, ld, first, nb, Eb,
                       Gb);
    return hipGetLastError();
}

int tetra_dos_slices(int64_t nt, int n_points) {
    // 1024 tetrahedra per slice (at most 256 slices), from nt alone: the
    // order in which a sample's contributions are added then does not depend
    // on the energy grid, and one energy gives the same bits alone or in a grid.
    (void)n_points;
    int64_t slices = (nt + 1023) / 1024;
    if (slices < 1) slices = 1;
    if (slices > 256) slices = 256;
    return static_cast<int>(slices);
}

hipError_t launch_tetra_dos(const double* pos, const int32_t* tet, const double* det, const uint8_t* in_wedge,
                            int64_t nt, const double* Eb, const double* Gb, int n_points, double e_min, double e_max,
                            double reduce_bz_factor, double spin_degeneracy, double wedge_factor, int2* window,
                            double* partial, double* energies, double* dos, hipStream_t s) {
    const double step = (e_max - e_min) / static_cast<double>(n_points - 1);
    const int slices = tetra_dos_slices(nt, n_points);
    if (nt > 0) {
        const unsigned tb = static_cast<unsigned>((nt + kThreads - 1) / kThreads);
        hipLaunchKernelGGL(tetra_window_kernel, dim3(tb), dim3(kThreads), 0, s, tet, in_wedge, nt, Eb, e_min, step,
                           n_points, window);
        const int64_t per_slice = (nt + slices - 1) / slices;
        const unsigned groups = static_cast<unsigned>((n_points + kEnergiesPerGroup - 1) / kEnergiesPerGroup);
        hipLaunchKernelGGL(tetra_dos_kernel, dim3(groups, slices), dim3(kThreads), 0, s, pos, tet, det, window, nt,
                           per_slice, Eb, Gb, e_min, step, n_points, partial);
    }
    const unsigned fb = static_cast<unsigned>((n_points + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(tetra_dos_finish_kernel, dim3(fb), dim3(kThreads), 0, s, partial, nt > 0 ? slices : 0,
                       n_points, e_min, step, reduce_bz_factor, spin_degeneracy, wedge_factor, energies, dos);
    return hipGetLastError();
}

}  // namespace uepm
