// Oracle for the dielectric function.  TEST INFRASTRUCTURE ONLY: tests and
// bench.py's CPU baseline call it as the checker; the product path never does.
//
// Restates DielectricFunction::compute_dielectric_function
// (src/EPP/DielectricFunction.cpp:83-165) in its loop and summation order:
// per q (|q| <= 1e-15 -> (1e-15, 1e-15, 1e-15)) and k, eigenpairs at k and
// k + q (SetMatrix(k, nonlocal), SOC off, Diagonalize(true)), then for the
// conduction bands c = 4 .. nb-1 (outer) and valence bands v = 0 .. 3 (inner)
// the weight pow(abs(x_c . x_v), 2) with the bilinear product Eigen's
// col(c).adjoint().dot(col(v)) computes, times the two Lorentzians of every
// energy, summed into the k row, and the k rows summed in k order.
//
// Pinned (tests/test_dielectric_host.py): the free-electron limit (plane-wave
// eigenvectors, overlaps 0 or 1, closed-form energies); the bilinear form
// against a numpy evaluation from the oracle's own eigenvectors; q = 0 for a
// real-symmetric H (Si), where every interband overlap vanishes.
//
// Also restates BZ_States (src/BZ_MESH/bz_states.cpp): eigenstates at the
// mesh vertices and at vertex + q (:24-85) and the tetrahedron-weighted
// dielectric function (:95-164).  Pinned (tests/test_bzstates_host.py): the
// tetrahedron sum equals |V|/4 times the single-k sums above at its four
// vertices, with volumes from numpy determinants; vertex k-points against a
// direct solve; the prefactor against the SI constants.
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "../include/uepm.h"

extern "C" int orc_solve(const uepm_material* m, const int32_t* G, int N, unsigned flags, int nb, const double* k,
                         int64_t nk, double* evals, double* evecs, double* grads, int nthreads);

extern "C" int orc_dielectric(const uepm_material* m, const int32_t* G, int N, int nonlocal, int nb, const double* k,
                              int64_t nk, const double* q, int64_t nq, const double* energies, int n_energies,
                              double eta, double* out, int nthreads) {
    using cplx = std::complex<double>;
    constexpr int first_conduction = 4;
    if (nb <= first_conduction) return 1;
    const unsigned flags = UEPM_VECTORS | (nonlocal ? UEPM_NONLOCAL : 0u);
    std::vector<double> Ek(nk * nb), Xk(2 * nk * nb * (size_t)N), Ekq(nk * nb), Xkq(2 * nk * nb * (size_t)N);
    if (orc_solve(m, G, N, flags, nb, k, nk, Ek.data(), Xk.data(), nullptr, nthreads) != 0) return 2;
    std::vector<double> kq(3 * nk), list_k(nk * (size_t)n_energies);
    for (int64_t iq = 0; iq < nq; ++iq) {
        double qv[3] = {q[3 * iq], q[3 * iq + 1], q[3 * iq + 2]};
        if (std::sqrt(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2]) <= 1e-15) qv[0] = qv[1] = qv[2] = 1e-15;
        for (int64_t i = 0; i < nk; ++i)
            for (int c = 0; c < 3; ++c) kq[3 * i + c] = k[3 * i + c] + qv[c];
        if (orc_solve(m, G, N, flags, nb, kq.data(), nk, Ekq.data(), Xkq.data(), nullptr, nthreads) != 0) return 2;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t ik = 0; ik < nk; ++ik) {
            const cplx* xk = reinterpret_cast<const cplx*>(Xk.data()) + ik * nb * (int64_t)N;
            const cplx* xq = reinterpret_cast<const cplx*>(Xkq.data()) + ik * nb * (int64_t)N;
            double* row = list_k.data() + ik * n_energies;
            for (int e = 0; e < n_energies; ++e) row[e] = 0.0;
            for (int c = first_conduction; c < nb; ++c)
                for (int v = 0; v < first_conduction; ++v) {
                    cplx dotp = 0.0;
                    for (int i = 0; i < N; ++i) dotp += xq[c * (int64_t)N + i] * xk[v * (int64_t)N + i];
                    const double overlap_integral = std::pow(std::abs(dotp), 2);
                    const double delta_energy = Ekq[ik * nb + c] - Ek[ik * nb + v];
                    for (int e = 0; e < n_energies; ++e) {
                        const double energy = energies[e];
                        const double f1 =
                            (delta_energy - energy) / ((delta_energy - energy) * (delta_energy - energy) + eta * eta);
                        const double f2 =
                            (delta_energy + energy) / ((delta_energy + energy) * (delta_energy + energy) + eta * eta);
                        row[e] += overlap_integral * (f1 + f2);
                    }
                }
        }
        double* o = out + iq * n_energies;
        for (int e = 0; e < n_energies; ++e) o[e] = 0.0;
        for (int64_t ik = 0; ik < nk; ++ik)
            for (int e = 0; e < n_energies; ++e) o[e] += list_k[ik * n_energies + e];
    }
    return 0;
}

namespace {
inline double signed_volume(const double* P, const int32_t* v) {  // Tetra::compute_signed_volume, mesh_tetra.cpp:128-132
    const double* p0 = P + 3 * v[0];
    double e[3][3];
    for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c) e[j][c] = P[3 * v[j + 1] + c] - p0[c];
    const double cx = e[1][1] * e[2][2] - e[1][2] * e[2][1];
    const double cy = e[1][2] * e[2][0] - e[1][0] * e[2][2];
    const double cz = e[1][0] * e[2][1] - e[1][1] * e[2][0];
    return (1.0 / 6.0) * (e[0][0] * cx + e[0][1] * cy + e[0][2] * cz);
}
}  // namespace

// xyz: nv x 3 vertex positions in 1/m; tet: nt x 4; q: reduced units.
// eps, sums: n_energies; total_volume: sum of |V|.
extern "C" int orc_bz_states_dielectric(const uepm_material* m, const int32_t* G, int N, int nonlocal, int nb,
                                        const double* xyz, int64_t nv, const int32_t* tet, int64_t nt,
                                        const double* q, const double* energies, int n_energies, double eta,
                                        double* eps, double* sums, double* total_volume, int nthreads) {
    using cplx = std::complex<double>;
    constexpr int index_first_conduction_band = 4;
    constexpr double q_e = 1.602176634e-19, eps_0 = 8.8541878128e-12;  // physical_constants.hpp:20,31
    if (nb <= index_first_conduction_band) return 1;
    const unsigned flags = UEPM_VECTORS | (nonlocal ? UEPM_NONLOCAL : 0u);
    const double a = 1.0e-10 * m->lattice_constant_angstrom;
    const double normalization_factor = 2.0 * M_PI / a;
    std::vector<double> k(3 * nv), kq(3 * nv);
    for (int64_t i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            k[3 * i + c] = (xyz[3 * i + c] * 1.0) / normalization_factor;
            kq[3 * i + c] = k[3 * i + c] + q[c];
        }
    std::vector<double> Ek(nv * nb), Xk(2 * nv * nb * (size_t)N), Ekq(nv * nb), Xkq(2 * nv * nb * (size_t)N);
    if (orc_solve(m, G, N, flags, nb, k.data(), nv, Ek.data(), Xk.data(), nullptr, nthreads) != 0) return 2;
    if (orc_solve(m, G, N, flags, nb, kq.data(), nv, Ekq.data(), Xkq.data(), nullptr, nthreads) != 0) return 2;
    // overlaps and energy differences per vertex (what the tetrahedron loop evaluates at each of its vertices)
    const int npair = (nb - index_first_conduction_band) * index_first_conduction_band;
    std::vector<double> ov(nv * npair), de(nv * npair);
    for (int64_t ik = 0; ik < nv; ++ik) {
        const cplx* xk = reinterpret_cast<const cplx*>(Xk.data()) + ik * nb * (int64_t)N;
        const cplx* xq = reinterpret_cast<const cplx*>(Xkq.data()) + ik * nb * (int64_t)N;
        int p = 0;
        for (int c = index_first_conduction_band; c < nb; ++c)
            for (int v = 0; v < index_first_conduction_band; ++v, ++p) {
                cplx dotp = 0.0;
                for (int i = 0; i < N; ++i) dotp += xq[c * (int64_t)N + i] * xk[v * (int64_t)N + i];
                ov[ik * npair + p] = std::pow(std::abs(dotp), 2);
                de[ik * npair + p] = Ekq[ik * nb + c] - Ek[ik * nb + v];
            }
    }
    std::vector<double> total(n_energies, 0.0), tetra_sum(n_energies);
    double tv = 0.0;
    for (int64_t t = 0; t < nt; ++t) {
        const int32_t* vtx = tet + 4 * t;
        const double volume_tetra = std::fabs(signed_volume(xyz, vtx));
        tv += volume_tetra;
        for (int e = 0; e < n_energies; ++e) tetra_sum[e] = 0.0;
        for (int j = 0; j < 4; ++j) {
            const int64_t index_k = vtx[j];
            for (int p = 0; p < npair; ++p) {
                const double overlap_integral = ov[index_k * npair + p];
                const double delta_energy = de[index_k * npair + p];
                for (int e = 0; e < n_energies; ++e) {
                    const double energy = energies[e];
                    const double f1 =
                        (delta_energy - energy) / ((delta_energy - energy) * (delta_energy - energy) + eta * eta);
                    const double f2 =
                        (delta_energy + energy) / ((delta_energy + energy) * (delta_energy + energy) + eta * eta);
                    const double total_factor = f1 + f2;
                    tetra_sum[e] += overlap_integral * total_factor;
                }
            }
        }
        for (int e = 0; e < n_energies; ++e) {
            tetra_sum[e] *= volume_tetra * 0.25;
            total[e] += tetra_sum[e];
        }
    }
    const double qs[3] = {q[0] * normalization_factor, q[1] * normalization_factor, q[2] * normalization_factor};
    const double len = std::sqrt(qs[0] * qs[0] + qs[1] * qs[1] + qs[2] * qs[2]);
    const double q_squared = len * len;
    const double coulomb_prefactor_eV = (q_e * q_e) / (eps_0 * q_squared) / q_e;
    const double prefactor = coulomb_prefactor_eV * (2.0 / std::pow(2.0 * M_PI, 3));
    for (int e = 0; e < n_energies; ++e) {
        if (sums) sums[e] = total[e];
        eps[e] = 1.0 + prefactor * total[e];
    }
    if (total_volume) *total_volume = tv;
    return 0;
}
