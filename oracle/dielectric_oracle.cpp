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
