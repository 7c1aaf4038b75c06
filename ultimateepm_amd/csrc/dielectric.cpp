// uepm_dielectric*: the dielectric function from eigenpairs at k and k + q
// (DielectricFunction, src/EPP/DielectricFunction.cpp).  The eigenpairs come
// from the plan's batched solve; overlaps and the Lorentzian sums run in
// dielectric.hip.  merge and Kramers-Kronig are host loops.
#include <cmath>
#include <string>
#include <vector>

#include "uepm_internal.hpp"

#pragma clang fp contract(off)

using uepm_host::DeviceBuffer;
using uepm_host::fail;

namespace {
constexpr int kValence = 4;
constexpr int64_t kChunk = 8192;  // k-points per solve pair: 2 x 8192 x nb x dim x 16 B of eigenvectors
}  // namespace

extern "C" {

int uepm_dielectric(uepm_plan* p, const double* k, int64_t nk, const double* q, int64_t nq, const double* energies,
                    int32_t n_energies, double eta_smearing, double* out) {
    if (!p || (nk > 0 && !k) || (nq > 0 && !q) || (n_energies > 0 && !energies) || nk < 0 || nq < 0 || n_energies < 0 ||
        (nq > 0 && n_energies > 0 && !out))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_dielectric: bad argument");
    if (!(p->flags & UEPM_VECTORS) || (p->flags & UEPM_SOC))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_dielectric: the plan needs UEPM_VECTORS and no UEPM_SOC");
    if (p->nb <= kValence)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_dielectric: the plan needs more than 4 bands (4 valence bands)");
    for (int64_t i = 0; i < nq * n_energies; ++i) out[i] = 0.0;
    if (nk == 0 || nq == 0 || n_energies == 0) return UEPM_OK;
    HIP_OK(hipSetDevice(p->ctx->device));
    const int nb = p->nb, dim = p->dim;
    const int64_t chunk = nk < kChunk ? nk : kChunk;
    const int npair = (nb - kValence) * kValence;
    DeviceBuffer dk, dkq, Ek, Ekq, Xk, Xkq, w, d, part, sum, den;
    HIP_OK(dk.ensure(sizeof(double) * 3 * chunk));
    HIP_OK(dkq.ensure(sizeof(double) * 3 * chunk));
    HIP_OK(Ek.ensure(sizeof(double) * nb * chunk));
    HIP_OK(Ekq.ensure(sizeof(double) * nb * chunk));
    HIP_OK(Xk.ensure(sizeof(double2) * (size_t)nb * dim * chunk));
    HIP_OK(Xkq.ensure(sizeof(double2) * (size_t)nb * dim * chunk));
    HIP_OK(w.ensure(sizeof(double) * npair * chunk));
    HIP_OK(d.ensure(sizeof(double) * npair * chunk));
    HIP_OK(part.ensure(sizeof(double) * (size_t)uepm::dielectric_slices(npair * chunk) * n_energies));
    HIP_OK(sum.ensure(sizeof(double) * nq * n_energies));
    HIP_OK(den.ensure(sizeof(double) * n_energies));
    hipStream_t s = nullptr;
    HIP_OK(hipMemcpy(den.p, energies, sizeof(double) * n_energies, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(sum.p, out, sizeof(double) * nq * n_energies, hipMemcpyHostToDevice));  // zeros
    // q -> (1e-15, 1e-15, 1e-15) when |q| <= 1e-15 (DielectricFunction.cpp:95-98)
    std::vector<double> qv(3 * nq);
    for (int64_t iq = 0; iq < nq; ++iq) {
        const double* qq = q + 3 * iq;
        const double len = std::sqrt(qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2]);
        for (int c = 0; c < 3; ++c) qv[3 * iq + c] = len <= 1e-15 ? 1e-15 : qq[c];
    }
    std::vector<double> kq(3 * chunk);
    for (int64_t c0 = 0; c0 < nk; c0 += chunk) {
        const int64_t n = (nk - c0 < chunk) ? nk - c0 : chunk;
        HIP_OK(hipMemcpyAsync(dk.p, k + 3 * c0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        int st = uepm_solve_device(p, dk.as<double>(), n, Ek.as<double>(), Xk.as<double>(), nullptr, s);
        if (st != UEPM_OK) return st;
        for (int64_t iq = 0; iq < nq; ++iq) {
            HIP_OK(hipStreamSynchronize(s));  // kq is rewritten below
            for (int64_t i = 0; i < n; ++i)
                for (int c = 0; c < 3; ++c) kq[3 * i + c] = k[3 * (c0 + i) + c] + qv[3 * iq + c];
            HIP_OK(hipMemcpyAsync(dkq.p, kq.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
            st = uepm_solve_device(p, dkq.as<double>(), n, Ekq.as<double>(), Xkq.as<double>(), nullptr, s);
            if (st != UEPM_OK) return st;
            HIP_OK(uepm::launch_dielectric_accumulate(Xk.as<double2>(), Xkq.as<double2>(), Ek.as<double>(),
                                                      Ekq.as<double>(), nullptr, nb, dim, n, den.as<double>(), n_energies,
                                                      eta_smearing, w.as<double>(), d.as<double>(),
                                                      part.as<double>(), sum.as<double>() + iq * n_energies, s));
        }
    }
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(out, sum.p, sizeof(double) * nq * n_energies, hipMemcpyDeviceToHost));
    return UEPM_OK;
}

// DielectricFunction::merge_results renormalisation (DielectricFunction.cpp:167-214):
// eps = 1 + (2 pi / |q|^2) S / total_nk, |q|^2 = 1e-14 when q = 0.
int uepm_dielectric_merge(const double* q, int64_t nq, int32_t n_energies, int64_t total_nk, double* eps) {
    if ((nq > 0 && (!q || !eps)) || nq < 0 || n_energies < 0 || total_nk <= 0)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_dielectric_merge: bad argument");
    const double renormalization = 1.0 / static_cast<double>(total_nk);
    for (int64_t iq = 0; iq < nq; ++iq) {
        const double* qq = q + 3 * iq;
        double q_squared = std::pow(std::sqrt(qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2]), 2);
        if (q_squared == 0.0) q_squared = 1e-14;
        for (int32_t e = 0; e < n_energies; ++e) {
            double& v = eps[iq * n_energies + e];
            v *= renormalization;
            v = 1.0 + (2.0 * M_PI / q_squared) * v;
        }
    }
    return UEPM_OK;
}

// DielectricFunction::apply_kramers_kronig (DielectricFunction.cpp:216-250) with
// create_kramers_matrix (:200-214): eps_imag = K (eps_real - 1),
// K_ij = -(2/pi) i / (j^2 - i^2) for i != j, 0 on the diagonal.
int uepm_kramers_kronig(const double* eps_real, int64_t nq, int32_t n_energies, double* eps_imag) {
    if ((nq > 0 && n_energies > 0 && (!eps_real || !eps_imag)) || nq < 0 || n_energies < 0)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_kramers_kronig: bad argument");
    const int64_t n = n_energies;
    std::vector<double> K((size_t)n * n);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
            K[i * n + j] = i == j ? 0.0
                                  : (double(i) / double(static_cast<long>(j * j) - static_cast<long>(i * i))) *
                                        (-2.0 / M_PI);
    std::vector<double> x(n);
    for (int64_t iq = 0; iq < nq; ++iq) {
        for (int64_t j = 0; j < n; ++j) x[j] = eps_real[iq * n + j] - 1.0;
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t j = 0; j < n; ++j) s += K[i * n + j] * x[j];
            eps_imag[iq * n + i] = s;
        }
    }
    return UEPM_OK;
}

}  // extern "C"
