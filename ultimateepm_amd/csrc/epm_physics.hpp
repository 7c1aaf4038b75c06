// epm_physics.hpp — per-element physics of the EPM Hamiltonian H(k) on gfx950.
//
// Device-side statement of the reference's element functions:
//   local V(G_i - G_j)      Pseudopotential::GetValue           src/EPP/Pseudopotential.cpp:24-56
//   kinetic diagonal        Hamiltonian::SetMatrix              src/EPP/Hamiltonian.cpp:61-74
//   non-local correction    Material::compute_pseudopotential_non_local_correction,
//                           F_l_function, F_2_function_gaussian src/EPP/Material.cpp:87-208
//   spherical Bessel j_l    generalized_bessel / J_neg_v        src/EPP/bessel_func.hpp:17-45
//   modified Bessel i_2     bessel_2nd_order_first_kind         src/EPP/bessel_func.hpp:47-52
//   spin-orbit 2x2 block    SpinOrbitCorrection                 src/EPP/SpinOrbitFunctional.cpp:20-138
//
// Expressions keep the reference's operation order and are compiled without
// FMA contraction (the pragma below), so an element differs from the
// reference only through the last-ulp behaviour of sin/cos/exp/sqrt and of
// the Bessel functions: those are evaluated here in closed form (with power
// series where the closed form cancels) instead of libstdc++'s
// std::cyl_bessel_j(l + 1/2, x), which the reference calls per element.  The
// one rearrangement is the non-local phase exp(2 pi i (tau/a) . (K_i - K_j)),
// formed as exp(i theta_i) exp(-i theta_j) from per-vector values (a few ulp,
// the size of the rounding of K_i - K_j itself).
//
// The per-basis-vector quantities (|K_i|, j_l(|K_i| R), SOC radial factors)
// are k-dependent but element-independent, so they are computed once per
// (k, i) and the O(N^2) element loop only combines them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace uepm {

// Everything an element needs, precomputed on the host from uepm_material
// with the reference's expressions (src/Common/physical_constants.hpp:18-41).
struct DevParams {
    int32_t N;          // basis size
    int32_t dim;        // N or 2N
    int32_t nl;         // non-local correction on
    int32_t soc;        // spin-orbit on
    int32_t well;       // 0 square, 1 gaussian (anything else: gaussian, as the reference's else-branch)
    int32_t nb;         // bands kept per k
    int32_t need_l0_a, need_l0_c;  // F_0 needed (R0 != 0) for anion / cation
    int32_t need_l2_a, need_l2_c;  // A2 != 0
    int32_t period_a, period_c;    // SOC radial-factor periods
    int32_t same_r0;    // need_l0_a && need_l0_c && R0_a == R0_c
    double a;           // lattice constant [m]
    double tau;         // a / 8 [m]
    double tpa;         // 2 pi / a
    double kin;         // (2 pi / a)^2 * hbar^2 / (2 m_e q_e): tpa * tpa * diag_factor
    double VS[4], VA[4];
    // non-local
    double fourier;     // 2 pi / a
    double diag_factor; // hbar^2 / (2 m_e q_e)
    double kF2;         // pow(kF, 2.0)
    double V_pre;       // 4 pi / Omega_at
    double tau_over_a;  // (a / 8) / a
    double alpha0_a, dfbeta_a, A2_a, R0_a, R2_a, gpre_a;  // dfbeta = diag_factor * beta_0
    double alpha0_c, dfbeta_c, A2_c, R0_c, R2_c, gpre_c;  // gpre = 5 pi^1.5 R2^3 / Omega
    // spin-orbit
    double soc_kfac;    // (2 pi) / a
    double xi_a, xi_c;  // radial extents
    double mu, alpha_mu;// mu_soc, alpha_soc * mu_soc
    double bohr;        // Bohr radius [m]
    // gradient
    double grad_pref;   // 2 * diag_factor * tpa^2
    double frac_to_phys;// 1 / tpa
};

#pragma clang fp contract(off)

// Spherical Bessel j_l(x), l in {-1, 0, 1, 2, 3}, x >= 0, with the special
// cases of generalized_bessel (bessel_func.hpp:29-45): x < 1e-14 gives 1 for
// l = 0 and 0 for l >= 1.  Closed forms (DLMF 10.49.3) above the cancellation
// region, the ascending series (DLMF 10.53.1) below it.
__device__ __forceinline__ double sph_series(int l, double x) {
    // x^l / (2l+1)!! * sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1))
    double dfact = 1.0;
    double xl = 1.0;
    for (int i = 1; i <= l; ++i) {
        dfact *= (double)(2 * i + 1);
        xl *= x;
    }
    const double h = -0.5 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 24; ++k) {
        term *= h / ((double)k * (double)(2 * l + 2 * k + 1));
        sum += term;
    }
    return xl / dfact * sum;
}

__device__ __forceinline__ double sph_j(int l, double x) {
    if (l == -1) return cos(x) / x;
    if (x < 1e-14) return l == 0 ? 1.0 : 0.0;
    const double lim = (l == 0) ? 0.25 : (double)l;
    if (x < lim) return sph_series(l, x);
    double s, c;
    sincos(x, &s, &c);
    const double ix = 1.0 / x;
    switch (l) {
        case 0: return s * ix;
        case 1: return (s * ix - c) * ix;
        case 2: return ((3.0 * ix * ix - 1.0) * s * ix - 3.0 * c * ix * ix);
        default: return ((15.0 * ix * ix * ix - 6.0 * ix) * s * ix - (15.0 * ix * ix - 1.0) * c * ix);
    }
}

// bessel_func.hpp:47-52
__device__ __forceinline__ double mod_i2(double x) {
    if (x == 0.0) return 0.0;
    return ((x * x + 3) * sinh(x) - 3 * x * cosh(x)) / (x * x * x);
}

// Per-(k, i) quantities for the radial integrals of one radius R:
// jm = j_{l-1}(|K| R), j0 = j_l(|K| R), jp = j_{l+1}(|K| R).
struct BesselTriple {
    double jm, j0, jp;
};

__device__ __forceinline__ BesselTriple bessel_triple(int l, double nK, double R) {
    const double x = nK * R;
    return {sph_j(l - 1, x), sph_j(l, x), sph_j(l + 1, x)};
}

// F_l_function (Material.cpp:87-109) from precomputed Bessel triples.
__device__ __forceinline__ double F_l(double n1, double n2, double R, int l, const BesselTriple& b1,
                                      const BesselTriple& b2) {
    constexpr double EPSILON = 1.0e-4;
    if (fabs(n1 - n2) > EPSILON) {
        const double pre = (R * R) / (n1 * n1 - n2 * n2);
        const double F = n1 * b1.jp * b2.j0 - n2 * b2.jp * b1.j0;
        return pre * F;
    } else if (n1 > EPSILON) {
        const double pre = (R * R * R) / (2.0);
        const double F = b1.j0 * b1.j0 - b1.jm * b1.jp;
        return pre * F;
    }
    return (l == 0) ? (R * R * R) / (3.0) : 0.0;
}

// F_2_function_gaussian (Material.cpp:111-116)
__device__ __forceinline__ double F_2_gauss(double n1, double n2, double R) {
    const double arg = 0.5 * (R * R) * n1 * n2;
    return mod_i2(arg) * exp(-0.25 * (n1 * n1 + n2 * n2) * R * R);
}

// Per-basis-vector state of one k-point.
struct KState {
    double kx, ky, kz;   // k + G_i, reduced
    double Kx, Ky, Kz;   // (k + G_i) * 2 pi / a
    double nK;           // |K|
    BesselTriple b0a, b0c, b2a, b2c;  // l = 0 / l = 2 triples at R0 / R2 of anion / cation
    double Ba, Bc;       // SOC radial factors B(|K|) of anion / cation
    double pc, ps;       // cos / sin of 2 pi (tau/a) (kx + ky + kz): the non-local phase of a
                         // pair (i, j) is exp(i (theta_i - theta_j))
};

// SpinOrbitFunctional.cpp:20-49 (period 2, 3, 4; others rejected on the host)
__device__ __forceinline__ double soc_radial(int period, double Klen, double bohr, double xi) {
    const double kappa = Klen * bohr / xi;
    const double q = 1.0 + kappa * kappa;
    if (period == 2) return 1.0 / (q * q * q);
    if (period == 3) return (5.0 - kappa * kappa) / (5.0 * (q * q * q * q));
    return (5.0 - 3.0 * kappa * kappa) / (5.0 * (q * q * q * q * q));
}

__device__ __forceinline__ void kstate_init(const DevParams& P, const double* k, const int32_t* G, int i, KState& s) {
    s.kx = k[0] + (double)G[3 * i];
    s.ky = k[1] + (double)G[3 * i + 1];
    s.kz = k[2] + (double)G[3 * i + 2];
    s.Kx = s.kx * P.fourier;
    s.Ky = s.ky * P.fourier;
    s.Kz = s.kz * P.fourier;
    s.nK = sqrt(s.Kx * s.Kx + s.Ky * s.Ky + s.Kz * s.Kz);
    if (P.nl) {
        sincos(2.0 * 3.14159265358979323846 *
                   (P.tau_over_a * s.kx + P.tau_over_a * s.ky + P.tau_over_a * s.kz),
               &s.ps, &s.pc);
        if (P.need_l0_a) s.b0a = bessel_triple(0, s.nK, P.R0_a);
        if (P.need_l0_c) s.b0c = P.same_r0 ? s.b0a : bessel_triple(0, s.nK, P.R0_c);
        if (P.well == 0) {
            if (P.need_l2_a) s.b2a = bessel_triple(2, s.nK, P.R2_a);
            if (P.need_l2_c) s.b2c = bessel_triple(2, s.nK, P.R2_c);
        }
    }
    if (P.soc) {
        // Kphys = kfac * K (SpinOrbitFunctional.cpp:109-110): same product as Kx above
        // because soc_kfac == fourier (both (2 pi) / a).
        const double px = s.kx * P.soc_kfac, py = s.ky * P.soc_kfac, pz = s.kz * P.soc_kfac;
        const double L = sqrt(px * px + py * py + pz * pz);
        s.Ba = soc_radial(P.period_a, L, P.bohr, P.xi_a);
        s.Bc = soc_radial(P.period_c, L, P.bohr, P.xi_c);
    }
}

// (cos, sin) of the phase of V(G) for integer G (Pseudopotential.cpp:24-56).
__device__ __forceinline__ double2 local_phase(const DevParams& P, int gx, int gy, int gz) {
    const double Gtau = P.tpa * (P.tau * (double)gx + P.tau * (double)gy + P.tau * (double)gz);
    double s, c;
    sincos(Gtau, &s, &c);
    return make_double2(c, s);
}

// The form factors vanish outside G^2 in {3, 4, 8, 11}, which bounds every
// component of G by 3: a table of local_phase over [-3, 3]^3 (343 entries,
// built on the device with the same expression) serves every pair.
constexpr int kVPhaseTab = 343;
__device__ __forceinline__ int local_phase_index(int gx, int gy, int gz) { return (gx + 3) * 49 + (gy + 3) * 7 + (gz + 3); }

// Pseudopotential::GetValue for integer G = G_i - G_j (Pseudopotential.cpp:24-56);
// tab: nullptr to evaluate the phase here, else the table above.
__device__ __forceinline__ void local_V(const DevParams& P, int gx, int gy, int gz, double& re, double& im,
                                        const double2* tab = nullptr) {
    const int G2 = gx * gx + gy * gy + gz * gz;
    double VS = 0.0, VA = 0.0;
    switch (G2) {
        case 3: VS = P.VS[0]; VA = P.VA[0]; break;
        case 4: VS = P.VS[1]; VA = P.VA[1]; break;
        case 8: VS = P.VS[2]; VA = P.VA[2]; break;
        case 11: VS = P.VS[3]; VA = P.VA[3]; break;
        default: re = 0.0; im = 0.0; return;  // VS = VA = 0: 0*cos = 0, 0*sin = 0
    }
    const double2 cs = tab ? tab[local_phase_index(gx, gy, gz)] : local_phase(P, gx, gy, gz);
    re = VS * cs.x;
    im = VA * cs.y;
}

// Material::compute_pseudopotential_non_local_correction (Material.cpp:139-208)
__device__ __forceinline__ void nonlocal(const DevParams& P, const KState& a, const KState& b, double& re, double& im) {
    const double n1 = a.nK, n2 = b.nK;
    double leg2 = 0.0;  // P_2(cos angle(K1, K2)), used by the l = 2 terms only
    if (P.need_l2_a || P.need_l2_c) {
        const double dp = a.Kx * b.Kx + a.Ky * b.Ky + a.Kz * b.Kz;
        const double np = n1 * n2;
        const double cosang = (np < 1.0e-13) ? 1.0 : dp / (n1 * n2);
        leg2 = 0.5 * (3 * cosang * cosang - 1);
    }
    const double nn = n1 * n2 - P.kF2;
    double Va = 0.0, Vc = 0.0;
    double F0a = 0.0;
    if (P.need_l0_a) {
        F0a = F_l(n1, n2, P.R0_a, 0, a.b0a, b.b0a);
        Va = P.V_pre * (P.alpha0_a + P.dfbeta_a * nn) * F0a;
    }
    if (P.need_l2_a) {
        double V2;
        if (P.well == 0) V2 = P.V_pre * P.A2_a * 5 * leg2 * F_l(n1, n2, P.R2_a, 2, a.b2a, b.b2a);
        else V2 = P.gpre_a * P.A2_a * leg2 * F_2_gauss(n1, n2, P.R2_a);
        Va += V2;
    }
    if (P.need_l0_c) {
        // the same radius gives the same F_0 (the Bessel triples are equal too)
        const double F0c = P.same_r0 ? F0a : F_l(n1, n2, P.R0_c, 0, a.b0c, b.b0c);
        Vc = P.V_pre * (P.alpha0_c + P.dfbeta_c * nn) * F0c;
    }
    if (P.need_l2_c) {
        double V2;
        if (P.well == 0) V2 = P.V_pre * P.A2_c * 5 * leg2 * F_l(n1, n2, P.R2_c, 2, a.b2c, b.b2c);
        else V2 = P.gpre_c * P.A2_c * leg2 * F_2_gauss(n1, n2, P.R2_c);
        Vc += V2;
    }
    const double Vs = 1.0 * (Va + Vc) / 2.0;
    const double Vas = 1.0 * (Va - Vc) / 2.0;
    // cos / sin of 2 pi (tau/a) sum_c (K_a - K_b)_c from the per-vector phases
    // (the reference forms the difference first; the two agree to rounding, and
    // the product form keeps the (b, a) value the exact conjugate)
    const double c = a.pc * b.pc + a.ps * b.ps;
    const double s = a.ps * b.pc - a.pc * b.ps;
    re = c * Vs;
    im = s * Vas;
}

// SpinOrbitCorrection::compute_soc_contribution (SpinOrbitFunctional.cpp:111-138):
// s[0..3] = (s00, s01, s10, s11) as (re, im) pairs; K = k + G reduced.
__device__ __forceinline__ void soc_block(const DevParams& P, const KState& a, const KState& b, int gx, int gy, int gz,
                                          double* s) {
    const double l1 = P.mu * a.Bc * b.Bc;
    const double l2 = P.alpha_mu * a.Ba * b.Ba;
    const double lsym = (l1 + l2) / 2.0;
    const double lanti = (l1 - l2) / 2.0;
    const double cz = a.kx * b.ky - a.ky * b.kx;
    const double cx = a.ky * b.kz - a.kz * b.ky;
    const double cy = a.kx * b.kz - a.kz * b.kx;
    const double Gtau = P.soc_kfac * (P.tau * (double)gx + P.tau * (double)gy + P.tau * (double)gz);
    double sn, cs;
    sincos(Gtau, &sn, &cs);
    const double pr = lanti * sn, pi = -lsym * cs;
    // (a + bi)(pr + i pi) = (a pr - b pi) + i (a pi + b pr)
    s[0] = cz * pr - 0.0 * pi;  s[1] = cz * pi + 0.0 * pr;
    s[2] = cx * pr - cy * pi;   s[3] = cx * pi + cy * pr;
    s[4] = cx * pr + cy * pi;   s[5] = cx * pi - cy * pr;
    const double mz = -(a.kx * b.ky - a.ky * b.kx);
    s[6] = mz * pr - 0.0 * pi;  s[7] = mz * pi + 0.0 * pr;
}

#pragma clang fp contract(on)

}  // namespace uepm
