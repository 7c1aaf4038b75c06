/*
 * ref_harness.cpp — extern "C" entry points onto the reference's own
 * standalone-compilable sources, built into oracle/_ref/ by oracle/Makefile
 * straight from /root/reference (nothing of the reference is copied here).
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle (epm_oracle.cpp) and to
 * write golden vectors under tests/golden/.  Only these reference files build
 * without Eigen / yaml-cpp / gmsh (absent from this image):
 *   src/EPP/bessel_func.hpp        (generalized_bessel, i_2)
 *   src/EPP/Pseudopotential.cpp    (Pseudopotential::GetValue)
 *   src/EPP/SymmetryPoints.cpp     (SymmetryPoints::GeneratePoints)
 */
#include <cstring>
#include <string>
#include <vector>

#include "Pseudopotential.h"
#include "SymmetryPoints.h"
#include "bessel_func.hpp"

using uepm::pseudopotential::Pseudopotential;
using uepm::pseudopotential::SymmetryPoints;

extern "C" {

double ref_generalized_bessel(double nu, double x) { return generalized_bessel(nu, x); }
double ref_generalized_bessel_cylindrical(double nu, double x) { return generalized_bessel_cylindrical(nu, x); }
double ref_i2(double x) { return bessel_2nd_order_first_kind(x); }

/* V(G) for integer G with the reference's tau = a/8 (1,1,1) and a in metres. */
void ref_local_V(const double* VSA /* V3S V4S V8S V11S V3A V4A V8A V11A */, const int* G, double a, double* out) {
    Pseudopotential p(VSA[0], VSA[1], VSA[2], VSA[3], VSA[4], VSA[5], VSA[6], VSA[7]);
    constexpr double one_eight = 1.0 / 8.0;
    Vector3D<double> tau{one_eight * a, one_eight * a, one_eight * a};
    auto z = p.GetValue(Vector3D<int>(G[0], G[1], G[2]), tau, a);
    out[0] = z.real();
    out[1] = z.imag();
}

/* Path through single-letter labels; returns the count or -1. */
int ref_kpath(const char* labels, unsigned nrPoints, double* out, unsigned* sympos) {
    std::vector<std::string> path;
    for (const char* c = labels; *c; ++c) path.emplace_back(1, *c);
    SymmetryPoints sp;
    std::vector<unsigned> pos;
    std::vector<Vector3D<double>> pts;
    try {
        pts = sp.GeneratePoints(path, nrPoints, pos);
    } catch (...) {
        return -1;
    }
    for (size_t i = 0; i < pts.size(); ++i) {
        out[3 * i] = pts[i].X;
        out[3 * i + 1] = pts[i].Y;
        out[3 * i + 2] = pts[i].Z;
    }
    for (size_t i = 0; i < pos.size(); ++i) sympos[i] = pos[i];
    return (int)pts.size();
}
}
