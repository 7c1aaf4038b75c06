#include "SymmetryPoints.h"

#include "device_context.h"

namespace uepm::pseudopotential {

SymmetryPoints::SymmetryPoints() {
    // fcc high-symmetry points in units of 2*pi/a (SymmetryPoints.cpp:21-28)
    symmetryPoints["L"] = SymmetryPoint("L", Vector3D<double>(0.5, 0.5, 0.5));
    symmetryPoints["G"] = SymmetryPoint("G", Vector3D<double>(0., 0., 0.));
    symmetryPoints["X"] = SymmetryPoint("X", Vector3D<double>(1., 0., 0.));
    symmetryPoints["W"] = SymmetryPoint("W", Vector3D<double>(1., 0.5, 0.));
    symmetryPoints["K"] = SymmetryPoint("K", Vector3D<double>(0.75, 0.75, 0.));
    symmetryPoints["U"] = SymmetryPoint("U", Vector3D<double>(1., 0.25, 0.25));
}

std::vector<Vector3D<double>> SymmetryPoints::GeneratePoints(const std::vector<std::string>& path, unsigned int nrPoints,
                                                             std::vector<unsigned int>& symmetryPointsPositions) {
    std::string labels;
    for (const auto& p : path) {
        if (p.size() != 1) throw std::invalid_argument("SymmetryPoints::GeneratePoints: unknown point " + p);
        labels += p;
    }
    std::vector<double> k(3 * static_cast<size_t>(nrPoints > 0 ? nrPoints : 1));
    std::vector<unsigned> pos(labels.size() ? labels.size() : 1);
    int count = 0;
    throw_on_error(uepm_kpath(labels.c_str(), nrPoints, k.data(), pos.data(), &count), "SymmetryPoints::GeneratePoints");
    std::vector<Vector3D<double>> out(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) out[i] = Vector3D<double>(k[3 * i], k[3 * i + 1], k[3 * i + 2]);
    symmetryPointsPositions.assign(pos.begin(), pos.begin() + static_cast<long>(labels.size()));
    return out;
}

}  // namespace uepm::pseudopotential
