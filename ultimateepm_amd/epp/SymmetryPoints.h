// SymmetryPoints.h — SymmetryPoints::GeneratePoints of the reference
// (src/EPP/SymmetryPoints.h:10-42, SymmetryPoints.cpp:21-147) over uepm_kpath.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "Vector3D.h"

namespace uepm::pseudopotential {

class SymmetryPoint {
 public:
    SymmetryPoint() = default;
    SymmetryPoint(const std::string& Name, const Vector3D<double> pos) : name(Name), position(pos) {}
    std::string name;
    Vector3D<double> position;
};

class SymmetryPoints {
 public:
    SymmetryPoints();
    std::map<std::string, SymmetryPoint> symmetryPoints;
    /* nrPoints points along `path` (single-letter names L, G, X, W, K, U);
     * every node is exact.  Throws std::invalid_argument for unknown names
     * or fewer points than nodes. */
    std::vector<Vector3D<double>> GeneratePoints(const std::vector<std::string>& path, unsigned int nrPoints,
                                                 std::vector<unsigned int>& symmetryPointsPositions);
};

}  // namespace uepm::pseudopotential
