#include "BandStructure.h"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <fstream>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "device_context.h"

namespace uepm::pseudopotential {

std::string BandStructure::get_path_as_string() const {
    std::string s;
    for (const auto& p : m_path) s += p;
    return s;
}

bool BandStructure::GenerateBasisVectors(unsigned int nn) {
    int count = 0;
    if (uepm_basis_vectors(static_cast<int>(nn), nullptr, 0, &count) != UEPM_OK) {
        std::cout << "Error: " << uepm_last_error() << std::endl;
        return false;
    }
    std::vector<int32_t> g(3 * static_cast<size_t>(count));
    throw_on_error(uepm_basis_vectors(static_cast<int>(nn), g.data(), count, &count), "GenerateBasisVectors");
    basisVectors.resize(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) basisVectors[i] = Vector3D<int>(g[3 * i], g[3 * i + 1], g[3 * i + 2]);
    return true;
}

void BandStructure::Initialize(const Material& material, std::size_t nb_bands, const std::vector<std::string>& path,
                               unsigned int nbPoints, unsigned int nearestNeighborsNumber,
                               bool enable_non_local_correction, bool enable_soc) {
    m_material = material;
    m_nb_bands = static_cast<unsigned int>(nb_bands);
    m_path = path;
    m_nb_points = nbPoints;
    m_nearestNeighborsNumber = nearestNeighborsNumber;
    m_enable_non_local_correction = enable_non_local_correction;
    m_enable_spin_orbit_coupling = enable_soc;
    m_kpoints.clear();
    m_energies.clear();
    if (m_enable_spin_orbit_coupling) m_nb_bands *= 2;  // BandStructure.cpp:78-82
    if (!GenerateBasisVectors(nearestNeighborsNumber))
        throw std::runtime_error("BandStructure::Initialize: GenerateBasisVectors failed");
    m_kpoints = symmetryPoints.GeneratePoints(m_path, m_nb_points, symmetryPointsPositions);
    m_nb_points = static_cast<unsigned int>(m_kpoints.size());
    if (m_nb_points == 0) throw std::runtime_error("BandStructure::Initialize: GeneratePoints failed. No points generated.");
}

void BandStructure::Initialize(const Material& material, std::size_t nb_bands,
                               const std::vector<Vector3D<double>>& list_k_points, unsigned int nearestNeighborsNumber,
                               bool enable_non_local_correction, bool enable_soc) {
    m_material = material;
    m_nb_bands = static_cast<unsigned int>(nb_bands);
    m_nb_points = static_cast<unsigned int>(list_k_points.size());
    m_nearestNeighborsNumber = nearestNeighborsNumber;
    m_enable_non_local_correction = enable_non_local_correction;
    m_enable_spin_orbit_coupling = enable_soc;
    if (m_enable_spin_orbit_coupling) m_nb_bands *= 2;
    m_kpoints = list_k_points;
    m_energies.clear();
    if (!GenerateBasisVectors(nearestNeighborsNumber))
        throw std::runtime_error("BandStructure::Initialize: GenerateBasisVectors failed");
}

// One batched solve over every k-point.  Levels beyond the matrix dimension
// are left at 0 (Compute_parallel, BandStructure.cpp:181-190) or dropped
// (Compute, BandStructure.cpp:141-145: clamp_rows).
void BandStructure::solve(bool compute_gradient, bool clamp_rows) {
    const auto start = std::chrono::high_resolution_clock::now();
    m_nb_points = static_cast<unsigned int>(m_kpoints.size());
    const int N = static_cast<int>(basisVectors.size());
    const int dim = m_enable_spin_orbit_coupling ? 2 * N : N;
    const int nb = std::min<int>(static_cast<int>(m_nb_bands), dim);
    const size_t nk = m_kpoints.size();
    m_energies.assign(nk, std::vector<double>(clamp_rows ? nb : m_nb_bands, 0.0));
    m_energies_gradient.clear();
    if (compute_gradient) m_energies_gradient.assign(nk, std::vector<Vector3D<double>>(m_nb_bands));
    if (nk == 0 || nb == 0) return;
    // a non-finite k-point has no eigen-decomposition: the reference's Eigen
    // solver reports NoConvergence there (Hamiltonian.cpp:130-133)
    for (const auto& kp : m_kpoints)
        if (!std::isfinite(kp.X) || !std::isfinite(kp.Y) || !std::isfinite(kp.Z))
            throw std::runtime_error("Eigenvalue decomposition failed");

    const uepm_material mat = m_material.to_uepm();
    std::vector<int32_t> g(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        g[3 * i] = basisVectors[i].X;
        g[3 * i + 1] = basisVectors[i].Y;
        g[3 * i + 2] = basisVectors[i].Z;
    }
    unsigned flags = (m_enable_non_local_correction ? UEPM_NONLOCAL : 0u) |
                     (m_enable_spin_orbit_coupling ? UEPM_SOC : 0u) | (compute_gradient ? UEPM_GRADIENTS : 0u) |
                     (m_mixed ? UEPM_MIXED : 0u);
    uepm_plan* plan = nullptr;
    throw_on_error(uepm_plan_create(device_context(), &mat, g.data(), N, flags, nb, &plan), "BandStructure::Compute");
    std::vector<double> k(3 * nk), ev(nk * (size_t)nb), gr;
    for (size_t i = 0; i < nk; ++i) {
        k[3 * i] = m_kpoints[i].X;
        k[3 * i + 1] = m_kpoints[i].Y;
        k[3 * i + 2] = m_kpoints[i].Z;
    }
    if (compute_gradient) gr.resize(3 * nk * (size_t)nb);
    // eigenvectors are needed for the gradients only; let the plan keep them on the device
    const int st = uepm_solve(plan, k.data(), static_cast<int64_t>(nk), ev.data(), nullptr,
                              compute_gradient ? gr.data() : nullptr);
    uepm_plan_destroy(plan);
    if (st == UEPM_ERR_NOT_CONVERGED) throw std::runtime_error("Eigenvalue decomposition failed");
    throw_on_error(st, "BandStructure::Compute");
    for (size_t i = 0; i < nk; ++i) {
        std::copy(ev.begin() + (long)(i * nb), ev.begin() + (long)((i + 1) * nb), m_energies[i].begin());
        if (compute_gradient)
            for (int b = 0; b < nb; ++b) {
                const double* q = gr.data() + 3 * (i * nb + b);
                m_energies_gradient[i][b] = Vector3D<double>(q[0], q[1], q[2]);
            }
    }
    const auto end = std::chrono::high_resolution_clock::now();
    m_computation_time_s = std::chrono::duration<double>(end - start).count();
}

void BandStructure::Compute(bool compute_gradient) { solve(compute_gradient, true); }

void BandStructure::Compute_parallel(bool compute_gradient, int /*nb_threads*/) { solve(compute_gradient, false); }

double BandStructure::AdjustValues(bool minConductionBandToZero) {
    // BandStructure.cpp:203-262, through uepm_adjust_values (every row has the same length)
    if (m_energies.empty()) return 0.0;
    const size_t nb = m_energies.front().size();
    std::vector<double> flat;
    flat.reserve(m_energies.size() * nb);
    for (const auto& row : m_energies) flat.insert(flat.end(), row.begin(), row.end());
    double band_gap = 0.0;
    throw_on_error(uepm_adjust_values(flat.data(), static_cast<int64_t>(m_energies.size()), static_cast<int32_t>(nb),
                                      m_enable_spin_orbit_coupling ? 1 : 0, minConductionBandToZero ? 1 : 0, &band_gap),
                   "BandStructure::AdjustValues");
    for (size_t k = 0; k < m_energies.size(); ++k)
        std::copy(flat.begin() + (long)(k * nb), flat.begin() + (long)((k + 1) * nb), m_energies[k].begin());
    return band_gap;
}

std::vector<double> BandStructure::get_band(unsigned int band_index) const {
    std::vector<double> res;
    res.reserve(m_energies.size());
    for (const auto& p : m_energies) res.push_back(p[band_index]);
    return res;
}

void BandStructure::print_results() const {
    for (const auto& p : m_energies) {
        for (double v : p) std::cout << v << " ";
        std::cout << std::endl;
    }
}

void BandStructure::export_k_points_to_file(std::string filename) const {
    std::ofstream file(filename);
    for (const auto& p : m_kpoints) file << p.Y << " " << p.X << " " << p.Z << std::endl;
}

namespace {
void write_bands(std::ostream& file, const std::vector<double>& v) {
    for (size_t i = 0; i < v.size(); ++i) file << (i ? "," : "") << v[i];
    file << std::endl;
}
void write_header(std::ostream& file, size_t nbands) {
    for (size_t i = 0; i + 1 < nbands; ++i) file << "band_" << i << ",";
    file << "band_" << nbands - 1 << std::endl;
}
}  // namespace

void BandStructure::export_result_in_file(const std::string& filename) const {
    // BandStructure.cpp:290-314
    std::ofstream file(filename);
    file << "# Material " << m_material.get_name() << std::endl;
    file << "# NBands " << m_nb_bands << std::endl;
    file << "# Nonlocal " << (m_enable_non_local_correction ? "Yes" : "No") << std::endl;
    file << "# Path " << get_path_as_string() << std::endl;
    file << "# Symmetry points:";
    for (unsigned p : symmetryPointsPositions) file << " " << p;
    file << std::endl;
    if (m_energies.empty()) return;
    write_header(file, m_energies.front().size());
    for (unsigned k = 0; k < m_nb_points && k < m_energies.size(); ++k) write_bands(file, m_energies[k]);
}

void BandStructure::export_result_in_file_with_kpoints(const std::string& filename) const {
    std::ofstream file(filename);
    file << "kx,ky,kz,";
    if (m_energies.empty()) return;
    write_header(file, m_energies.front().size());
    for (unsigned k = 0; k < m_nb_points && k < m_energies.size(); ++k) {
        file << m_kpoints[k].Y << "," << m_kpoints[k].X << "," << m_kpoints[k].Z << ",";
        write_bands(file, m_energies[k]);
    }
}

std::string BandStructure::path_band_filename() const {
    // BandStructure.cpp:331-345
    const std::string path_string = get_path_as_string();
    return "EPM_" + m_material.get_name() + "_nb_bands_" + std::to_string(m_energies.front().size()) + "_path_" +
           path_string + "_size_basis_" + std::to_string(basisVectors.size());
}

void export_vector_bands_result_in_file(const std::string& filename, std::vector<std::vector<double>> results) {
    std::ofstream file(filename);
    for (const auto& p : results) {
        for (double v : p) file << v << ",";
        file << std::endl;
    }
}

}  // namespace uepm::pseudopotential
