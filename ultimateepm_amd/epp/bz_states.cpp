// bz_states.cpp — BZ_States drop-in over uepm_bzmesh_compute_states /
// uepm_bzmesh_states / uepm_bzmesh_states_dielectric (semantics in
// include/uepm.h, file:line of src/BZ_MESH/bz_states.cpp at each entry point).
#include "bz_states.hpp"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "device_context.h"
#include "uepm.h"

namespace uepm::mesh_bz {

using uepm::pseudopotential::MatrixXcd;
using uepm::pseudopotential::throw_on_error;
using uepm::pseudopotential::VectorXd;

BZ_States::~BZ_States() {
    if (m_plan) uepm_plan_destroy(m_plan);
}

// Hamiltonian(m_material, m_basisVectors) per thread with SetMatrix(k, false)
// and Diagonalize(true) (bz_states.cpp:30-47): one plan, local, SOC off,
// m_nb_bands_total eigenvector columns.
uepm_plan* BZ_States::plan() {
    if (m_basisVectors.empty()) throw std::invalid_argument("BZ_States: no basis vectors (set_basis_vectors)");
    if (m_nb_bands_total <= 0) throw std::invalid_argument("BZ_States: no band count (set_nb_bands)");
    const int N = static_cast<int>(m_basisVectors.size());
    if (m_plan) {
        if (m_plan_nb == m_nb_bands_total && m_plan_basis == m_basisVectors) return m_plan;
        uepm_plan_destroy(m_plan);
        m_plan = nullptr;
    }
    const uepm_material mat = m_material.to_uepm();
    std::vector<int32_t> g(3 * static_cast<size_t>(N));
    for (int i = 0; i < N; ++i) {
        g[3 * i] = m_basisVectors[i].X;
        g[3 * i + 1] = m_basisVectors[i].Y;
        g[3 * i + 2] = m_basisVectors[i].Z;
    }
    throw_on_error(uepm_plan_create(uepm::pseudopotential::device_context(), &mat, g.data(), N, UEPM_VECTORS,
                                    m_nb_bands_total, &m_plan),
                   "BZ_States::compute_eigenstates");
    m_plan_nb = m_nb_bands_total;
    m_plan_basis = m_basisVectors;
    return m_plan;
}

void BZ_States::compute_eigenstates(int /*nb_threads*/) {
    throw_on_error(uepm_bzmesh_compute_states(mesh(), plan(), nullptr), "BZ_States::compute_eigenstates");
    m_host_vectors = false;
    m_host_values[0] = false;
}

void BZ_States::compute_shifted_eigenstates(const Vector3D<double>& q_shift, int /*nb_threads*/) {
    const double normalization_factor = 2.0 * M_PI / m_material.get_lattice_constant_meter();
    m_q_shift = q_shift * normalization_factor;  // kept in 1/m like the reference (bz_states.cpp:55-57)
    const double q[3] = {q_shift.X, q_shift.Y, q_shift.Z};
    throw_on_error(uepm_bzmesh_compute_states(mesh(), plan(), q), "BZ_States::compute_shifted_eigenstates");
    m_host_values[1] = false;
}

const std::vector<MatrixXcd>& BZ_States::get_eigen_states() const {
    if (m_host_vectors) return m_eigenvectors_k;
    int32_t nb = 0, dim = 0;
    throw_on_error(uepm_bzmesh_states(mesh(), 0, nullptr, nullptr, &nb, &dim), "BZ_States::get_eigen_states");
    std::vector<double> buf(2 * m_nv * static_cast<size_t>(nb) * dim);
    throw_on_error(uepm_bzmesh_states(mesh(), 0, nullptr, buf.data(), nullptr, nullptr), "BZ_States::get_eigen_states");
    m_eigenvectors_k.assign(m_nv, MatrixXcd());
    for (std::size_t v = 0; v < m_nv; ++v) {
        MatrixXcd& M = m_eigenvectors_k[v];
        M.resize(dim, nb);
        const double* x = buf.data() + 2 * v * static_cast<size_t>(nb) * dim;
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i < dim; ++i) M(i, j) = {x[2 * (j * dim + i)], x[2 * (j * dim + i) + 1]};
    }
    m_host_vectors = true;
    return m_eigenvectors_k;
}

const std::vector<VectorXd>& BZ_States::get_eigenvalues(bool shifted) const {
    const int s = shifted ? 1 : 0;
    if (m_host_values[s]) return m_eigenvalues[s];
    int32_t nb = 0;
    throw_on_error(uepm_bzmesh_states(mesh(), s, nullptr, nullptr, &nb, nullptr), "BZ_States::get_eigenvalues");
    std::vector<double> buf(m_nv * static_cast<size_t>(nb));
    throw_on_error(uepm_bzmesh_states(mesh(), s, buf.data(), nullptr, nullptr, nullptr), "BZ_States::get_eigenvalues");
    m_eigenvalues[s].assign(m_nv, VectorXd());
    for (std::size_t v = 0; v < m_nv; ++v) {
        m_eigenvalues[s][v].resize(nb);
        for (int j = 0; j < nb; ++j) m_eigenvalues[s][v](j) = buf[v * nb + j];
    }
    m_host_values[s] = true;
    return m_eigenvalues[s];
}

const double BZ_States::compute_fermi_level(double /*doping_concentration*/, double /*temperature*/) const {
    return 0.0;
}

void BZ_States::compute_dielectric_function(const std::vector<double>& list_energies, double eta_smearing,
                                            int /*nb_threads*/) {
    m_list_energies = list_energies;
    m_dielectric_function_real.assign(list_energies.size(), 0.0);
    double total_volume = 0.0;
    throw_on_error(uepm_bzmesh_states_dielectric(mesh(), list_energies.data(), static_cast<int32_t>(list_energies.size()),
                                                 eta_smearing, m_dielectric_function_real.data(), nullptr,
                                                 &total_volume),
                   "BZ_States::compute_dielectric_function");
    // the reference's report (bz_states.cpp:142-162)
    const double a = m_material.get_lattice_constant_meter();
    const double V_BZ = std::pow(2.0 * M_PI, 3) / (a * a * a);
    std::cout << "Total volume (k-space) integrated: " << total_volume << std::endl;
    std::cout << "Expected BZ volume: " << V_BZ << "\n";
    std::cout << "Integrated volume: " << total_volume << "\n";
    std::cout << "Ratio (integrated / expected): " << (total_volume / V_BZ) << "\n";
    if (!m_dielectric_function_real.empty()) std::cout << "EPS[0] = " << m_dielectric_function_real[0] << std::endl;
}

void BZ_States::export_dielectric_function(const std::string& prefix) const {
    throw_on_error(uepm_write_dielectric_csv(prefix.c_str(), m_list_energies.data(), m_dielectric_function_real.data(),
                                             static_cast<int32_t>(m_dielectric_function_real.size())),
                   "BZ_States::export_dielectric_function");
}

void BZ_States::export_full_eigenstates() const {
    throw_on_error(uepm_bzmesh_export_states(mesh(), nullptr), "BZ_States::export_full_eigenstates");
}

}  // namespace uepm::mesh_bz
