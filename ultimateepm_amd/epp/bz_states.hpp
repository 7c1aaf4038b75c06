// bz_states.hpp — drop-in for uepm::mesh_bz::BZ_States
// (src/BZ_MESH/bz_states.hpp:20-96): eigenstates at every mesh vertex and at
// vertex + q, and the dielectric function integrated over the tetrahedra.
// The eigenpairs stay on the GPU inside the mesh object (uepm_bzmesh_
// compute_states); get_eigen_states copies them to the host on demand.
//
// Differences kept visible: the reference's eigenvalue vectors hold all N
// eigenvalues of each vertex; here they hold the nb_bands of set_nb_bands,
// the only ones the dielectric function reads.  nb_threads is accepted and
// ignored (one batched GPU solve replaces the OpenMP loop).
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "Vector3D.h"
#include "bz_mesh.hpp"
#include "eigen_types.h"

struct uepm_plan;

namespace uepm::mesh_bz {

class BZ_States : public MeshBZ {
 public:
    explicit BZ_States(const uepm::pseudopotential::Material& material) : MeshBZ(material) {}
    BZ_States(const BZ_States&) = delete;
    BZ_States& operator=(const BZ_States&) = delete;
    ~BZ_States();

    void set_nb_bands(int nb_bands) { m_nb_bands_total = nb_bands; }
    void set_basis_vectors(const std::vector<Vector3D<int>>& basis_vectors) { m_basisVectors = basis_vectors; }
    const std::vector<Vector3D<int>>& get_basis_vectors() const { return m_basisVectors; }
    void compute_eigenstates(int nb_threads = 1);
    void compute_shifted_eigenstates(const Vector3D<double>& q_shift, int nb_threads = 1);

    const std::vector<double>& get_energies() const { return m_list_energies; }
    void set_energies(const std::vector<double>& energies) { m_list_energies = energies; }

    // eigenvectors at the vertices: dim x nb_bands per vertex
    const std::vector<uepm::pseudopotential::MatrixXcd>& get_eigen_states() const;
    // eigenvalues at the vertices (k, or k + q when shifted): nb_bands per vertex
    const std::vector<uepm::pseudopotential::VectorXd>& get_eigenvalues(bool shifted = false) const;

    // A stub returning 0 in the reference (bz_states.cpp:166-169).
    const double compute_fermi_level(double doping_concentration, double temperature) const;

    void compute_dielectric_function(const std::vector<double>& energies, double eta_smearing, int nb_threads = 1);
    void export_dielectric_function(const std::string& prefix) const;
    const std::vector<double>& get_dielectric_function_real() const { return m_dielectric_function_real; }

    // A stub returning 1 in the reference (bz_states.hpp:87-90).
    std::complex<double> get_dielectric_function(const Vector3D<double>& q, double energy) const {
        (void)q;
        (void)energy;
        return 1.0;
    }

    void export_full_eigenstates() const;

 protected:
    std::vector<Vector3D<int>> m_basisVectors;
    int m_nb_bands_total = 0;
    Vector3D<double> m_q_shift;
    std::vector<double> m_list_energies;
    std::vector<double> m_dielectric_function_real;

 private:
    uepm_plan* m_plan = nullptr;
    int m_plan_nb = 0;
    std::vector<Vector3D<int>> m_plan_basis;
    mutable std::vector<uepm::pseudopotential::MatrixXcd> m_eigenvectors_k;
    mutable std::vector<uepm::pseudopotential::VectorXd> m_eigenvalues[2];
    mutable bool m_host_vectors = false, m_host_values[2] = {false, false};

    uepm_plan* plan();
};

}  // namespace uepm::mesh_bz
