// DielectricFunction.hpp — drop-in for uepm::pseudopotential::DielectricFunction
// (src/EPP/DielectricFunction.hpp:19-208).  compute_dielectric_function runs
// the eigen-solves at k and k + q, the overlaps and the Lorentzian sums on the
// GPU (uepm_dielectric); merge_results and apply_kramers_kronig use the host
// routines of the same library.  Eigen's per-k eigenvector matrices are not
// kept: the device holds them for the length of the call.
#pragma once

#include <string>
#include <vector>

#include "Material.h"
#include "Vector3D.h"

namespace uepm::pseudopotential {

bool is_in_irreducible_wedge(const Vector3D<double>& k);
bool is_in_first_BZ(const Vector3D<double>& k, bool one_eighth = false);

class DielectricFunction {
 protected:
    std::vector<Vector3D<int>> m_basisVectors;
    std::vector<Vector3D<double>> m_kpoints;
    const Material& m_material;
    const int m_nb_bands;
    bool m_nonlocal_epm = false;
    std::vector<Vector3D<double>> m_qpoints;
    std::vector<double> m_energies;
    std::string m_export_prefix = "dielectric_function";
    std::size_t m_offset_k_index = 0;
    std::size_t m_nb_kpoints = 0;
    std::vector<std::vector<double>> m_dielectric_function_real;
    std::vector<std::vector<double>> m_dielectric_function_imag;

 public:
    DielectricFunction(const Material& material, const std::vector<Vector3D<int>>& basisVectors, const int nb_bands);
    DielectricFunction(const DielectricFunction&) = default;

    void set_non_local_epm(const bool new_value) { m_nonlocal_epm = new_value; }
    // std::random_device seeded, as the reference (not reproducible run to run)
    void generate_k_points_random(std::size_t nb_points);
    void generate_k_points_grid(std::size_t Nx, std::size_t Ny, std::size_t Nz, double shift, bool irreducible_wedge);
    const std::vector<Vector3D<double>>& get_kpoints() const { return m_kpoints; }
    std::size_t get_offset_k_index() const { return m_offset_k_index; }
    std::size_t get_nb_kpoints() const { return m_nb_kpoints; }
    void set_offset_k_index(std::size_t offset_k_index) { m_offset_k_index = offset_k_index; }
    void set_nb_kpoints(std::size_t nb_kpoints) { m_nb_kpoints = nb_kpoints; }
    void set_qpoints(const std::vector<Vector3D<double>>& qpoints) { m_qpoints = qpoints; }
    void set_energies(const std::vector<double>& energies) { m_energies = energies; }
    void set_export_prefix(const std::string& prefix) { m_export_prefix = prefix; }

    // Sums over k in [offset, offset + nb_kpoints); one row per q is appended.
    void compute_dielectric_function(double eta_smearing = 1e-2, int mpi_rank = 0);
    void clear_eigen_states() {}

    const std::vector<std::vector<double>>& get_dielectric_function() const { return m_dielectric_function_real; }
    const std::vector<std::vector<double>>& get_dielectric_function_imag() const { return m_dielectric_function_imag; }
    const std::vector<double> get_flat_dielectric_function() const;

    static DielectricFunction merge_results(DielectricFunction RootDielectricFunction,
                                            const std::vector<std::vector<std::vector<double>>>& dielectric_function_results,
                                            std::vector<int> nb_kpoints_per_instance);
    // The reference also writes its 6 x 6 debug matrix to Bkramers_matrix.dat; not here.
    void apply_kramers_kronig();
    void export_kpoints(const std::string& filename) const;
    void export_dielectric_function_at_q(const std::string& filename, std::size_t idx_q, bool name_auto) const;
    void export_dielectric_function(const std::string& filename, bool name_auto) const;
};

}  // namespace uepm::pseudopotential
