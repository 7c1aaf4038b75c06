// BandStructure.h — drop-in for uepm::pseudopotential::BandStructure
// (src/EPP/BandStructure.h:12-106).  Compute and Compute_parallel run the
// whole k-set as one batched solve on the GPU (the thread count is accepted
// and ignored); results, AdjustValues and the export formats follow the
// reference (BandStructure.cpp:59-345).
#pragma once

#include <string>
#include <vector>

#include "Material.h"
#include "SymmetryPoints.h"
#include "Vector3D.h"

namespace uepm::pseudopotential {

class BandStructure {
 public:
    BandStructure() = default;

    void Initialize(const Material& material, std::size_t nb_bands, const std::vector<std::string>& path,
                    unsigned int nrPoints, unsigned int nearestNeighborsNumber, bool enable_non_local_correction,
                    bool enable_soc);
    void Initialize(const Material& material, std::size_t nb_bands, const std::vector<Vector3D<double>>& list_k_points,
                    unsigned int nearestNeighborsNumber, bool enable_non_local_correction, bool enable_soc);

    const std::vector<std::string>& GetPath() const { return m_path; }
    std::string get_path_as_string() const;
    unsigned int GetPointsNumber() const { return static_cast<unsigned int>(m_kpoints.size()); }
    void Compute(bool compute_gradient = false);
    void Compute_parallel(bool compute_gradient = false, int nb_threads = 1);
    double AdjustValues(bool minConductionBandToZero = false);

    void print_results() const;
    std::string path_band_filename() const;
    void export_k_points_to_file(std::string filename) const;
    void export_result_in_file(const std::string& filename) const;
    void export_result_in_file_with_kpoints(const std::string& filename) const;

    unsigned int get_number_of_bands() const { return m_nb_bands; }
    std::vector<double> get_band(unsigned int band_index) const;
    double get_energy_at_k_band(unsigned int band_index, unsigned int index_k) const { return m_energies[index_k][band_index]; }

    const std::vector<Vector3D<int>>& get_basis_vectors() const { return basisVectors; }
    const std::vector<Vector3D<double>>& get_kpoints() const { return m_kpoints; }
    const std::vector<std::vector<double>>& get_band_energies() const { return m_energies; }
    const std::vector<std::vector<Vector3D<double>>>& get_band_energy_gradients() const { return m_energies_gradient; }
    void set_kpoints(const std::vector<Vector3D<double>>& kpoints) { m_kpoints = kpoints; }

    double get_computation_time_s() const { return m_computation_time_s; }
    bool is_soc_enabled() const { return m_enable_spin_orbit_coupling; }

    /* Not in the reference: solve in the precision form (complex64 reduction +
     * FP64 refinement, eigenvalues within 0.1 meV of FP64) for nb_bands <= 24. */
    void set_mixed_precision(bool on) { m_mixed = on; }

 private:
    SymmetryPoints symmetryPoints;
    std::vector<unsigned int> symmetryPointsPositions;
    std::vector<std::string> m_path;
    unsigned int m_nb_points = 0;

    Material m_material;
    unsigned int m_nb_bands = 10;
    unsigned int m_nearestNeighborsNumber = 10;
    bool m_enable_non_local_correction = false;
    bool m_enable_spin_orbit_coupling = false;
    bool m_mixed = false;

    std::vector<Vector3D<int>> basisVectors;
    std::vector<Vector3D<double>> m_kpoints;
    std::vector<std::vector<double>> m_energies;
    std::vector<std::vector<Vector3D<double>>> m_energies_gradient;

    double m_computation_time_s = -1.0;

    bool GenerateBasisVectors(unsigned int nearestNeighborsNumber);
    void solve(bool compute_gradient, bool clamp_rows);
};

void export_vector_bands_result_in_file(const std::string& filename, std::vector<std::vector<double>> results);

}  // namespace uepm::pseudopotential
