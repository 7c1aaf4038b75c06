// bz_mesh.hpp — drop-in for the band-solve side of uepm::mesh_bz::MeshBZ
// (src/BZ_MESH/bz_mesh.hpp:96-360): the gmsh mesh of the Brillouin zone, bands
// over its vertices (BandsOnBZ), the k-star fan-out from the irreducible
// wedge, band views in and out, and the tetrahedron density of states
// (DOS_MeshBZ).  Everything runs in libuepm.so (uepm_bzmesh_*); the band
// solve goes through BandStructure, the DOS through the GPU kernel.  gmsh is
// not needed: MSH 4.1 ASCII is read and written directly.
//
// Not carried over: octree search, random k draws, electron-phonon and
// impact-ionisation rates (consumers outside the band solve), and the binary
// MSH format.  The eigenstates and dielectric function of BZ_States derive
// from this class (bz_states.hpp).
#pragma once

#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include "BandStructure.h"
#include "Material.h"

struct uepm_bzmesh;

namespace uepm::mesh_bz {

enum class MeshParticleType { valence, conduction };

class MeshBZ {
 public:
    MeshBZ() = default;
    explicit MeshBZ(const uepm::pseudopotential::Material& material) : m_material(material) {}
    MeshBZ(const MeshBZ&) = delete;
    MeshBZ& operator=(const MeshBZ&) = delete;
    MeshBZ(MeshBZ&& o) noexcept;
    MeshBZ& operator=(MeshBZ&& o) noexcept;
    ~MeshBZ();

    int get_number_threads_mesh_ops() const noexcept { return m_nb_threads_mesh_ops; }
    void set_number_threads_mesh_ops(int nb_threads) noexcept { m_nb_threads_mesh_ops = nb_threads; }

    // Reads the mesh, then the k-star file "<stem>_kstar_ibz_to_bz.txt" of the
    // working directory, which must exist (bz_mesh.cpp:67-191, 198-210).
    void read_mesh_geometry_from_msh_file(const std::string& filename, bool normalize_by_fourier_factor = true);
    // The file name argument is not used: the reference derives it from the mesh.
    void load_kstar_ibz_to_bz(const std::string& filename = "kstar_ibz_to_bz.txt");
    // Not in the reference (apps/bz_meshing.cpp writes it when it builds a mesh):
    // the k-star file of a mesh closed under the 48 operations of O_h.
    void write_kstar_ibz_to_bz(const std::string& filename) const;

    void compute_band_structure_over_mesh(uepm::pseudopotential::BandStructure& band_structure, bool use_iwedge = true);

    void export_selected_bands_to_gmsh(const std::string& out_filename, std::size_t nb_valence_to_export,
                                       std::size_t nb_conduction_to_export, bool highest_valence_as_band0,
                                       const std::string& model_name_or_msh_path, bool write_gradients) const;
    void export_energies_and_gradients_to_vtk(const std::string& filename) const;
    void read_mesh_bands_from_msh_file(const std::string& filename, int nb_conduction_bands = -1,
                                       int nb_valence_bands = -1, bool auto_shift_conduction_band = false,
                                       bool set_positive_valence_band = false);
    void apply_scissor(double scissor_value);

    double compute_mesh_volume() const;
    double compute_dos_at_energy_and_band(double iso_energy, int band_index, bool use_interp = false,
                                          bool use_iw = false) const;
    std::vector<std::vector<double>> compute_dos_band_at_band(int band_index, double min_energy, double max_energy,
                                                              std::size_t nb_points, bool use_interp = false,
                                                              bool use_iw = false) const;
    std::vector<std::vector<double>> compute_dos_band_at_band_auto(int band_index, std::size_t nb_points,
                                                                   bool use_interp = false,
                                                                   bool use_iw = false) const;

    double get_reduce_bz_factor() const noexcept { return m_reduce_bz_factor; }
    double get_volume() const noexcept { return m_total_volume; }
    double si_to_reduced_scale() const noexcept;
    std::size_t get_number_vertices() const noexcept { return m_nv; }
    std::size_t get_number_elements() const noexcept { return m_nt; }
    std::size_t get_number_vertices_in_irreducible_wedge() const;
    std::size_t get_number_bands_total() const;
    std::size_t get_number_valence_bands() const;
    std::size_t get_number_conduction_bands() const;
    std::size_t get_global_band_index(std::size_t local_band_index, MeshParticleType type) const;
    std::pair<double, double> get_min_max_energy_at_band(const int& band_index) const;
    // energies [vertex][band] and gradients [vertex][band] (eV m)
    std::vector<std::vector<double>> get_band_energies() const;
    std::vector<std::vector<Vector3D<double>>> get_energy_gradients() const;
    void print_band_info() const;

 protected:
    uepm::pseudopotential::Material m_material;
    int m_nb_threads_mesh_ops = 1;
    uepm_bzmesh* m_mesh = nullptr;
    std::size_t m_nv = 0, m_nt = 0;
    double m_total_volume = 0.0, m_reduce_bz_factor = 1.0;
    double m_spin_degeneracy = 2.0;

    uepm_bzmesh* mesh() const;
};

}  // namespace uepm::mesh_bz
