// bz_mesh.cpp — MeshBZ drop-in over the uepm_bzmesh_* C ABI.  Method
// semantics follow src/BZ_MESH/bz_mesh.cpp (file:line in include/uepm.h at
// each entry point); this file maps calls, containers and exceptions.
#include "bz_mesh.hpp"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "device_context.h"
#include "uepm.h"

namespace uepm::mesh_bz {

using uepm::pseudopotential::throw_on_error;

MeshBZ::MeshBZ(MeshBZ&& o) noexcept { *this = std::move(o); }

MeshBZ& MeshBZ::operator=(MeshBZ&& o) noexcept {
    if (this != &o) {
        if (m_mesh) uepm_bzmesh_destroy(m_mesh);
        m_material = o.m_material;
        m_nb_threads_mesh_ops = o.m_nb_threads_mesh_ops;
        m_mesh = o.m_mesh;
        m_nv = o.m_nv;
        m_nt = o.m_nt;
        m_total_volume = o.m_total_volume;
        m_reduce_bz_factor = o.m_reduce_bz_factor;
        m_spin_degeneracy = o.m_spin_degeneracy;
        o.m_mesh = nullptr;
    }
    return *this;
}

MeshBZ::~MeshBZ() {
    if (m_mesh) uepm_bzmesh_destroy(m_mesh);
}

uepm_bzmesh* MeshBZ::mesh() const {
    if (!m_mesh) throw std::runtime_error("MeshBZ: no mesh read yet");
    return m_mesh;
}

double MeshBZ::si_to_reduced_scale() const noexcept {
    return m_material.get_lattice_constant_meter() / (2.0 * M_PI);
}

void MeshBZ::read_mesh_geometry_from_msh_file(const std::string& filename, bool normalize_by_fourier_factor) {
    // A host without a device still reads meshes and writes files; the DOS
    // then reports the missing device.
    uepm_ctx* ctx = nullptr;
    try {
        ctx = uepm::pseudopotential::device_context();
    } catch (const std::runtime_error&) {
        ctx = nullptr;
    }
    uepm_bzmesh* m = nullptr;
    throw_on_error(uepm_bzmesh_read_msh(ctx, filename.c_str(), m_material.get_lattice_constant_meter(),
                                        normalize_by_fourier_factor ? 1 : 0, &m),
                   "MeshBZ::read_mesh_geometry_from_msh_file");
    if (m_mesh) uepm_bzmesh_destroy(m_mesh);
    m_mesh = m;
    int64_t nv = 0, nt = 0;
    uepm_bzmesh_counts(m, &nv, &nt, nullptr);
    m_nv = static_cast<std::size_t>(nv);
    m_nt = static_cast<std::size_t>(nt);
    uepm_bzmesh_geometry(m, nullptr, nullptr, nullptr, nullptr, &m_total_volume, &m_reduce_bz_factor);
    load_kstar_ibz_to_bz();
}

void MeshBZ::load_kstar_ibz_to_bz(const std::string& /*filename*/) {
    throw_on_error(uepm_bzmesh_load_kstar(mesh(), nullptr), "MeshBZ::load_kstar_ibz_to_bz");
}

void MeshBZ::write_kstar_ibz_to_bz(const std::string& filename) const {
    throw_on_error(uepm_bzmesh_write_kstar(mesh(), filename.c_str()), "MeshBZ::write_kstar_ibz_to_bz");
}

void MeshBZ::compute_band_structure_over_mesh(uepm::pseudopotential::BandStructure& band_structure, bool use_iwedge) {
    int64_t n = 0;
    throw_on_error(uepm_bzmesh_kpoints(mesh(), use_iwedge ? 1 : 0, nullptr, &n), "MeshBZ::compute_band_structure_over_mesh");
    std::vector<double> k(3 * static_cast<size_t>(n));
    uepm_bzmesh_kpoints(m_mesh, use_iwedge ? 1 : 0, k.data(), &n);
    std::vector<Vector3D<double>> kpoints(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) kpoints[i] = Vector3D<double>(k[3 * i], k[3 * i + 1], k[3 * i + 2]);
    band_structure.set_kpoints(kpoints);
    const bool compute_gradient = true;
    band_structure.Compute_parallel(compute_gradient, m_nb_threads_mesh_ops);
    const bool set_cond_band_zero = false;
    band_structure.AdjustValues(set_cond_band_zero);
    const auto& E = band_structure.get_band_energies();
    const auto& G = band_structure.get_band_energy_gradients();
    const int nb = E.empty() ? 0 : static_cast<int>(E.front().size());
    std::vector<double> e(static_cast<size_t>(n) * nb), g(3 * static_cast<size_t>(n) * nb);
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < nb; ++b) {
            e[i * nb + b] = E[i][b];
            g[3 * (i * nb + b)] = G[i][b].X;
            g[3 * (i * nb + b) + 1] = G[i][b].Y;
            g[3 * (i * nb + b) + 2] = G[i][b].Z;
        }
    throw_on_error(uepm_bzmesh_set_mesh_bands(m_mesh, use_iwedge ? 1 : 0, e.data(), g.data(), nb),
                   "MeshBZ::compute_band_structure_over_mesh");
    print_band_info();
}

void MeshBZ::export_selected_bands_to_gmsh(const std::string& out_filename, std::size_t nb_valence_to_export,
                                           std::size_t nb_conduction_to_export, bool highest_valence_as_band0,
                                           const std::string& model_name_or_msh_path, bool write_gradients) const {
    throw_on_error(uepm_bzmesh_export_msh(mesh(), out_filename.c_str(), static_cast<int64_t>(nb_valence_to_export),
                                          static_cast<int64_t>(nb_conduction_to_export), highest_valence_as_band0 ? 1 : 0,
                                          model_name_or_msh_path.c_str(), write_gradients ? 1 : 0),
                   "MeshBZ::export_selected_bands_to_gmsh");
}

void MeshBZ::export_energies_and_gradients_to_vtk(const std::string& filename) const {
    throw_on_error(uepm_bzmesh_export_vtk(mesh(), filename.c_str()), "MeshBZ::export_energies_and_gradients_to_vtk");
}

void MeshBZ::read_mesh_bands_from_msh_file(const std::string& filename, int nb_conduction_bands, int nb_valence_bands,
                                           bool auto_shift_conduction_band, bool set_positive_valence_band) {
    throw_on_error(uepm_bzmesh_read_bands_msh(mesh(), filename.c_str(), nb_conduction_bands, nb_valence_bands,
                                              auto_shift_conduction_band ? 1 : 0, set_positive_valence_band ? 1 : 0),
                   "MeshBZ::read_mesh_bands_from_msh_file");
}

void MeshBZ::apply_scissor(double scissor_value) {
    throw_on_error(uepm_bzmesh_apply_scissor(mesh(), scissor_value), "MeshBZ::apply_scissor");
}

double MeshBZ::compute_mesh_volume() const { return m_total_volume; }

double MeshBZ::compute_dos_at_energy_and_band(double iso_energy, int band_index, bool use_interp, bool use_iw) const {
    if (use_interp) return 0.0;  // Tetra::interpolate_dos_at_energy_per_band returns 0 (mesh_tetra.cpp:578-581)
    double e[2], d[2];
    throw_on_error(uepm_bzmesh_dos(mesh(), band_index, 2, iso_energy, iso_energy + 1.0 + std::fabs(iso_energy), m_reduce_bz_factor,
                                   m_spin_degeneracy, use_iw ? 1 : 0, e, d),
                   "MeshBZ::compute_dos_at_energy_and_band");
    return d[0];
}

std::vector<std::vector<double>> MeshBZ::compute_dos_band_at_band(int band_index, double min_energy, double max_energy,
                                                                  std::size_t nb_points, bool /*use_interp*/,
                                                                  bool /*use_iw*/) const {
    // The reference evaluates compute_dos_at_energy_and_band(energy, band_index)
    // with its defaults here, whatever use_interp and use_iw say (bz_mesh.cpp:876).
    std::vector<double> e(nb_points), d(nb_points);
    throw_on_error(uepm_bzmesh_dos(mesh(), band_index, static_cast<int32_t>(nb_points), min_energy, max_energy,
                                   m_reduce_bz_factor, m_spin_degeneracy, 0, e.data(), d.data()),
                   "MeshBZ::compute_dos_band_at_band");
    return {e, d};
}

std::vector<std::vector<double>> MeshBZ::compute_dos_band_at_band_auto(int band_index, std::size_t nb_points,
                                                                       bool use_interp, bool use_iw) const {
    std::vector<double> e(nb_points), d(nb_points);
    throw_on_error(uepm_bzmesh_dos_auto(mesh(), band_index, static_cast<int32_t>(nb_points), use_interp ? 1 : 0,
                                        use_iw ? 1 : 0, e.data(), d.data()),
                   "MeshBZ::compute_dos_band_at_band_auto");
    return {e, d};
}

std::size_t MeshBZ::get_number_vertices_in_irreducible_wedge() const {
    int64_t n = 0;
    uepm_bzmesh_wedge_vertices(mesh(), nullptr, &n);
    return static_cast<std::size_t>(n);
}

std::size_t MeshBZ::get_number_bands_total() const {
    int32_t nb = 0;
    uepm_bzmesh_counts(mesh(), nullptr, nullptr, &nb);
    return static_cast<std::size_t>(nb);
}

std::size_t MeshBZ::get_number_valence_bands() const {
    int64_t nv = 0;
    uepm_bzmesh_band_groups(mesh(), &nv, nullptr, nullptr, nullptr);
    return static_cast<std::size_t>(nv);
}

std::size_t MeshBZ::get_number_conduction_bands() const {
    int64_t nc = 0;
    uepm_bzmesh_band_groups(mesh(), nullptr, nullptr, &nc, nullptr);
    return static_cast<std::size_t>(nc);
}

std::size_t MeshBZ::get_global_band_index(std::size_t local_band_index, MeshParticleType type) const {
    int64_t nv = 0, vs = 0, nc = 0, cs = 0;
    uepm_bzmesh_band_groups(mesh(), &nv, &vs, &nc, &cs);
    if (type == MeshParticleType::valence) {
        if (local_band_index >= static_cast<std::size_t>(nv))
            throw std::out_of_range("Local valence band index out of range.");
        return static_cast<std::size_t>(vs) + local_band_index;
    }
    if (local_band_index >= static_cast<std::size_t>(nc))
        throw std::out_of_range("Local conduction band index out of range.");
    return static_cast<std::size_t>(cs) + local_band_index;
}

std::pair<double, double> MeshBZ::get_min_max_energy_at_band(const int& band_index) const {
    double mn = 0.0, mx = 0.0;
    throw_on_error(uepm_bzmesh_band_info(mesh(), band_index, nullptr, nullptr, &mn, &mx),
                   "MeshBZ::get_min_max_energy_at_band");
    return {mn, mx};
}

std::vector<std::vector<double>> MeshBZ::get_band_energies() const {
    int32_t nb = 0;
    throw_on_error(uepm_bzmesh_bands(mesh(), nullptr, nullptr, &nb), "MeshBZ::get_band_energies");
    std::vector<double> e(m_nv * static_cast<size_t>(nb));
    uepm_bzmesh_bands(m_mesh, e.data(), nullptr, &nb);
    std::vector<std::vector<double>> out(m_nv);
    for (size_t v = 0; v < m_nv; ++v) out[v].assign(e.begin() + v * nb, e.begin() + (v + 1) * nb);
    return out;
}

std::vector<std::vector<Vector3D<double>>> MeshBZ::get_energy_gradients() const {
    int32_t nb = 0;
    throw_on_error(uepm_bzmesh_bands(mesh(), nullptr, nullptr, &nb), "MeshBZ::get_energy_gradients");
    std::vector<double> g(3 * m_nv * static_cast<size_t>(nb));
    uepm_bzmesh_bands(m_mesh, nullptr, g.data(), &nb);
    std::vector<std::vector<Vector3D<double>>> out(m_nv, std::vector<Vector3D<double>>(nb));
    for (size_t v = 0; v < m_nv; ++v)
        for (int b = 0; b < nb; ++b) {
            const double* q = &g[3 * (v * nb + b)];
            out[v][b] = Vector3D<double>(q[0], q[1], q[2]);
        }
    return out;
}

void MeshBZ::print_band_info() const {
    const int nb = static_cast<int>(get_number_bands_total());
    std::cout << "\nBand info:\nIndex  Type        LocalIdx   GlobalIdx    MinE(eV)   MaxE(eV)\n";
    for (int b = 0; b < nb; ++b) {
        int32_t ic = 0;
        int64_t li = 0;
        double mn = 0, mx = 0;
        if (uepm_bzmesh_band_info(m_mesh, b, &ic, &li, &mn, &mx) != UEPM_OK) break;
        std::cout << b << "  " << (ic ? "Conduction" : "Valence") << "  " << li << "  " << b << "  " << mn << "  " << mx
                  << "\n";
    }
    std::cout << std::endl;
}

}  // namespace uepm::mesh_bz
