// test_dropin.cpp — exercises the drop-in EPP classes the way the reference's
// callers use them (apps/EmpiricalPseudoPotentialMain.cpp:68-80,
// src/BZ_MESH/bz_mesh.cpp:1255-1258).
//
//   test_dropin cpu <materials.yaml> [<reference-style materials.yaml>]
//       host-only checks: loader, k path, basis, exceptions (no device needed)
//   test_dropin gpu <materials.yaml> <local-cohen.yaml> <outdir>
//       band structures through the GPU; results written for the Python checker
//   test_dropin mesh <local-cohen.yaml> <bz_mesh.msh> <outdir>
//       apps/BandsOnBZ.cpp:85-140 then apps/DOS_MeshBZ.cpp:60-150 through MeshBZ
//   test_dropin states <cohen.yaml> <bz_mesh.msh> <outdir>
//       apps/fullstates.cpp:42-104 through BZ_States
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "BandStructure.h"
#include "DielectricFunction.hpp"
#include "Hamiltonian.h"
#include "Material.h"
#include "SymmetryPoints.h"
#include "bz_mesh.hpp"
#include "bz_states.hpp"
#include "csv_utils.hpp"

#include <unistd.h>

using namespace uepm::pseudopotential;

static int g_fail = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                             \
        }                                                                         \
    } while (0)

template <class E, class F>
static bool throws(F f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static const double kRyd = (4.3597447222071e-18 / 1.602176634e-19) / 2.0;

static void check_materials(const std::string& file) {
    Materials ms;
    ms.load_material_parameters(file);
    CHECK(ms.materials.count("Si") && ms.materials.count("Ge") && ms.materials.count("GaAs") && ms.materials.count("InP"));
    const Material& si = ms.materials.at("Si");
    CHECK(si.get_name() == "Si");
    CHECK(std::fabs(si.get_lattice_constant_meter() - 5.43e-10) < 1e-22);
    CHECK(std::fabs(si.get_pseudopotential().m_V3S - (-0.257 * kRyd)) < 1e-14);
    CHECK(si.is_non_local_parameters_populated());
    CHECK(std::fabs(si.get_non_local_parameters().m_R0_anion - 1.06e-10) < 1e-22);
    CHECK(si.get_non_local_parameters().m_well_type == non_local_well_type::square);
    const Material& ge = ms.materials.at("Ge");
    CHECK(ge.get_non_local_parameters().m_well_type == non_local_well_type::gaussian);
    CHECK(ge.is_spin_orbit_parameters_populated());
    CHECK(ge.get_spin_orbit_parameters().m_period_anion == 4);
    CHECK(std::fabs(ge.get_spin_orbit_parameters().m_mu - 0.00069) < 1e-15);
    const uepm_material c = ge.to_uepm();
    CHECK(c.nl_well_type == 1 && c.soc_populated == 1 && c.period_cation == 4);
}

static int run_cpu(int argc, char** argv) {
    check_materials(argv[2]);
    if (argc > 3) check_materials(argv[3]);
    CHECK(throws<std::runtime_error>([] { Materials m; m.load_material_parameters("/nonexistent/materials.yaml"); }));

    SymmetryPoints sp;
    std::vector<unsigned> pos;
    const std::vector<std::string> path{"L", "G", "X", "W", "K", "U", "L", "W", "X", "K", "G"};
    auto pts = sp.GeneratePoints(path, 2000, pos);
    CHECK(pts.size() == 2000 && pos.size() == path.size());
    for (size_t i = 0; i < path.size(); ++i) CHECK(pts[pos[i]] == sp.symmetryPoints[path[i]].position);
    CHECK(throws<std::invalid_argument>([&] {
        std::vector<unsigned> p2;
        sp.GeneratePoints({"L", "Q"}, 20, p2);
    }));

    Materials ms;
    ms.load_material_parameters(argv[2]);
    BandStructure bs;
    bs.Initialize(ms.materials.at("Si"), 8, path, 100, 12, true, false);
    CHECK(bs.get_basis_vectors().size() == 181);
    CHECK(bs.GetPointsNumber() == 100);
    CHECK(bs.get_path_as_string() == "LGXWKULWXKG");
    BandStructure b2;
    b2.Initialize(ms.materials.at("Ge"), 8, std::vector<Vector3D<double>>{Vector3D<double>(0.1, 0.2, 0.3)}, 10, true, true);
    CHECK(b2.get_number_of_bands() == 16 && b2.get_basis_vectors().size() == 137);
    CHECK(throws<std::runtime_error>([&] { BandStructure b3; b3.Initialize(ms.materials.at("Si"), 8, path, 100, 1, false, false); }));
    // a non-finite k-point: "Eigenvalue decomposition failed" (Hamiltonian.cpp:130-133), before any device work
    {
        const double nan = std::nan("");
        const std::vector<Vector3D<int>> G = bs.get_basis_vectors();
        Hamiltonian h(ms.materials.at("Si"), G);
        h.SetMatrix(Vector3D<double>(0.1, nan, 0.2), true, false);
        bool msg_ok = false;
        try {
            h.Diagonalize(false);
        } catch (const std::runtime_error& e) {
            msg_ok = std::string(e.what()) == "Eigenvalue decomposition failed";
        }
        CHECK(msg_ok);
        BandStructure b4;
        b4.Initialize(ms.materials.at("Si"), 8,
                      std::vector<Vector3D<double>>{Vector3D<double>(0.0, 0.0, 0.0),
                                                    Vector3D<double>(std::numeric_limits<double>::infinity(), 0, 0)},
                      10, false, false);
        CHECK(throws<std::runtime_error>([&] { b4.Compute(); }));
    }

    Vector3D<double> a(1, 2, 3), b(-2, 0.5, 4);
    CHECK(a * b == -2 + 1 + 12);
    const Vector3D<double> cr = a % b;
    CHECK(cr.X == 2 * 4 - 3 * 0.5 && cr.Y == 3 * -2 - 1 * 4 && cr.Z == 1 * 0.5 - 2 * -2);
    CHECK(std::fabs(a.Length() - std::sqrt(14.0)) < 1e-15);
    std::printf("cpu checks: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}

static void write_energies(const std::string& fn, const BandStructure& bs) {
    std::ofstream f(fn);
    f << std::setprecision(17);
    for (const auto& row : bs.get_band_energies()) {
        for (size_t i = 0; i < row.size(); ++i) f << (i ? " " : "") << row[i];
        f << "\n";
    }
}

static int run_gpu(char** argv) {
    Materials ms, cohen;
    ms.load_material_parameters(argv[2]);
    cohen.load_material_parameters(argv[3]);
    const std::string out = argv[4];
    const std::vector<std::string> path{"L", "G", "X", "W", "K", "U", "L", "W", "X", "K", "G"};

    // 1. the band-structure app's call sequence: path, local, Compute_parallel, AdjustValues, export
    {
        BandStructure bs;
        bs.Initialize(cohen.materials.at("Si"), 16, path, 200, 10, false, false);
        bs.Compute_parallel(false, 8);
        write_energies(out + "/path_raw.txt", bs);
        const double gap = bs.AdjustValues();
        CHECK(gap > 0.5 && gap < 1.5);
        bs.export_result_in_file(out + "/" + bs.path_band_filename() + ".txt");
        std::printf("band gap %.6f eV, file %s\n", gap, bs.path_band_filename().c_str());
    }
    // 2. BandsOnBZ-style: explicit k list, non-local + SOC, gradients
    {
        std::vector<Vector3D<double>> ks{Vector3D<double>(0.1, 0.2, 0.3), Vector3D<double>(-0.4, 0.05, 0.6),
                                         Vector3D<double>(0.5, 0.5, 0.5), Vector3D<double>(0, 0, 0)};
        BandStructure bs;
        bs.Initialize(ms.materials.at("Ge"), 8, ks, 8, true, true);
        bs.set_kpoints(ks);
        bs.Compute_parallel(true, 4);
        write_energies(out + "/ge_soc.txt", bs);
        std::ofstream f(out + "/ge_soc_grad.txt");
        f << std::setprecision(17);
        for (const auto& row : bs.get_band_energy_gradients())
            for (const auto& g : row) f << g.X << " " << g.Y << " " << g.Z << "\n";
    }
    // 3. per-k Hamiltonian (DielectricFunction / BZ_States style): all eigenpairs
    {
        BandStructure bs;
        bs.Initialize(ms.materials.at("GaAs"), 4, std::vector<Vector3D<double>>{}, 6, true, false);
        Hamiltonian h(ms.materials.at("GaAs"), bs.get_basis_vectors());
        const Vector3D<double> k(0.13, -0.27, 0.41);
        h.SetMatrix(k, true, false);
        CHECK(!h.check_matrix_is_symmetric());  // Hermitian, not real-symmetric
        h.Diagonalize(true);
        const auto& ev = h.eigenvalues();
        const auto& V = h.get_eigenvectors();
        CHECK(ev.size() == 59 && V.rows() == 59 && V.cols() == 59);
        std::ofstream f(out + "/gaas_all.txt");
        f << std::setprecision(17);
        for (long i = 0; i < ev.size(); ++i) f << ev(i) << (i + 1 < ev.size() ? " " : "\n");
        for (long j = 0; j < V.cols(); ++j)
            for (long i = 0; i < V.rows(); ++i) f << V(i, j).real() << " " << V(i, j).imag() << "\n";
        const auto g = h.compute_gradient_at_level(k, 3);
        f << g.X << " " << g.Y << " " << g.Z << "\n";
        CHECK(throws<std::invalid_argument>([&] { h.compute_gradient_at_level(Vector3D<double>(0, 0, 0), 3); }));
        h.Diagonalize(false);
        CHECK(throws<std::runtime_error>([&] { h.compute_gradient_at_level(k, 3); }));
    }
    // 4. the reference's error for an unsupported spin-orbit period (InP, period 5)
    {
        BandStructure bs;
        bs.Initialize(ms.materials.at("InP"), 4, std::vector<Vector3D<double>>{Vector3D<double>(0.1, 0, 0)}, 6, true, true);
        CHECK(throws<std::runtime_error>([&] { bs.Compute_parallel(false, 1); }));
    }
    // 5. the precision form through the same class
    {
        std::vector<Vector3D<double>> ks;
        for (int i = 0; i < 64; ++i) ks.emplace_back(0.01 * i, 0.3 - 0.004 * i, 0.1 + 0.007 * i);
        BandStructure a, b;
        a.Initialize(ms.materials.at("Si"), 12, ks, 12, true, false);
        b.Initialize(ms.materials.at("Si"), 12, ks, 12, true, false);
        b.set_mixed_precision(true);
        a.Compute_parallel();
        b.Compute_parallel();
        double worst = 0;
        for (size_t i = 0; i < ks.size(); ++i)
            for (int j = 0; j < 12; ++j)
                worst = std::max(worst, std::fabs(a.get_band_energies()[i][j] - b.get_band_energies()[i][j]));
        std::printf("mixed vs fp64 max |dE| = %.3e eV\n", worst);
        CHECK(worst <= 1e-4);
    }
    // 6. DielectricFunction as apps/epsilon.cpp:195-367 drives it (one rank)
    {
        const Material& si = ms.materials.at("Si");
        BandStructure bs;
        bs.Initialize(si, 16, std::vector<Vector3D<double>>{}, 8, true, false);
        DielectricFunction df(si, bs.get_basis_vectors(), 16);
        df.generate_k_points_grid(9, 9, 9, 0.0, false);
        const std::size_t nk = df.get_kpoints().size();
        df.set_export_prefix(out + "/Si_");
        df.set_qpoints({Vector3D<double>(0, 0, 0), Vector3D<double>(0.05, 0.02, 0.0)});
        std::vector<double> en;
        for (int i = 0; i < 101; ++i) en.push_back(0.2 * i);
        df.set_energies(en);
        df.set_offset_k_index(0);
        df.set_nb_kpoints(nk);
        df.set_non_local_epm(true);
        df.compute_dielectric_function(0.05, 0);
        CHECK(df.get_dielectric_function().size() == 2);
        {
            std::ofstream f(out + "/dielectric_k.txt");
            f << std::setprecision(17);
            for (const auto& k : df.get_kpoints()) f << k.X << " " << k.Y << " " << k.Z << "\n";
        }
        std::ofstream f(out + "/dielectric_raw.txt");
        f << std::setprecision(17);
        for (const auto& row : df.get_dielectric_function())
            for (size_t e = 0; e < row.size(); ++e) f << row[e] << (e + 1 < row.size() ? " " : "\n");
        auto merged = DielectricFunction::merge_results(df, {df.get_dielectric_function()}, {static_cast<int>(nk)});
        merged.apply_kramers_kronig();
        merged.export_dielectric_function("", true);
        CHECK(std::ifstream(out + "/Si__00001_0.050000_0.020000_0.000000.csv").good());
        CHECK(throws<std::runtime_error>([&] { DielectricFunction::merge_results(df, {}, {}); }));
    }
    std::printf("gpu checks: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}

static void write_matrix(const std::string& file, const std::vector<std::vector<double>>& rows) {
    std::ofstream f(file);
    f << std::setprecision(17);
    for (const auto& r : rows) {
        for (size_t i = 0; i < r.size(); ++i) f << r[i] << (i + 1 < r.size() ? " " : "\n");
    }
}

static int run_mesh(char** argv) {
    Materials cohen;
    cohen.load_material_parameters(argv[2]);
    const Material& si = cohen.materials.at("Si");
    const std::string msh = argv[3];
    if (chdir(argv[4]) != 0) return 2;  // the k-star file is looked up in the working directory
    using uepm::mesh_bz::MeshBZ;
    // the k-star file must exist, as in the reference; the mesh is kept to write it
    {
        MeshBZ m(si);
        CHECK(throws<std::runtime_error>([&] { m.read_mesh_geometry_from_msh_file(msh); }));
        m.write_kstar_ibz_to_bz("bz_mesh_kstar_ibz_to_bz.txt");
    }
    const int nrLevels = 4 + 12 + 4;
    // BandsOnBZ, full zone
    {
        MeshBZ mesh(si);
        mesh.set_number_threads_mesh_ops(8);
        mesh.read_mesh_geometry_from_msh_file(msh);
        CHECK(mesh.get_number_vertices() == 14705 && mesh.get_number_elements() == 73728);
        CHECK(mesh.get_number_vertices_in_irreducible_wedge() == 405);
        CHECK(std::fabs(mesh.get_reduce_bz_factor() - 1.0) < 1e-9);
        BandStructure bs;
        bs.Initialize(si, nrLevels, std::vector<Vector3D<double>>{}, 10, false, false);
        mesh.compute_band_structure_over_mesh(bs, false);
        CHECK(mesh.get_number_valence_bands() == 4 && mesh.get_number_conduction_bands() == 16);
        write_matrix("mesh_E_full.txt", mesh.get_band_energies());
        mesh.export_selected_bands_to_gmsh("bz_mesh_bands.msh", 4, 12, true, msh, true);
        mesh.export_energies_and_gradients_to_vtk("bz_mesh_bands.vtk");
        CHECK(throws<std::out_of_range>([&] { mesh.get_global_band_index(16, uepm::mesh_bz::MeshParticleType::conduction); }));
    }
    // BandsOnBZ -w: the wedge, fanned out
    {
        MeshBZ mesh(si);
        mesh.read_mesh_geometry_from_msh_file(msh);
        BandStructure bs;
        bs.Initialize(si, nrLevels, std::vector<Vector3D<double>>{}, 10, false, false);
        mesh.compute_band_structure_over_mesh(bs);  // use_iwedge defaults to true
        CHECK(bs.get_band_energies().size() == 405);
        write_matrix("mesh_E_iw.txt", mesh.get_band_energies());
    }
    // DOS_MeshBZ: read the bands back, shift, scissor, DOS per band, CSV
    {
        MeshBZ mesh(si);
        mesh.read_mesh_geometry_from_msh_file(msh);
        mesh.read_mesh_bands_from_msh_file("bz_mesh_bands.msh", -1, -1, true, false);
        mesh.apply_scissor(1.12);
        const int nb = static_cast<int>(mesh.get_number_bands_total());
        CHECK(nb == 16);
        std::vector<std::vector<double>> cols;
        std::vector<std::string> header;
        for (int b = 0; b < nb; ++b) {
            auto ed = mesh.compute_dos_band_at_band_auto(b, 200, false, false);
            cols.push_back(ed[0]);
            cols.push_back(ed[1]);
            header.push_back("energy_band_" + std::to_string(b));
            header.push_back("dos_band_" + std::to_string(b));
        }
        uepm::utils::export_multiple_vector_to_csv("DOS_bz_mesh_bands.csv", header, cols);
        const double d = mesh.compute_dos_at_energy_and_band(cols[2][100], 1);
        CHECK(d == cols[3][100]);
        CHECK(mesh.compute_dos_at_energy_and_band(cols[2][100], 1, true) == 0.0);
        CHECK(throws<std::out_of_range>([&] { mesh.compute_dos_band_at_band_auto(nb, 10); }));
    }
    std::printf("mesh checks: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}

// apps/fullstates.cpp:42-104: eigenstates at the vertices and at vertex + q,
// the dielectric function over the tetrahedra, its CSV.
static int run_states(char** argv) {
    Materials materials;
    materials.load_material_parameters(argv[2]);
    const Material current_material = materials.materials.at("Si");
    const std::string msh = argv[3];
    if (chdir(argv[4]) != 0) return 2;
    {
        uepm::mesh_bz::MeshBZ m(current_material);
        CHECK(throws<std::runtime_error>([&] { m.read_mesh_geometry_from_msh_file(msh); }));
        m.write_kstar_ibz_to_bz("bz_mesh_kstar_ibz_to_bz.txt");
    }
    const int nb_bands_to_use = 12;
    uepm::mesh_bz::BZ_States my_bz_mesh(current_material);
    my_bz_mesh.set_nb_bands(nb_bands_to_use);
    BandStructure band_structure{};
    const int nb_nearest_neighbors = 10;
    band_structure.Initialize(current_material, nb_bands_to_use, {}, nb_nearest_neighbors, false, false);
    my_bz_mesh.set_basis_vectors(band_structure.get_basis_vectors());
    my_bz_mesh.read_mesh_geometry_from_msh_file(msh);
    std::cout << "Mesh volume: " << my_bz_mesh.compute_mesh_volume() << std::endl;
    my_bz_mesh.compute_eigenstates(8);
    Vector3D<double> q_shift = Vector3D<double>{3e-14, 0.0, 0.0};
    my_bz_mesh.compute_shifted_eigenstates(q_shift, 8);
    std::vector<double> list_energy;
    const double min_energy = 0.0, max_energy = 20.0, energy_step = 0.01;
    for (double energy = min_energy; energy <= max_energy + energy_step; energy += energy_step)
        list_energy.push_back(energy);
    my_bz_mesh.compute_dielectric_function(list_energy, 0.05, 8);
    my_bz_mesh.export_dielectric_function("./TEST_DIELECTRIC_FUNCTION_");
    CHECK(my_bz_mesh.get_dielectric_function_real().size() == list_energy.size());
    const auto& X = my_bz_mesh.get_eigen_states();
    CHECK(X.size() == 14705 && X[0].rows() == 137 && X[0].cols() == nb_bands_to_use);
    const auto& E = my_bz_mesh.get_eigenvalues();
    const auto& Eq = my_bz_mesh.get_eigenvalues(true);
    CHECK(E.size() == 14705 && Eq.size() == 14705 && E[0].size() == nb_bands_to_use);
    {
        std::ofstream f("states_E.txt");
        f << std::setprecision(17);
        for (const auto& v : E) {
            for (int j = 0; j < v.size(); ++j) f << (j ? " " : "") << v(j);
            f << "\n";
        }
        double nrm = 0.0;
        for (int i = 0; i < X[7].rows(); ++i) nrm += std::norm(X[7](i, 3));
        CHECK(std::fabs(nrm - 1.0) < 1e-12);
    }
    CHECK(my_bz_mesh.compute_fermi_level(1e21, 300.0) == 0.0);
    CHECK(my_bz_mesh.get_dielectric_function(Vector3D<double>{0.1, 0.0, 0.0}, 1.0) == 1.0);
    {
        uepm::mesh_bz::BZ_States no_basis(current_material);
        no_basis.set_nb_bands(nb_bands_to_use);
        no_basis.read_mesh_geometry_from_msh_file(msh);
        CHECK(throws<std::invalid_argument>([&] { no_basis.compute_eigenstates(1); }));
        CHECK(throws<std::invalid_argument>([&] { no_basis.compute_dielectric_function(list_energy, 0.05, 1); }));
    }
    std::printf("states checks: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc >= 5 && std::string(argv[1]) == "mesh") return run_mesh(argv);
    if (argc >= 5 && std::string(argv[1]) == "states") return run_states(argv);
    if (argc >= 3 && std::string(argv[1]) == "cpu") return run_cpu(argc, argv);
    if (argc >= 5 && std::string(argv[1]) == "gpu") return run_gpu(argv);
    std::fprintf(stderr,
                 "usage: test_dropin cpu <materials.yaml> [...] | gpu <materials.yaml> <cohen.yaml> <outdir>"
                 " | mesh <cohen.yaml> <bz_mesh.msh> <outdir> | states <cohen.yaml> <bz_mesh.msh> <outdir>\n");
    return 2;
}
