// test_dropin.cpp — exercises the drop-in EPP classes the way the reference's
// callers use them (apps/EmpiricalPseudoPotentialMain.cpp:68-80,
// src/BZ_MESH/bz_mesh.cpp:1255-1258).
//
//   test_dropin cpu <materials.yaml> [<reference-style materials.yaml>]
//       host-only checks: loader, k path, basis, exceptions (no device needed)
//   test_dropin gpu <materials.yaml> <local-cohen.yaml> <outdir>
//       band structures through the GPU; results written for the Python checker
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "BandStructure.h"
#include "Hamiltonian.h"
#include "Material.h"
#include "SymmetryPoints.h"

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
    std::printf("gpu checks: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc >= 3 && std::string(argv[1]) == "cpu") return run_cpu(argc, argv);
    if (argc >= 5 && std::string(argv[1]) == "gpu") return run_gpu(argv);
    std::fprintf(stderr, "usage: test_dropin cpu <materials.yaml> [...] | gpu <materials.yaml> <cohen.yaml> <outdir>\n");
    return 2;
}
