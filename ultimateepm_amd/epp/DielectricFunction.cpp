// DielectricFunction.cpp — the drop-in over uepm_dielectric /
// uepm_dielectric_merge / uepm_kramers_kronig.  Sampling, the k slice,
// result bookkeeping and file formats follow src/EPP/DielectricFunction.cpp.
#include "DielectricFunction.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>

#include "device_context.h"
#include "uepm.h"

namespace uepm::pseudopotential {

bool is_in_irreducible_wedge(const Vector3D<double>& k) {
    // DielectricFunction.cpp:31-33
    return (k.Z >= 0.0) && (k.Y >= k.Z) && (k.X >= k.Y) && (k.X <= 1.0) && (k.X + k.Y + k.Z <= 3.0 / 2.0);
}

bool is_in_first_BZ(const Vector3D<double>& k, bool one_eighth) {
    // DielectricFunction.cpp:35-40
    const bool cond_1 = std::fabs(k.X) <= 1.0 && std::fabs(k.Y) <= 1.0 && std::fabs(k.Z) <= 1.0;
    const bool cond_2 = std::fabs(k.X) + std::fabs(k.Y) + std::fabs(k.Z) <= 3.0 / 2.0;
    const bool cond_eighth = (k.X >= 0.0 && k.Y >= 0.0 && k.Z >= 0.0);
    return cond_1 && cond_2 && (one_eighth ? cond_eighth : true);
}

DielectricFunction::DielectricFunction(const Material& material, const std::vector<Vector3D<int>>& basisVectors,
                                       const int nb_bands)
    : m_basisVectors(basisVectors), m_material(material), m_nb_bands(nb_bands) {}

void DielectricFunction::generate_k_points_random(std::size_t nb_points) {
    // DielectricFunction.cpp:47-57
    std::random_device rd;
    std::mt19937 gen(rd());
    std::uniform_real_distribution<> dis(-1, 1.0);
    while (m_kpoints.size() < nb_points) {
        const double x = dis(gen), y = dis(gen), z = dis(gen);
        Vector3D<double> k(x, y, z);
        if (is_in_first_BZ(k)) m_kpoints.push_back(k);
    }
}

void DielectricFunction::generate_k_points_grid(std::size_t Nx, std::size_t Ny, std::size_t Nz, double shift,
                                                bool irreducible_wedge) {
    // DielectricFunction.cpp:59-74
    m_kpoints.clear();
    const double min = -1.0 - shift, max = 1.0 + shift;
    for (std::size_t i = 0; i + 1 < Nx; ++i)
        for (std::size_t j = 0; j + 1 < Ny; ++j)
            for (std::size_t k = 0; k + 1 < Nz; ++k) {
                Vector3D<double> kv(min + (max - min) * i / static_cast<double>(Nx - 1),
                                    min + (max - min) * j / static_cast<double>(Ny - 1),
                                    min + (max - min) * k / static_cast<double>(Nz - 1));
                if (is_in_first_BZ(kv) && (!irreducible_wedge || is_in_irreducible_wedge(kv))) m_kpoints.push_back(kv);
            }
}

void DielectricFunction::compute_dielectric_function(double eta_smearing, int mpi_rank) {
    if (m_offset_k_index + m_nb_kpoints > m_kpoints.size())
        throw std::out_of_range("DielectricFunction::compute_dielectric_function: k slice beyond the k-points");
    const int N = static_cast<int>(m_basisVectors.size());
    std::vector<int32_t> g(3 * static_cast<size_t>(N));
    for (int i = 0; i < N; ++i) {
        g[3 * i] = m_basisVectors[i].X;
        g[3 * i + 1] = m_basisVectors[i].Y;
        g[3 * i + 2] = m_basisVectors[i].Z;
    }
    const uepm_material mat = m_material.to_uepm();
    uepm_plan* plan = nullptr;
    throw_on_error(uepm_plan_create(device_context(), &mat, g.data(), N,
                                    UEPM_VECTORS | (m_nonlocal_epm ? UEPM_NONLOCAL : 0u), m_nb_bands, &plan),
                   "DielectricFunction::compute_dielectric_function");
    std::vector<double> k(3 * m_nb_kpoints), q(3 * m_qpoints.size());
    for (std::size_t i = 0; i < m_nb_kpoints; ++i) {
        const auto& kv = m_kpoints[m_offset_k_index + i];
        k[3 * i] = kv.X;
        k[3 * i + 1] = kv.Y;
        k[3 * i + 2] = kv.Z;
    }
    for (std::size_t i = 0; i < m_qpoints.size(); ++i) {
        q[3 * i] = m_qpoints[i].X;
        q[3 * i + 1] = m_qpoints[i].Y;
        q[3 * i + 2] = m_qpoints[i].Z;
    }
    std::vector<double> out(m_qpoints.size() * m_energies.size());
    const int st = uepm_dielectric(plan, k.data(), static_cast<int64_t>(m_nb_kpoints), q.data(),
                                   static_cast<int64_t>(m_qpoints.size()), m_energies.data(),
                                   static_cast<int32_t>(m_energies.size()), eta_smearing, out.data());
    uepm_plan_destroy(plan);
    throw_on_error(st, "DielectricFunction::compute_dielectric_function");
    for (std::size_t iq = 0; iq < m_qpoints.size(); ++iq) {
        m_dielectric_function_real.emplace_back(out.begin() + (long)(iq * m_energies.size()),
                                                out.begin() + (long)((iq + 1) * m_energies.size()));
        if (mpi_rank == 0)
            std::cout << "Computed dielectric function for q = " << m_qpoints[iq] << " -> " << iq + 1 << "/"
                      << m_qpoints.size() << std::endl;
    }
}

const std::vector<double> DielectricFunction::get_flat_dielectric_function() const {
    std::vector<double> result;
    for (const auto& row : m_dielectric_function_real) result.insert(result.end(), row.begin(), row.end());
    return result;
}

DielectricFunction DielectricFunction::merge_results(
    DielectricFunction RootDielectricFunction, const std::vector<std::vector<std::vector<double>>>& results,
    std::vector<int> nb_kpoints_per_instance) {
    // DielectricFunction.cpp:167-214: add the instances' rows, then renormalise
    if (results.empty()) throw std::runtime_error("No results to merge");
    if (results.size() != nb_kpoints_per_instance.size())
        throw std::runtime_error("Number of results and number of k-points per instance do not match");
    int64_t total = 0;
    for (int n : nb_kpoints_per_instance) total += n;
    std::vector<std::vector<double>> sum;
    for (std::size_t inst = 0; inst < results.size(); ++inst)
        for (std::size_t iq = 0; iq < results[inst].size(); ++iq) {
            if (inst == 0) {
                sum.push_back(results[inst][iq]);
            } else {
                for (std::size_t e = 0; e < results[inst][iq].size(); ++e) sum[iq][e] += results[inst][iq][e];
            }
        }
    for (std::size_t iq = 0; iq < sum.size(); ++iq) {
        const auto& qv = RootDielectricFunction.m_qpoints.at(iq);
        const double qq[3] = {qv.X, qv.Y, qv.Z};
        throw_on_error(uepm_dielectric_merge(qq, 1, static_cast<int32_t>(sum[iq].size()), total, sum[iq].data()),
                       "DielectricFunction::merge_results");
    }
    DielectricFunction merged = RootDielectricFunction;
    merged.m_dielectric_function_real = sum;
    return merged;
}

void DielectricFunction::apply_kramers_kronig() {
    m_dielectric_function_imag.assign(m_dielectric_function_real.size(), std::vector<double>(m_energies.size()));
    for (std::size_t iq = 0; iq < m_qpoints.size() && iq < m_dielectric_function_real.size(); ++iq)
        throw_on_error(uepm_kramers_kronig(m_dielectric_function_real[iq].data(), 1,
                                           static_cast<int32_t>(m_energies.size()), m_dielectric_function_imag[iq].data()),
                       "DielectricFunction::apply_kramers_kronig");
}

void DielectricFunction::export_kpoints(const std::string& filename) const {
    std::ofstream file(filename);
    file << "X,Y,Z" << std::endl;
    for (const auto& k : m_kpoints) file << k.X << "," << k.Y << "," << k.Z << std::endl;
}

void DielectricFunction::export_dielectric_function_at_q(const std::string& filename, std::size_t idx_q,
                                                         bool name_auto) const {
    std::string outname = filename;
    if (name_auto) {
        // "{prefix}_{idx:05}_{qx:.6f}_{qy:.6f}_{qz:.6f}.csv" (DielectricFunction.cpp:258-263)
        char buf[256];
        std::snprintf(buf, sizeof buf, "_%05zu_%.6f_%.6f_%.6f.csv", idx_q, m_qpoints[idx_q].X, m_qpoints[idx_q].Y,
                      m_qpoints[idx_q].Z);
        outname = m_export_prefix + buf;
    }
    std::ofstream outfile(outname);
    outfile << "Energy (eV),EpsilonReal,EpsilonImaginary" << std::endl;
    for (std::size_t e = 0; e < m_energies.size(); ++e)
        outfile << m_energies[e] << "," << m_dielectric_function_real[idx_q][e] << ","
                << m_dielectric_function_imag.at(idx_q).at(e) << std::endl;
}

void DielectricFunction::export_dielectric_function(const std::string& filename, bool name_auto) const {
    for (std::size_t iq = 0; iq < m_qpoints.size(); ++iq) export_dielectric_function_at_q(filename, iq, name_auto);
}

}  // namespace uepm::pseudopotential
