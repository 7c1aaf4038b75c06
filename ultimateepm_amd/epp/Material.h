// Material.h — material parameters as the reference's EPP API exposes them
// (src/EPP/Material.h:13-187, NonLocalParameters.hpp:27-131,
// SpinOrbitParameters.hpp:22-101, Pseudopotential.h:8-40), with the same
// names, units after loading (eV, metres, raw SOC values) and getters.
// Parameter files are read by a schema-specific YAML reader (yaml_reader.h)
// in place of yaml-cpp.
#pragma once

#include <cmath>
#include <map>
#include <string>

#include "uepm.h"

namespace uepm::pseudopotential {

enum class non_local_well_type { square, gaussian, unknown };

struct NonLocalParameters {
    double m_alpha_0_cation = 0.0, m_beta_0_cation = 0.0, m_A2_cation = 0.0;
    double m_alpha_0_anion = 0.0, m_beta_0_anion = 0.0, m_A2_anion = 0.0;
    double m_R0_cation = 0.0, m_R2_cation = 0.0, m_R0_anion = 0.0, m_R2_anion = 0.0;
    non_local_well_type m_well_type = non_local_well_type::unknown;
};

struct SpinOrbitParameters {
    unsigned int m_period_anion = 0;
    unsigned int m_period_cation = 0;
    double m_radial_extent_cation = 0;
    double m_radial_extent_anion = 0;
    double m_alpha = 0;
    double m_mu = 0;
};

/* Local form factors in eV for |G|^2 = 3, 4, 8, 11. */
class Pseudopotential {
 public:
    Pseudopotential(double V3S = 0, double V4S = 0, double V8S = 0, double V11S = 0, double V3A = 0, double V4A = 0,
                    double V8A = 0, double V11A = 0)
        : m_V3S(V3S), m_V4S(V4S), m_V8S(V8S), m_V11S(V11S), m_V3A(V3A), m_V4A(V4A), m_V8A(V8A), m_V11A(V11A) {}
    double m_V3S, m_V4S, m_V8S, m_V11S, m_V3A, m_V4A, m_V8A, m_V11A;
};

class Material {
 public:
    Material() = default;
    Material(const std::string& Name, double a, double V3S, double V4S, double V8S, double V11S, double V3A = 0,
             double V4A = 0, double V8A = 0, double V11A = 0)
        : m_name(Name), m_lattice_constant(a), m_pseudopotential(V3S, V4S, V8S, V11S, V3A, V4A, V8A, V11A) {}

    const std::string& get_name() const { return m_name; }
    double get_lattice_constant_meter() const { return 1.0e-10 * m_lattice_constant; }
    double get_fourier_factor() const { return 2.0 * M_PI / (1.0e-10 * m_lattice_constant); }
    double get_atomic_volume_angstrom() const { return (1.0 / 8.0) * std::pow(m_lattice_constant, 3.0); }
    double get_atomic_volume() const { return (1.0 / 8.0) * std::pow(get_lattice_constant_meter(), 3.0); }
    double get_fermi_momentum() const { return std::pow(96 * M_PI * M_PI, 1.0 / 3.0) / get_lattice_constant_meter(); }

    const Pseudopotential& get_pseudopotential() const { return m_pseudopotential; }
    const NonLocalParameters& get_non_local_parameters() const { return m_non_local_parameters; }
    const SpinOrbitParameters& get_spin_orbit_parameters() const { return m_spin_orbit_parameters; }
    bool is_non_local_parameters_populated() const { return m_is_non_local_parameters_populated; }
    bool is_spin_orbit_parameters_populated() const { return m_is_spin_orbit_parameters_populated; }
    void set_is_non_local_parameters_populated(bool v) { m_is_non_local_parameters_populated = v; }
    void set_is_spin_orbit_parameters_populated(bool v) { m_is_spin_orbit_parameters_populated = v; }

    /* Setters used by the loader (the reference fills these from YAML nodes). */
    void set_non_local_parameters(const NonLocalParameters& p) { m_non_local_parameters = p; }
    void set_spin_orbit_parameters(const SpinOrbitParameters& p) { m_spin_orbit_parameters = p; }

    /* The C-ABI view of this material (include/uepm.h). */
    uepm_material to_uepm() const;

 protected:
    std::string m_name;
    double m_lattice_constant = 0.0;
    Pseudopotential m_pseudopotential;
    NonLocalParameters m_non_local_parameters;
    SpinOrbitParameters m_spin_orbit_parameters;
    bool m_is_non_local_parameters_populated = false;
    bool m_is_spin_orbit_parameters_populated = false;
};

class Materials {
 public:
    Materials() = default;
    /* Materials::load_material_parameters (src/EPP/Material.cpp:33-71): form
     * factors and non-local energies Ry -> eV, radii Angstrom -> m, SOC raw.
     * Throws std::runtime_error on unreadable files or an unknown well type. */
    void load_material_parameters(const std::string& filename);
    std::map<std::string, Material> materials;
};

}  // namespace uepm::pseudopotential
