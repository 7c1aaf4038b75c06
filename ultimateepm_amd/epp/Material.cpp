// Material.cpp — Materials::load_material_parameters (src/EPP/Material.cpp:33-71,
// NonLocalParameters.hpp:91-116, SpinOrbitParameters.hpp:80-87) and the C-ABI view.
#include "Material.h"

#include <cstdint>
#include <stdexcept>

#include "yaml_reader.h"

namespace uepm::pseudopotential {

namespace {
// CODATA 2018 (src/Common/physical_constants.hpp:18-41): Ry = E_h / 2 in eV
constexpr double kRydToEV = (4.3597447222071e-18 / 1.602176634e-19) / 2.0;
constexpr double kAngstromToM = 1.0e-10;

double req(const yaml::Node* n, const char* key) {
    const yaml::Node* v = n ? n->get(key) : nullptr;
    if (!v) throw std::runtime_error(std::string("Materials::load_material_parameters: missing key ") + key);
    return v->as_double(key);
}

// an unsigned integer value (the reference converts these with as<unsigned int>, which refuses other values)
unsigned req_unsigned(const yaml::Node* n, const char* key) {
    const double x = req(n, key);
    if (!(x >= 0.0 && x <= 4294967295.0 && x == static_cast<double>(static_cast<uint64_t>(x))))
        throw std::runtime_error(std::string("Materials::load_material_parameters: ") + key +
                                 " is not an unsigned integer");
    return static_cast<unsigned>(x);
}
}  // namespace

void Materials::load_material_parameters(const std::string& filename) {
    auto root = yaml::load_file(filename);
    const yaml::Node* list = root->get("materials");
    if (!list || list->kind != yaml::Node::Seq)
        throw std::runtime_error("Materials::load_material_parameters: no 'materials' list in " + filename);
    for (const auto& mp : list->seq) {
        const yaml::Node* m = mp.get();
        const yaml::Node* sym = m->get("symbol");
        if (!sym) throw std::runtime_error("Materials::load_material_parameters: material without symbol");
        const std::string symbol = sym->as_string();
        const double a = req(m, "lattice-constant");
        const yaml::Node* pp = m->get("pseudo-potential-parameters");
        const double V3S = kRydToEV * req(pp, "V3S"), V8S = kRydToEV * req(pp, "V8S"), V11S = kRydToEV * req(pp, "V11S");
        const double V3A = kRydToEV * req(pp, "V3A"), V4A = kRydToEV * req(pp, "V4A"), V11A = kRydToEV * req(pp, "V11A");
        const double V4S = pp->get("V4S") ? kRydToEV * req(pp, "V4S") : 0.0;
        const double V8A = pp->get("V8A") ? kRydToEV * req(pp, "V8A") : 0.0;
        materials[symbol] = Material(symbol, a, V3S, V4S, V8S, V11S, V3A, V4A, V8A, V11A);

        if (const yaml::Node* nl = m->get("non-local-parameters")) {
            NonLocalParameters p;
            p.m_alpha_0_cation = kRydToEV * req(nl, "alpha_0_cation");
            p.m_alpha_0_anion = kRydToEV * req(nl, "alpha_0_anion");
            p.m_beta_0_cation = req(nl, "beta_0_cation");
            p.m_beta_0_anion = req(nl, "beta_0_anion");
            p.m_A2_cation = kRydToEV * req(nl, "A2_cation");
            p.m_A2_anion = kRydToEV * req(nl, "A2_anion");
            p.m_R0_cation = req(nl, "R0_cation") * kAngstromToM;
            p.m_R2_cation = req(nl, "R2_cation") * kAngstromToM;
            p.m_R0_anion = req(nl, "R0_anion") * kAngstromToM;
            p.m_R2_anion = req(nl, "R2_anion") * kAngstromToM;
            const yaml::Node* wt = nl->get("well_type");
            const std::string w = wt ? wt->as_string() : "";
            if (w == "square")
                p.m_well_type = non_local_well_type::square;
            else if (w == "gaussian")
                p.m_well_type = non_local_well_type::gaussian;
            else
                throw std::runtime_error("Unknown non-local well type");
            materials[symbol].set_non_local_parameters(p);
            materials[symbol].set_is_non_local_parameters_populated(true);
        }
        if (const yaml::Node* so = m->get("spin-orbit-parameters")) {
            SpinOrbitParameters p;
            p.m_period_anion = req_unsigned(so, "period_anion");
            p.m_period_cation = req_unsigned(so, "period_cation");
            p.m_radial_extent_anion = req(so, "radial_extent_anion");
            p.m_radial_extent_cation = req(so, "radial_extent_cation");
            p.m_alpha = req(so, "alpha_soc");
            p.m_mu = req(so, "mu_soc");
            materials[symbol].set_spin_orbit_parameters(p);
            materials[symbol].set_is_spin_orbit_parameters_populated(true);
        }
    }
}

uepm_material Material::to_uepm() const {
    uepm_material c{};
    c.lattice_constant_angstrom = m_lattice_constant;
    const Pseudopotential& p = m_pseudopotential;
    const double vs[4] = {p.m_V3S, p.m_V4S, p.m_V8S, p.m_V11S}, va[4] = {p.m_V3A, p.m_V4A, p.m_V8A, p.m_V11A};
    for (int i = 0; i < 4; ++i) {
        c.V_S[i] = vs[i];
        c.V_A[i] = va[i];
    }
    const NonLocalParameters& n = m_non_local_parameters;
    c.nl_populated = m_is_non_local_parameters_populated ? 1 : 0;
    c.nl_well_type = n.m_well_type == non_local_well_type::square ? 0 : n.m_well_type == non_local_well_type::gaussian ? 1 : 2;
    c.alpha_0_cation = n.m_alpha_0_cation;
    c.beta_0_cation = n.m_beta_0_cation;
    c.A2_cation = n.m_A2_cation;
    c.alpha_0_anion = n.m_alpha_0_anion;
    c.beta_0_anion = n.m_beta_0_anion;
    c.A2_anion = n.m_A2_anion;
    c.R0_cation = n.m_R0_cation;
    c.R2_cation = n.m_R2_cation;
    c.R0_anion = n.m_R0_anion;
    c.R2_anion = n.m_R2_anion;
    const SpinOrbitParameters& s = m_spin_orbit_parameters;
    c.soc_populated = m_is_spin_orbit_parameters_populated ? 1 : 0;
    c.period_anion = static_cast<int32_t>(s.m_period_anion);
    c.period_cation = static_cast<int32_t>(s.m_period_cation);
    c.radial_extent_anion = s.m_radial_extent_anion;
    c.radial_extent_cation = s.m_radial_extent_cation;
    c.alpha_soc = s.m_alpha;
    c.mu_soc = s.m_mu;
    return c;
}

}  // namespace uepm::pseudopotential
