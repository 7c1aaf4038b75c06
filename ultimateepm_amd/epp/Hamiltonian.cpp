#include "Hamiltonian.h"

#include <cmath>
#include <stdexcept>

#include "device_context.h"

namespace uepm::pseudopotential {

// Plans per (non-local, SOC, vectors) combination, created on first use.
struct Hamiltonian::Plans {
    std::array<uepm_plan*, 8> p{};
    ~Plans() {
        for (auto* q : p)
            if (q) uepm_plan_destroy(q);
    }
};

Hamiltonian::Hamiltonian(const Material& material, const std::vector<Vector3D<int>>& basisVectors)
    : m_material(material), m_basisVectors(basisVectors), m_plans(std::make_unique<Plans>()) {}

Hamiltonian::Hamiltonian(const Hamiltonian& o)
    : m_material(o.m_material),
      m_basisVectors(o.m_basisVectors),
      m_plans(std::make_unique<Plans>()),
      m_k(o.m_k),
      m_nl(o.m_nl),
      m_soc(o.m_soc),
      m_has_vectors(o.m_has_vectors),
      m_eigenvalues(o.m_eigenvalues),
      m_eigenvectors(o.m_eigenvectors),
      m_gradients(o.m_gradients) {}

Hamiltonian::Hamiltonian(Hamiltonian&&) noexcept = default;
Hamiltonian::~Hamiltonian() = default;

void Hamiltonian::SetMatrix(const Vector3D<double>& k, bool add_non_local_correction, bool enable_soc) {
    m_k = k;
    m_nl = add_non_local_correction;
    m_soc = enable_soc;
}

namespace {
std::vector<int32_t> flat_basis(const std::vector<Vector3D<int>>& G) {
    std::vector<int32_t> g(3 * G.size());
    for (size_t i = 0; i < G.size(); ++i) {
        g[3 * i] = G[i].X;
        g[3 * i + 1] = G[i].Y;
        g[3 * i + 2] = G[i].Z;
    }
    return g;
}
}  // namespace

void Hamiltonian::Diagonalize(bool keep_eigenvectors) {
    // a non-finite k-point has no eigen-decomposition: the reference's Eigen
    // solver reports NoConvergence there (Hamiltonian.cpp:130-133)
    if (!std::isfinite(m_k.X) || !std::isfinite(m_k.Y) || !std::isfinite(m_k.Z))
        throw std::runtime_error("Eigenvalue decomposition failed");
    const int N = static_cast<int>(m_basisVectors.size());
    const int dim = m_soc ? 2 * N : N;
    const int slot = (m_nl ? 1 : 0) | (m_soc ? 2 : 0) | (keep_eigenvectors ? 4 : 0);
    uepm_plan*& plan = m_plans->p[slot];
    if (!plan) {
        const uepm_material mat = m_material.to_uepm();
        const std::vector<int32_t> g = flat_basis(m_basisVectors);
        unsigned flags = (m_nl ? UEPM_NONLOCAL : 0u) | (m_soc ? UEPM_SOC : 0u) |
                         (keep_eigenvectors ? (UEPM_VECTORS | UEPM_GRADIENTS) : 0u);
        throw_on_error(uepm_plan_create(device_context(), &mat, g.data(), N, flags, dim, &plan),
                       "Hamiltonian::Diagonalize");
    }
    const double k[3] = {m_k.X, m_k.Y, m_k.Z};
    m_eigenvalues.resize(dim);
    std::vector<double> vec, grad;
    if (keep_eigenvectors) {
        vec.resize(2 * static_cast<size_t>(dim) * dim);
        grad.resize(3 * static_cast<size_t>(dim));
    }
    const int st = uepm_solve(plan, k, 1, m_eigenvalues.data(), keep_eigenvectors ? vec.data() : nullptr,
                              keep_eigenvectors ? grad.data() : nullptr);
    if (st == UEPM_ERR_NOT_CONVERGED) throw std::runtime_error("Eigenvalue decomposition failed");
    throw_on_error(st, "Hamiltonian::Diagonalize");
    m_has_vectors = keep_eigenvectors;
    if (keep_eigenvectors) {
        // uepm layout [band][row] -> column-major (row, band)
        m_eigenvectors.resize(dim, dim);
        for (int b = 0; b < dim; ++b)
            for (int r = 0; r < dim; ++r)
                m_eigenvectors(r, b) = {vec[2 * ((size_t)b * dim + r)], vec[2 * ((size_t)b * dim + r) + 1]};
        m_gradients = std::move(grad);
    } else {
        m_eigenvectors.resize(0, 0);
        m_gradients.clear();
    }
}

bool Hamiltonian::check_matrix_is_symmetric() const {
    const int N = static_cast<int>(m_basisVectors.size());
    const int dim = m_soc ? 2 * N : N;
    uepm_ctx* ctx = device_context();
    const uepm_material mat = m_material.to_uepm();
    const std::vector<int32_t> g = flat_basis(m_basisVectors);
    uepm_plan* plan = nullptr;
    throw_on_error(uepm_plan_create(ctx, &mat, g.data(), N, (m_nl ? UEPM_NONLOCAL : 0u) | (m_soc ? UEPM_SOC : 0u), 1,
                                    &plan),
                   "Hamiltonian::check_matrix_is_symmetric");
    const size_t bytes = 16 * (size_t)dim * dim;
    void *dk = nullptr, *dH = nullptr;
    std::vector<double> H(2 * (size_t)dim * dim);
    const double k[3] = {m_k.X, m_k.Y, m_k.Z};
    int st = uepm_device_alloc(ctx, 24, &dk);
    if (st == UEPM_OK) st = uepm_device_alloc(ctx, bytes, &dH);
    if (st == UEPM_OK) st = uepm_memcpy_htod(ctx, dk, k, 24);
    if (st == UEPM_OK) st = uepm_assemble_device(plan, static_cast<double*>(dk), 1, static_cast<double*>(dH), nullptr);
    if (st == UEPM_OK) st = uepm_stream_synchronize(ctx, nullptr);
    if (st == UEPM_OK) st = uepm_memcpy_dtoh(ctx, H.data(), dH, bytes);
    if (dH) uepm_device_free(ctx, dH);
    if (dk) uepm_device_free(ctx, dk);
    uepm_plan_destroy(plan);
    throw_on_error(st, "Hamiltonian::check_matrix_is_symmetric");
    // Eigen's isApprox: ||A - A^T|| <= 1e-12 * min(||A||, ||A^T||) (Frobenius)
    double diff = 0.0, norm = 0.0;
    for (int j = 0; j < dim; ++j)
        for (int i = 0; i < dim; ++i) {
            const size_t a = 2 * ((size_t)j * dim + i), b = 2 * ((size_t)i * dim + j);
            const double dr = H[a] - H[b], di = H[a + 1] - H[b + 1];
            diff += dr * dr + di * di;
            norm += H[a] * H[a] + H[a + 1] * H[a + 1];
        }
    return diff <= 1e-24 * norm;
}

Vector3D<double> Hamiltonian::compute_gradient_at_level(const Vector3D<double>& k_point, unsigned int level_index) const {
    if (!m_has_vectors) throw std::runtime_error("compute_gradient_at_level: call Diagonalize(true) first");
    if (!(k_point == m_k))
        throw std::invalid_argument("compute_gradient_at_level: k differs from the k of the last SetMatrix");
    if (level_index >= static_cast<unsigned>(m_eigenvalues.size()))
        throw std::out_of_range("compute_gradient_at_level: level index out of range");
    const double* g = m_gradients.data() + 3 * (size_t)level_index;
    return Vector3D<double>(g[0], g[1], g[2]);
}

}  // namespace uepm::pseudopotential
