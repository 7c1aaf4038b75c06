// Hamiltonian.h — drop-in for uepm::pseudopotential::Hamiltonian
// (src/EPP/Hamiltonian.h:10-32).  SetMatrix records k and the terms;
// Diagonalize solves H(k) on the GPU through the C ABI (all eigenvalues,
// ascending, as Eigen's SelfAdjointEigenSolver returns them, and optionally
// all eigenvectors as columns).  Like the reference, the object keeps
// references to the material and the basis (the caller owns both) and is not
// thread-safe; use one object per thread.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "Material.h"
#include "Vector3D.h"
#include "eigen_types.h"

namespace uepm::pseudopotential {

class Hamiltonian {
 public:
    Hamiltonian(const Material& material, const std::vector<Vector3D<int>>& basisVectors);
    Hamiltonian(const Hamiltonian& other);
    Hamiltonian(Hamiltonian&&) noexcept;
    ~Hamiltonian();

    /* The k-independent local matrix lives in the device plan; kept for callers. */
    void SetConstantNonDiagonalMatrix() {}

    void SetMatrix(const Vector3D<double>& k, bool add_non_local_correction = false, bool enable_soc = false);
    /* Throws std::runtime_error("Eigenvalue decomposition failed") like Hamiltonian.cpp:130-133. */
    void Diagonalize(bool keep_eigenvectors = false);

    const VectorXd& get_eigenvalues() const { return m_eigenvalues; }
    const MatrixXcd& get_eigenvectors() const { return m_eigenvectors; }
    const VectorXd& eigenvalues() const { return m_eigenvalues; }

    /* H(k) of the last SetMatrix equals its transpose (Hamiltonian.h:22). */
    bool check_matrix_is_symmetric() const;

    /* Hellmann-Feynman gradient of `level` at the k of the last SetMatrix, in
     * eV m (Hamiltonian.cpp:146-194); needs Diagonalize(true). */
    Vector3D<double> compute_gradient_at_level(const Vector3D<double>& k_point, unsigned int level_index) const;

 protected:
    const Material& m_material;
    const std::vector<Vector3D<int>>& m_basisVectors;

 private:
    struct Plans;
    std::unique_ptr<Plans> m_plans;
    Vector3D<double> m_k;
    bool m_nl = false, m_soc = false, m_has_vectors = false;
    VectorXd m_eigenvalues;
    MatrixXcd m_eigenvectors;
    std::vector<double> m_gradients;
};

}  // namespace uepm::pseudopotential
