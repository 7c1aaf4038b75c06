// eigen_types.h — the two dense result types the reference's Hamiltonian
// hands out (Eigen::VectorXd, Eigen::MatrixXcd, src/EPP/Hamiltonian.h:18-21),
// with the accessors its callers use: rows(), cols(), size(), (i), (i, j),
// data().  Column-major like Eigen.  Eigen itself is not a dependency.
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

namespace uepm::pseudopotential {

class VectorXd {
 public:
    VectorXd() = default;
    explicit VectorXd(std::ptrdiff_t n) : v_(static_cast<size_t>(n)) {}
    std::ptrdiff_t rows() const { return static_cast<std::ptrdiff_t>(v_.size()); }
    std::ptrdiff_t cols() const { return 1; }
    std::ptrdiff_t size() const { return rows(); }
    double operator()(std::ptrdiff_t i) const { return v_[static_cast<size_t>(i)]; }
    double& operator()(std::ptrdiff_t i) { return v_[static_cast<size_t>(i)]; }
    double operator[](std::ptrdiff_t i) const { return v_[static_cast<size_t>(i)]; }
    const double* data() const { return v_.data(); }
    double* data() { return v_.data(); }
    void resize(std::ptrdiff_t n) { v_.assign(static_cast<size_t>(n), 0.0); }

 private:
    std::vector<double> v_;
};

class MatrixXcd {
 public:
    MatrixXcd() = default;
    MatrixXcd(std::ptrdiff_t r, std::ptrdiff_t c) : r_(r), c_(c), v_(static_cast<size_t>(r * c)) {}
    std::ptrdiff_t rows() const { return r_; }
    std::ptrdiff_t cols() const { return c_; }
    std::ptrdiff_t size() const { return r_ * c_; }
    const std::complex<double>& operator()(std::ptrdiff_t i, std::ptrdiff_t j) const {
        return v_[static_cast<size_t>(i + j * r_)];
    }
    std::complex<double>& operator()(std::ptrdiff_t i, std::ptrdiff_t j) { return v_[static_cast<size_t>(i + j * r_)]; }
    const std::complex<double>* data() const { return v_.data(); }
    std::complex<double>* data() { return v_.data(); }
    void resize(std::ptrdiff_t r, std::ptrdiff_t c) {
        r_ = r;
        c_ = c;
        v_.assign(static_cast<size_t>(r * c), {});
    }

 private:
    std::ptrdiff_t r_ = 0, c_ = 0;
    std::vector<std::complex<double>> v_;
};

}  // namespace uepm::pseudopotential
