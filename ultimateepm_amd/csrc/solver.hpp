// solver.hpp — launch interface between the host C ABI (uepm_abi.cpp) and the
// device stages (solver_f64.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "epm_physics.hpp"

namespace uepm {

// Layout of the blocked reduction for one matrix size: leading columns [0, c0)
// in global scratch, trailing columns in LDS (lds_bytes dynamic LDS per workgroup).
struct TriPlan {
    int n;
    int bs;  // threads per workgroup (256 or 512)
    int c0;
    int lds_bytes;
    int64_t scratch_elems;  // complex elements of global scratch per workgroup
};
TriPlan tri_plan(int n, int elem_bytes, int lds_limit, int bs);
// s_end < n - 1 (a multiple of tridiag_blocked_panel(element bytes)) stops after
// step s_end - 1 and leaves the updated trailing block A[s_end:, s_end:] (lower
// triangle) at its place in Vout, for the register grids to continue.
int tridiag_blocked_panel(int elem_bytes);
hipError_t launch_tridiag_blocked_f64(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                      double2* Vout, double* dd, double* de, double2* dtau, hipStream_t s,
                                      int s_end = 1 << 30);
hipError_t launch_tridiag_blocked_c64(const TriPlan& tp, const double2* dH, int ld, int64_t nk, void* scratch,
                                      float2* Vout, double* dd, double* de, float2* dtau, hipStream_t s,
                                      int s_end = 1 << 30);

hipError_t launch_assemble(const DevParams& P, const int32_t* dG, const double* dk, int64_t nk, double2* dH, int ld,
                           hipStream_t s);
// Failure reporting.  A solve's device status is two words: the number of
// k-points whose eigen-solve failed and the lowest index among them (all ones
// when none); launch_status_reset clears it.  Inside a chunk, kstat (one int
// per k-point of the chunk) carries the tridiagonal stage's flag to the last
// stage, which adds the chunk's failures (index offset k0) to the status.
hipError_t launch_status_reset(unsigned long long* status, hipStream_t s);
// abstol_rel > 0 ends bisection at brackets narrower than abstol_rel ||T||_inf
// (the precision form, whose T carries complex64 rounding); 0 bisects to the
// rounding level of each eigenvalue.
// inv_sweeps: inverse-iteration sweeps for the eigenvectors of T (0: eigenvalues only).
// A non-finite T gives NaN eigenvalues and vectors and is flagged in kstat
// (when not null) or else added to status (when not null).
hipError_t launch_tri_eig(const double* dd, const double* de, int n, int nb, int64_t nk, double* devals, double* dZ,
                          double* dW, int inv_sweeps, double abstol_rel, int* kstat, unsigned long long* status,
                          int64_t k0, hipStream_t s);
// FP64 Rayleigh-Ritz refinement of nbr approximate eigenvectors X (nk x nbr x n,
// overwritten) against the FP64 Hamiltonians H (nk x n x n column-major):
// writes the lowest nb refined eigenvalues / vectors / gradients.  Adds to
// status the k-points kstat flags, whose Jacobi sweeps do not converge or
// whose Ritz values are not finite.
hipError_t launch_refine(const double2* dH, int n, int ld, double2* dX, int nbr, int nb, double* devals,
                         double2* devecs, double* dgrads, const DevParams& P, const double* dk, const int32_t* dG,
                         int64_t nk, const int* kstat, unsigned long long* status, int64_t k0, hipStream_t s);

// Name of the refinement kernel launch_refine runs for nbr Ritz vectors.
const char* refine_kernel_name(int nbr);

// Register-resident complex64 reduction for dim <= tridiag_reg_max_n(): same
// outputs as launch_tridiag_blocked_c64 (reflectors in Vout column-major with ld).
int tridiag_reg_max_n();
// Trailing size the largest grid hands to the next smaller one (0: none) and
// the complex64 hand-over workspace per k-point (bytes) `work` must provide.
int tridiag_reg_split(int n);
size_t tridiag_reg_work_bytes(int n);
hipError_t launch_tridiag_reg_c64(const double2* dH, int n, int ld, int64_t nk, float2* Vout, double* dd, double* de,
                                  float2* dtau, float2* work, hipStream_t s);
// Complex64 register grids after the blocked reduction for n > 288: the
// trailing size they take over (a panel boundary of the blocked kernel; n
// itself for n <= 288) and the launch on a complex64
// block left in place by the blocked kernel (arguments as for the FP64 form
// below; the hand-over workspace is tridiag_reg_work_bytes(m)).
int tridiag_reg_c64_tail(int n);
hipError_t launch_tridiag_reg_c64_block(const float2* H, int64_t hk, int m, int ld, int64_t nk, float2* V, int ldv,
                                        int64_t vk, double* dd, double* de, float2* dtau, int64_t dk, float2* work,
                                        hipStream_t s);
// The same register-resident reduction in FP64 (grids 192 .. 64): trailing
// size it takes over for a matrix of size n (n itself when n <= 192), its
// hand-over workspace per k-point (bytes), and the
// launch on a block of size m whose element (0, 0) sits at H (leading
// dimension ld, per-k stride hk; hermitian_full: both triangles hold the
// Hermitian block, else only its lower triangle); reflectors, d, e, tau at the
// block's offset of the full matrix's outputs (V may alias H: each stage loads
// its block before it writes, and writes only below the sub-diagonal).
int tridiag_reg_c128_tail(int n);
size_t tridiag_reg_c128_work_bytes(int m);
hipError_t launch_tridiag_reg_c128(const double2* H, int64_t hk, int m, int ld, bool hermitian_full, int64_t nk,
                                   double2* V, int ldv, int64_t vk, double* dd, double* de, double2* dtau, int64_t dk,
                                   double2* work, hipStream_t s);

// Eigenvectors of H from those of T: cluster re-orthogonalisation of Z
// ([k][band][row], n x nb per k), then x = Q z with the reflectors of the
// reduction (FP64: over H; precision form: complex64 Vout), gradients when dgrads.
hipError_t launch_cluster_orth(const double* dd, const double* de, const double* devals, double* dZ, int n, int nb,
                               int64_t nk, hipStream_t s);
// bands_per_group: bands per workgroup (16 lanes each), 4, 8 or 16, or 0 for the
// blocked MFMA back-transform (compact WY on FP64 MFMA, sixteen bands per pass),
// which needs apply_q_wy_f64_fits under the per-workgroup LDS limit lds_limit.
// The FP64 form's back-transform is its last stage: it adds the k-points kstat
// flags to status (both may be null).
hipError_t launch_apply_q_f64(const double2* dV, int n, int ld, const double2* dtau, const double* dZ, int nb,
                              double2* dX, double* dgrads, const DevParams& P, const double* dk, const int32_t* dG,
                              int64_t nk, const int* kstat, unsigned long long* status, int64_t k0,
                              int bands_per_group, int lds_limit, hipStream_t s);
bool apply_q_wy_f64_fits(int n, int lds_limit);
// Precision form: bands_per_group 0 selects the blocked MFMA back-transform
// (compact WY, sixteen reflectors per block), which needs apply_q_wy_fits
// under the per-workgroup LDS limit lds_limit (bytes).
hipError_t launch_apply_q_c64(const float2* dV, int n, int ld, const float2* dtau, const double* dZ, int nb,
                              double2* dX, const DevParams& P, const double* dk, const int32_t* dG, int64_t nk,
                              int bands_per_group, int lds_limit, hipStream_t s);
bool apply_q_wy_fits(int n, int nb, int lds_limit);

// Device-to-device copy (16-byte aligned pointers take the vector path).
hipError_t launch_copy_bytes(void* dst, const void* src, uint64_t bytes, int num_cus, hipStream_t s);

// Tetrahedron density of states on a Brillouin-zone mesh (tetra_dos.hip).
// pos: nv x 3 (1/m); tet: nt x 4 vertex indices; det: nt (6 x signed volume).
hipError_t launch_tetra_det(const double* pos, const int32_t* tet, int64_t nt, double* det, hipStream_t s);
// Vertex-major energies E (nv x ld) / gradients G (nv x ld x 3, may be null),
// bands [first, first + nb) -> band-major Eb (nb x nv) / Gb (nb x nv x 3).
hipError_t launch_bands_to_band_major(const double* E, const double* G, int64_t nv, int ld, int first, int nb,
                                      double* Eb, double* Gb, hipStream_t s);
// Slices of the tetrahedron list per energy group (a function of nt only); the
// partial-sum workspace holds tetra_dos_slices(nt, n_points) x n_points
// doubles, the window nt int2.
int tetra_dos_slices(int64_t nt, int n_points);
hipError_t launch_tetra_dos(const double* pos, const int32_t* tet, const double* det, const uint8_t* in_wedge,
                            int64_t nt, const double* Eb, const double* Gb, int n_points, double e_min, double e_max,
                            double reduce_bz_factor, double spin_degeneracy, double wedge_factor, int2* window,
                            double* partial, double* energies, double* dos, hipStream_t s);

// Dielectric-function sums (dielectric.hip): eigenpairs at k (Ek, Xk) and
// k + q (Ekq, Xkq), nk x nb and nk x nb x dim band-contiguous; adds the
// n_energies Lorentzian sums into `sum`; kweight (nk doubles, or null for 1)
// scales the pair weights of each k.  Workspace: weight and delta
// nk x (nb - 4) x 4 doubles each, partial dielectric_slices(...) x n_energies.
int dielectric_slices(int64_t nitems);
hipError_t launch_dielectric_accumulate(const double2* Xk, const double2* Xkq, const double* Ek, const double* Ekq,
                                        const double* kweight, int nb, int dim, int64_t nk, const double* energies,
                                        int n_energies,
                                        double eta, double* weight, double* delta, double* partial, double* sum,
                                        hipStream_t s);

}  // namespace uepm
