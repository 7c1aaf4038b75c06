/*
 * uepm.h — C ABI of the MI355X plane-wave EPM band solver.
 *
 * The reference computes bands through the C++ classes
 *   uepm::pseudopotential::Hamiltonian   (src/EPP/Hamiltonian.h:10-32)
 *   uepm::pseudopotential::BandStructure (src/EPP/BandStructure.h:12-106)
 * with one Hamiltonian object per OpenMP thread and a loop over k
 * (src/EPP/BandStructure.cpp:151-201).  This header is the boundary that
 * replaces that loop: plain C types, caller-owned buffers, int status codes,
 * no exceptions and no Eigen / yaml-cpp / torch types.
 *
 * Units follow the reference: k and G in reduced units of 2*pi/a, energies in
 * eV, lengths in metres inside the physics (lattice constant in Angstrom as in
 * Material::m_lattice_constant, src/EPP/Material.h:26).
 *
 * Threading: a context may be shared between threads; a plan may not (the
 * reference's Hamiltonian is likewise one-per-thread,
 * src/EPP/BandStructure.cpp:168-171).
 */
#ifndef UEPM_H
#define UEPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
    UEPM_OK                = 0,
    UEPM_ERR_INVALID_ARG   = 1, /* std::invalid_argument in the reference  */
    UEPM_ERR_RUNTIME       = 2, /* std::runtime_error in the reference     */
    UEPM_ERR_OUT_OF_RANGE  = 3, /* std::out_of_range in the reference      */
    UEPM_ERR_DEVICE        = 4, /* a HIP call failed                        */
    UEPM_ERR_NOT_CONVERGED = 5, /* eigen-solve did not converge; reference:
                                   "Eigenvalue decomposition failed",
                                   src/EPP/Hamiltonian.cpp:130-133          */
    UEPM_ERR_NO_MEMORY     = 6
};

/* ---- plan flags ---------------------------------------------------------- */
#define UEPM_NONLOCAL   0x1u  /* SetMatrix(k, add_non_local_correction=true) */
#define UEPM_SOC        0x2u  /* SetMatrix(k, ..., enable_soc=true): 2N x 2N */
#define UEPM_VECTORS    0x4u  /* Diagonalize(keep_eigenvectors=true)         */
#define UEPM_GRADIENTS  0x8u  /* compute_gradient_at_level for each band;
                                 implies UEPM_VECTORS                          */
#define UEPM_MIXED      0x10u /* precision form: complex64 reduction + FP64
                                 refinement (eigenvalues within 0.1 meV)     */

/* Material parameters in the units the reference holds after loading
 * (Materials::load_material_parameters, src/EPP/Material.cpp:33-71;
 *  NonLocalParameters::populate_non_local_parameters,
 *  src/EPP/NonLocalParameters.hpp:91-116;
 *  SpinOrbitParameters::populate_from_yaml, src/EPP/SpinOrbitParameters.hpp:80-87). */
typedef struct uepm_material {
    double lattice_constant_angstrom;  /* Material::m_lattice_constant        */
    double V_S[4];                     /* eV, |G|^2 = 3, 4, 8, 11             */
    double V_A[4];                     /* eV, |G|^2 = 3, 4, 8, 11             */
    /* non-local (Pötz–Vogl / Chelikowsky–Cohen) */
    int32_t nl_populated;              /* 0: all-zero parameters              */
    int32_t nl_well_type;              /* 0 square, 1 gaussian, 2 unknown     */
    double alpha_0_cation, beta_0_cation, A2_cation; /* eV, 1, eV           */
    double alpha_0_anion, beta_0_anion, A2_anion;    /* eV, 1, eV           */
    double R0_cation, R2_cation, R0_anion, R2_anion; /* metres              */
    /* spin-orbit (Pötz–Vogl), raw YAML values */
    int32_t soc_populated;
    int32_t period_anion, period_cation;
    double radial_extent_anion, radial_extent_cation, alpha_soc, mu_soc;
} uepm_material;

typedef struct uepm_ctx uepm_ctx;
typedef struct uepm_plan uepm_plan;

/* Last error message of the calling thread ("" when none). */
const char* uepm_last_error(void);

/* Context on one HIP device.  Replaces nothing in the reference (it has no
 * device); holds the device ordinal and its properties. */
int uepm_ctx_create(int device, uepm_ctx** out);
int uepm_ctx_destroy(uepm_ctx* ctx);

/* Basis generation: BandStructure::GenerateBasisVectors
 * (src/EPP/BandStructure.cpp:26-57).  Writes up to `capacity` integer
 * triples (x,y,z) of G = i*b1 + j*b2 + k*b3, |G|^2 <= G2[n-1], in the
 * reference's i,j,k loop order.  *count receives the basis size N.
 * n outside [2, 116] -> UEPM_ERR_RUNTIME (the reference returns false and
 * Initialize throws runtime_error). Host-only. */
int uepm_basis_vectors(int nearest_neighbors, int32_t* G_out, int capacity, int* count);

/* k path: SymmetryPoints::GeneratePoints (src/EPP/SymmetryPoints.cpp:47-147).
 * labels: one letter per node from {L,G,X,W,K,U} (e.g. "LGXWKULWXKG").
 * Writes nrPoints x 3 doubles to k_out and strlen(labels) node indices to
 * sym_positions; *count receives the number of points.  Fewer than two
 * labels, nrPoints below the label count, or an unknown label ->
 * UEPM_ERR_INVALID_ARG (std::invalid_argument in the reference). Host-only. */
int uepm_kpath(const char* labels, unsigned nrPoints, double* k_out, unsigned* sym_positions, int* count);

/* Seeded uniform sample of the first Brillouin zone: rejection sampling in
 * [-1,1)^3 kept by is_in_first_BZ (src/EPP/DielectricFunction.cpp:35-57).
 * The reference seeds std::mt19937 from std::random_device; this sampler is
 * deterministic: splitmix64 draws from `seed`, taken x, y, z in order,
 * u = (draw >> 11) * 2^-53, k = 2u - 1.  Host-only. */
int uepm_bz_sample(uint64_t seed, int64_t n, double* k_out);

/* A solve plan: one material, one basis, fixed flags and band count.
 * Replaces the per-thread Hamiltonian objects of Compute_parallel
 * (src/EPP/BandStructure.cpp:168-171) and the k-independent local matrix
 * built by Hamiltonian::SetConstantNonDiagonalMatrix (Hamiltonian.cpp:32-48).
 * G: N x 3 int32 (host); nb: bands kept per k (after the SOC doubling the
 * reference applies in Initialize, BandStructure.cpp:78-82, i.e. nb counts
 * rows of the 2N x 2N spectrum when UEPM_SOC is set).
 * SOC with a period outside {2,3,4} -> UEPM_ERR_RUNTIME
 * (SpinOrbitFunctional.cpp:59,71). */
int uepm_plan_create(uepm_ctx* ctx, const uepm_material* mat, const int32_t* G, int N,
                     unsigned flags, int nb, uepm_plan** out);
int uepm_plan_destroy(uepm_plan* plan);
/* Matrix dimension (N, or 2N with SOC). */
int uepm_plan_dim(const uepm_plan* plan);

/* Batched band solve over nk k-points: one call replaces
 * BandStructure::Compute_parallel (BandStructure.cpp:151-201), i.e. per k
 * SetMatrix + Diagonalize + copy of the lowest nb eigenvalues
 * (+ compute_gradient_at_level, Hamiltonian.cpp:146-194).
 *   k      : nk x 3 doubles, reduced units
 *   evals  : nk x nb doubles, ascending, raw (before AdjustValues)
 *   evecs  : nk x nb x dim complex128 as interleaved (re,im) doubles, each
 *            band's vector contiguous and of unit norm (phase arbitrary, as
 *            Eigen's); NULL unless the plan has UEPM_VECTORS
 *   grads  : nk x nb x 3 doubles in eV*m; NULL unless UEPM_GRADIENTS
 * Host buffers; the call is synchronous. */
int uepm_solve(uepm_plan* plan, const double* k, int64_t nk, double* evals, double* evecs,
               double* grads);

/* Same on device buffers, enqueued on `stream` (a hipStream_t; NULL = the
 * default stream).  Returns after enqueueing. */
int uepm_solve_device(uepm_plan* plan, const double* d_k, int64_t nk, double* d_evals,
                      double* d_evecs, double* d_grads, void* stream);

/* Hamiltonian::SetMatrix + the hermitisation step of Diagonalize
 * (Hamiltonian.cpp:50-127): writes the full dim x dim complex128 matrix
 * 0.5*(H + H^dagger) for each k, column-major, interleaved (re,im).
 * d_H: nk * dim * dim * 2 doubles.  Device buffers, enqueued on `stream`. */
int uepm_assemble_device(uepm_plan* plan, const double* d_k, int64_t nk, double* d_H, void* stream);

/* Per-stage device timing of the last uepm_solve_device call, enabled with
 * uepm_plan_set_timing(plan, 1): for stage index `stage` (0-based; beyond
 * the last -> UEPM_ERR_OUT_OF_RANGE) the summed kernel time in milliseconds
 * (hipEvents recorded on the launch stream), the number of launches and the
 * kernel's name.  No counterpart in the reference (it times the whole k loop
 * in seconds, BandStructure.cpp:153,198-199). */
int uepm_plan_set_timing(uepm_plan* plan, int enable);
int uepm_plan_last_timing(const uepm_plan* plan, int stage, double* ms, int64_t* launches,
                          const char** name);

/* ---- configuration, tuning and workspace -----------------------------------
 * No counterpart in the reference (its solve has no tunables). */
typedef struct uepm_config {
    int32_t tridiag;   /* 0 automatic; 1 LDS-blocked reduction; 2 register-resident
                          complex64 reduction (UEPM_MIXED plans, dim <= 192)          */
    int32_t tri_block; /* LDS-blocked reduction workgroup size: 256 or 512 (0 = auto) */
    int32_t bt_bands;  /* bands per back-transform workgroup: 4, 8 or 16 (0 = 16)    */
    int32_t max_chunk; /* k-points per launch (0 = automatic: <= 4096, <= 4 GiB of H) */
} uepm_config;
int uepm_plan_get_config(const uepm_plan* plan, uepm_config* cfg);
/* Values outside the ranges above -> UEPM_ERR_INVALID_ARG (plan unchanged). */
int uepm_plan_set_config(uepm_plan* plan, const uepm_config* cfg);
/* Autotuner: times every candidate configuration on the device k-points
 * d_k[0..nk) (one warm-up and two timed solves each, on `stream`), keeps the
 * fastest in the plan and returns it in *chosen (may be NULL). */
int uepm_plan_tune(uepm_plan* plan, const double* d_k, int64_t nk, void* stream, uepm_config* chosen);
/* Device workspace that solves of up to nk k-points need, and its allocation
 * ahead of time: after uepm_plan_reserve(plan, nk) such solves allocate nothing
 * (required before they are captured into a graph). */
int uepm_plan_workspace_bytes(const uepm_plan* plan, int64_t nk, uint64_t* bytes);
int uepm_plan_reserve(uepm_plan* plan, int64_t nk);

/* ---- batch graphs ----------------------------------------------------------
 * Solves of many plans (materials, cutoffs, SOC on/off: uneven dims) captured
 * once into one hipGraph and replayed with one launch.  Items of one plan run
 * in order on one branch; plans are spread over `branches` concurrent branches,
 * heaviest first.  Every plan's workspace is reserved at creation, so the graph
 * allocates nothing.  Plans must not be used elsewhere while the graph runs.
 * No counterpart in the reference (its k loop is per material). */
typedef struct uepm_batch_item {
    uepm_plan* plan;
    const double* d_k;  /* nk x 3, device */
    int64_t nk;
    double* d_evals;    /* nk x nb, device */
    double* d_evecs;    /* NULL or nk x nb x dim x 2 (plan with UEPM_VECTORS) */
    double* d_grads;    /* NULL or nk x nb x 3 (plan with UEPM_GRADIENTS) */
} uepm_batch_item;
typedef struct uepm_batch uepm_batch;
int uepm_batch_create(uepm_ctx* ctx, const uepm_batch_item* items, int nitems, int branches, uepm_batch** out);
int uepm_batch_launch(uepm_batch* batch, void* stream);
int uepm_batch_destroy(uepm_batch* batch);

/* ---- serving ----------------------------------------------------------------
 * Continuous batching of band-structure requests on one device: a scheduler
 * thread admits queued requests in arrival order while an in-flight slot is
 * free and the plan workspaces fit `hbm_budget_bytes`; every admitted request
 * runs on its own stream, so requests overlap on the device.  Plans are pooled
 * per (material, basis, flags, bands) with resident workspaces; idle plans are
 * evicted oldest first when the budget needs room.  A request that cannot fit
 * even alone fails with UEPM_ERR_NO_MEMORY.  Eigenvalues only (no vectors).
 * No counterpart in the reference. */
typedef struct uepm_request {
    const uepm_material* material;
    int32_t nearest_neighbors; /* basis shells, as in uepm_basis_vectors     */
    uint32_t flags;            /* UEPM_NONLOCAL | UEPM_SOC | UEPM_MIXED       */
    int32_t nb;
    const double* k;           /* host, nk x 3 (copied at submission)        */
    int64_t nk;
} uepm_request;
typedef struct uepm_server_stats {
    uint64_t submitted, completed, kpoints, evictions, used_bytes;
    int32_t inflight, queued;
} uepm_server_stats;
typedef struct uepm_server uepm_server;
int uepm_server_create(uepm_ctx* ctx, uint64_t hbm_budget_bytes, int max_inflight, uepm_server** out);
int uepm_server_submit(uepm_server* server, const uepm_request* request, uint64_t* ticket);
/* Blocks until the request is done: evals (nk x nb, may be NULL) and the
 * latency from submission to completion, and the part spent queued. */
int uepm_server_wait(uepm_server* server, uint64_t ticket, double* evals, double* latency_s, double* queue_s);
int uepm_server_stats_get(uepm_server* server, uepm_server_stats* stats);
/* Drains the queue and the running requests, then frees everything. */
int uepm_server_destroy(uepm_server* server);

/* ---- device plumbing ------------------------------------------------------
 * Device memory, streams and synchronisation on the context's device, for
 * callers without a GPU runtime of their own (scripts, tests, FFI hosts).
 * No counterpart in the reference, which has no device. */
int uepm_device_alloc(uepm_ctx* ctx, uint64_t bytes, void** ptr);
int uepm_device_free(uepm_ctx* ctx, void* ptr);
int uepm_memcpy_htod(uepm_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int uepm_memcpy_dtoh(uepm_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int uepm_memset_device(uepm_ctx* ctx, void* dst, int value, uint64_t bytes);
/* Device-to-device copy on `stream` (0: the default stream), asynchronous;
 * both pointers 16-byte aligned. */
int uepm_memcpy_dtod(uepm_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);
int uepm_stream_create(uepm_ctx* ctx, void** stream);
int uepm_stream_destroy(uepm_ctx* ctx, void* stream);
int uepm_stream_synchronize(uepm_ctx* ctx, void* stream);
int uepm_device_synchronize(uepm_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* UEPM_H */
