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
 * (SpinOrbitFunctional.cpp:59,71).  A matrix dimension (N, or 2N with SOC)
 * above UEPM_MAX_DIM -> UEPM_ERR_OUT_OF_RANGE: the reductions and
 * back-transforms keep up to 9 rows per lane of a wave (n = 15 shells with SOC
 * is 566, n = 16 is 614; the reference's Eigen solver has no such limit). */
#define UEPM_MAX_DIM 576
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
 * Host buffers; the call is synchronous.  A non-finite k-point ->
 * UEPM_ERR_INVALID_ARG before any device work.  A k-point whose eigen-solve
 * fails (H(k) not finite, e.g. overflow at huge |k|) gets NaN eigenvalues and
 * vectors; the others are solved and every output is written, and the call
 * returns UEPM_ERR_NOT_CONVERGED (the reference's Eigen solver reports
 * NoConvergence and Diagonalize throws "Eigenvalue decomposition failed",
 * Hamiltonian.cpp:130-133). */
int uepm_solve(uepm_plan* plan, const double* k, int64_t nk, double* evals, double* evecs,
               double* grads);

/* Same on device buffers, enqueued on `stream` (a hipStream_t; NULL = the
 * default stream).  Returns after enqueueing.  A solve of more than one chunk
 * (uepm_config.lanes) also runs on a stream the plan owns, forked from and
 * joined back into `stream`: later work on `stream` sees every result.
 * Solves of one plan are ordered whatever streams they are enqueued on: each
 * waits for the plan's previous solve (outside graph capture), since they
 * share the plan's workspace.  k-points are not checked on the host; failed
 * k-points (NaN eigenvalues and vectors) are counted on the device and read
 * with uepm_plan_status. */
int uepm_solve_device(uepm_plan* plan, const double* d_k, int64_t nk, double* d_evals,
                      double* d_evecs, double* d_grads, void* stream);

/* Failures of the plan's last solve: waits for `stream` (the stream the solve
 * was enqueued on), then writes the number of k-points whose eigen-solve
 * failed and the lowest index among them (-1 when none); either pointer may
 * be NULL.  Returns UEPM_ERR_NOT_CONVERGED when any failed, else UEPM_OK.
 * Replaces the per-k ComputationInfo check of Hamiltonian::Diagonalize
 * (Hamiltonian.cpp:130-133). */
int uepm_plan_status(uepm_plan* plan, void* stream, int64_t* n_failed, int64_t* first_failed);

/* Same as uepm_solve_device, without allocating.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * call makes no hipMalloc / hipFree and creates no stream or event: the chunk
 * workspace is a UEPM_NOALLOC_SCRATCH_BYTES device array of the library that
 * each GPU holds from its first uepm_ctx_create until the process ends.  The
 * solve runs in chunks of the most k-points (at most 4096) whose workspace
 * fits the array (uepm_plan_workspace_bytes(plan, 1) per k-point), one chunk at
 * a time on `stream`; results are those of uepm_solve_device.  A plan that
 * needs more than the array for one k-point -> UEPM_ERR_OUT_OF_RANGE.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads, other plans) race and must be ordered by the
 * caller.  Stage timing (uepm_plan_set_timing) is not recorded.  Failures are
 * read with uepm_plan_status as after uepm_solve_device. */
#define UEPM_NOALLOC_SCRATCH_BYTES (256ull << 20)
int uepm_solve_device_noalloc(uepm_plan* plan, const double* d_k, int64_t nk, double* d_evals,
                              double* d_evecs, double* d_grads, void* stream);

/* Hamiltonian::SetMatrix + the hermitisation step of Diagonalize
 * (Hamiltonian.cpp:50-127): writes the full dim x dim complex128 matrix
 * 0.5*(H + H^dagger) for each k, column-major, interleaved (re,im).
 * d_H: nk * dim * dim * 2 doubles.  Device buffers, enqueued on `stream`. */
int uepm_assemble_device(uepm_plan* plan, const double* d_k, int64_t nk, double* d_H, void* stream);

/* Per-stage device timing of the last uepm_solve_device call, enabled with
 * uepm_plan_set_timing(plan, 1): for stage index `stage` (0-based; beyond
 * the last -> UEPM_ERR_OUT_OF_RANGE) the kernel time in milliseconds, the
 * number of launches and the kernel's name.  The time is the sum over chunks
 * of the interval between hipEvents recorded around the stage on the stream
 * its chunk ran on; with two lanes those intervals overlap in time, so the
 * stage sums add up to about twice the wall time.  In timing mode the solve
 * synchronises `stream` before it returns.  No counterpart in the reference
 * (it times the whole k loop in seconds, BandStructure.cpp:153,198-199). */
int uepm_plan_set_timing(uepm_plan* plan, int enable);
int uepm_plan_last_timing(const uepm_plan* plan, int stage, double* ms, int64_t* launches,
                          const char** name);

/* ---- configuration, tuning and workspace -----------------------------------
 * No counterpart in the reference (its solve has no tunables). */
typedef struct uepm_config {
    int32_t tridiag;   /* 0 automatic (UEPM_MIXED plans above dim 288: LDS-blocked
                          leading steps, then the register grids); 1 LDS-blocked
                          reduction only; 2 register-resident complex64 reduction
                          (UEPM_MIXED plans, dim <= 288)                              */
    int32_t tri_block; /* LDS-blocked reduction workgroup size: 256 or 512 (0 = 512)  */
    int32_t bt_bands;  /* back-transform: 0 automatic (above dim 192, and UEPM_MIXED
                          plans with nb + 4 > 16: the blocked MFMA kernel while X
                          fits the LDS -- complex64 while nb + 4 <= 32, FP64 in
                          passes of 16 bands up to dim 524 at 160 KB; otherwise
                          16); 4, 8 or 16 bands per workgroup of the
                          reflector-by-reflector kernel                               */
    int32_t max_chunk; /* k-points per launch (0 = automatic: <= 4096, <= 4 GiB of H) */
    int32_t lanes;     /* chunks in flight: 0 automatic (2 when a solve spans two or
                          more chunks: alternate chunks on the caller's stream and on
                          a plan-owned stream, twice the workspace); 1 one chunk    */
} uepm_config;
int uepm_plan_get_config(const uepm_plan* plan, uepm_config* cfg);
/* Values outside the ranges above, or a plan held by a batch graph ->
 * UEPM_ERR_INVALID_ARG (plan unchanged). */
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
 * allocates nothing.  The graph addresses the plans' workspaces, so until
 * uepm_batch_destroy each plan is held: uepm_plan_destroy, _set_config and
 * _tune return UEPM_ERR_INVALID_ARG, and so does a solve of the plan that
 * would need a larger workspace than the one reserved.  Other solves of a held
 * plan reuse that workspace: they must not run while the graph runs.  Each
 * replay clears and fills the plans' status words (uepm_plan_status).
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
 * free; every admitted request runs on its slot's stream, so requests overlap
 * on the device.  Each of the max_inflight slots owns a workspace arena of
 * hbm_budget_bytes / max_inflight bytes, allocated on its first request and
 * kept; a request runs in chunks whose workspace fits it, so a running server
 * neither allocates nor frees device memory per request.  Plans are pooled per
 * (material, basis, flags, bands) and hold no workspace; idle plans are
 * destroyed oldest first only past a pool of 512 (stats: evictions).  A request
 * one of whose k-points needs more workspace than a slot's share fails with
 * UEPM_ERR_NO_MEMORY.  used_bytes: the arenas allocated.  Eigenvalues only.
 * A request with a non-finite k-point is refused at submission
 * (UEPM_ERR_INVALID_ARG); one whose eigen-solve fails for some k-point ends
 * with UEPM_ERR_NOT_CONVERGED, its other eigenvalues returned.
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

/* ---- Brillouin-zone mesh files (gmsh MSH 4.1, ASCII) ------------------------
 * Replaces the gmsh calls of MeshBZ::read_mesh_geometry_from_msh_file
 * (src/BZ_MESH/bz_mesh.cpp:67-191: getNodes, getElements of dimension 3),
 * MeshBZ::read_mesh_bands_from_msh_file (:331-425: the NodeData views) and the
 * view writes of MeshBZ::export_selected_bands_to_gmsh (:1426-1545).  Host-only.
 * Vertex index = position of the node in the file; tetrahedra come back as
 * 0-based vertex indices.  Binary MSH, MSH 2 and volume elements other than
 * 4-node tetrahedra -> UEPM_ERR_INVALID_ARG; a malformed file -> UEPM_ERR_RUNTIME
 * (the reference's runtime_error). */
typedef struct uepm_msh uepm_msh;
int uepm_msh_read(const char* path, uepm_msh** out);
int uepm_msh_counts(const uepm_msh* msh, int64_t* n_nodes, int64_t* n_tetra, int32_t* n_views);
/* tags: n_nodes (may be NULL); xyz: n_nodes x 3 as written in the file (may be NULL) */
int uepm_msh_nodes(const uepm_msh* msh, uint64_t* tags, double* xyz);
int uepm_msh_tetrahedra(const uepm_msh* msh, int32_t* vertex_index /* n_tetra x 4 */);
/* Views in file order: name (NUL-terminated, truncated to name_capacity) and
 * components per node (1: band energies, 3: gradients); data is
 * n_nodes x ncomp in vertex order, NaN where the view has no value. */
int uepm_msh_view_info(const uepm_msh* msh, int32_t view, char* name, int32_t name_capacity, int32_t* ncomp);
int uepm_msh_view_data(const uepm_msh* msh, int32_t view, double* data);
void uepm_msh_free(uepm_msh* msh);
/* Appends one NodeData view (`ncomp` values per node, ASCII, %.17g) to
 * out_path.  When out_path does not exist it is first created with the mesh
 * sections of mesh_path (its data views left out), as gmsh's
 * PostProcessing.SaveMesh does for the reference's first view. */
int uepm_msh_append_node_view(const char* out_path, const char* mesh_path, const char* name, int32_t ncomp,
                              const uint64_t* tags, int64_t n, const double* data);

/* ---- Brillouin-zone mesh driver and tetrahedron DOS -------------------------
 * Replaces MeshBZ (src/BZ_MESH/bz_mesh.hpp) on the band-solve side: the mesh
 * of the zone, the bands at its vertices (BandsOnBZ), the k-star fan-out from
 * the irreducible wedge, band views in and out, and the density of states
 * (DOS_MeshBZ).  Positions are held as MeshBZ holds them after reading (1/m
 * when normalised by 2 pi / a, bz_mesh.cpp:90-105); energies in eV, gradients
 * in eV m.  Vertex coordinates, bands and the DOS workspace go to the device
 * at the first device call.  ctx may be NULL for a host-only mesh (reading,
 * k-star maps, fan-out, files); its device calls -> UEPM_ERR_INVALID_ARG.
 * Calls on one mesh are not thread-safe (as a plan). */
typedef struct uepm_bzmesh uepm_bzmesh;
/* From arrays: xyz nv x 3 (1/m); tet nt x 4 vertex indices in [0, nv);
 * tetra_in_wedge nt flags or NULL.  No lattice constant: only the explicit-
 * range DOS (uepm_bzmesh_dos) and band setters apply. */
int uepm_bzmesh_create(uepm_ctx* ctx, const double* xyz, int64_t nv, const int32_t* tet, int64_t nt,
                       const uint8_t* tetra_in_wedge, uepm_bzmesh** out);
/* MeshBZ::read_mesh_geometry_from_msh_file (bz_mesh.cpp:67-191) without its
 * final k-star load: positions x 2 pi / a (normalize != 0), degenerate
 * tetrahedra (|V| < 1e-12) dropped, wedge flag per tetrahedron from its
 * barycenter, total volume and reduce-BZ factor V / ((2 pi)^3 / (a^3 / 4)). */
int uepm_bzmesh_read_msh(uepm_ctx* ctx, const char* path, double lattice_constant_m, int normalize_by_fourier_factor,
                         uepm_bzmesh** out);
int uepm_bzmesh_destroy(uepm_bzmesh* mesh);
int uepm_bzmesh_counts(const uepm_bzmesh* mesh, int64_t* n_vertices, int64_t* n_tetra, int32_t* n_bands);
/* Any output may be NULL: xyz nv x 3, tet nt x 4, wedge flags nt, node tags nv. */
int uepm_bzmesh_geometry(const uepm_bzmesh* mesh, double* xyz, int32_t* tet, uint8_t* tetra_in_wedge, uint64_t* tags,
                         double* total_volume, double* reduce_bz_factor);
/* MeshBZ::load_kstar_ibz_to_bz (bz_mesh.cpp:198-272); path NULL reads
 * "<mesh stem>_kstar_ibz_to_bz.txt" in the working directory, as the
 * reference.  Missing file, a row whose vertex is outside the wedge, orbit
 * sizes not adding up to the vertex count or an index out of range ->
 * UEPM_ERR_RUNTIME. */
int uepm_bzmesh_load_kstar(uepm_bzmesh* mesh, const char* path);
/* Not in MeshBZ: writes that file for a mesh closed under O_h (the format of
 * apps/bz_meshing.cpp:644-657): every wedge vertex with its orbit under the
 * 48 signed permutations.  An orbit point that is not a vertex ->
 * UEPM_ERR_RUNTIME. */
int uepm_bzmesh_write_kstar(const uepm_bzmesh* mesh, const char* path);
/* Wedge vertices in k-star file order (ids may be NULL to query count). */
int uepm_bzmesh_wedge_vertices(const uepm_bzmesh* mesh, int64_t* ids, int64_t* count);
/* k-points of compute_band_structure_over_mesh (bz_mesh.cpp:1236-1252),
 * reduced units: the wedge vertices (use_iwedge) or all; k may be NULL. */
int uepm_bzmesh_kpoints(const uepm_bzmesh* mesh, int use_iwedge, double* k, int64_t* count);
/* Adjusted E (count x nb) and G (count x nb x 3) at those k-points -> the
 * vertices, the wedge fan-out with gradients mapped by the signed permutation
 * between the two points (distribute_energies_from_iw_wedge_to_full_bz,
 * bz_mesh.cpp:1383-1424, apply_same_symmetry_operation :1314-1381), band
 * extrema, valence = minimum below 1e-6 eV (:1280-1301). */
int uepm_bzmesh_set_mesh_bands(uepm_bzmesh* mesh, int use_iwedge, const double* E, const double* G, int32_t nb);
/* compute_band_structure_over_mesh (bz_mesh.cpp:1230-1303) in one call: solve
 * on the GPU with a plan carrying UEPM_GRADIENTS, uepm_adjust_values (no
 * conduction shift), uepm_bzmesh_set_mesh_bands.  band_gap may be NULL. */
int uepm_bzmesh_compute_bands(uepm_bzmesh* mesh, uepm_plan* plan, int use_iwedge, double* band_gap);
/* Host copies, vertex-major: E nv x nb, G nv x nb x 3 (either may be NULL). */
int uepm_bzmesh_bands(const uepm_bzmesh* mesh, double* E, double* G, int32_t* nb);
int uepm_bzmesh_band_info(const uepm_bzmesh* mesh, int32_t band, int32_t* is_conduction, int64_t* local_index,
                          double* min_energy, double* max_energy);
int uepm_bzmesh_band_groups(const uepm_bzmesh* mesh, int64_t* n_valence, int64_t* valence_start,
                            int64_t* n_conduction, int64_t* conduction_start);
/* Host E (nv x nb) and G (nv x nb x 3), vertex-major; finite; both required.
 * Band info as after a solve. */
int uepm_bzmesh_set_bands(uepm_bzmesh* mesh, const double* E, const double* G, int32_t nb);
/* Device E (nv x ld) and G (nv x ld x 3) as uepm_solve_device writes them,
 * bands [first, first + nb); asynchronous on stream.  Leaves no host copy:
 * only the explicit-range DOS applies afterwards. */
int uepm_bzmesh_set_bands_device(uepm_bzmesh* mesh, const double* d_E, const double* d_G, int32_t ld, int32_t first,
                                 int32_t nb, void* stream);
/* MeshBZ::compute_dos_at_energy_and_band (bz_mesh.cpp:844-862, use_interp =
 * false) at e_min + i (e_max - e_min)/(n_points - 1), i < n_points: per
 * tetrahedron Tetra::compute_tetra_dos_energy_band (mesh_tetra.cpp:477-525),
 * summed, x reduce_bz_factor, x spin_degeneracy, and with wedge_only over the
 * wedge tetrahedra only, x 48.  States / (eV m^3); energies may be NULL. */
int uepm_bzmesh_dos(uepm_bzmesh* mesh, int32_t band, int32_t n_points, double e_min, double e_max,
                    double reduce_bz_factor, double spin_degeneracy, int wedge_only, double* energies, double* dos);
int uepm_bzmesh_dos_device(uepm_bzmesh* mesh, int32_t band, int32_t n_points, double e_min, double e_max,
                           double reduce_bz_factor, double spin_degeneracy, int wedge_only, double* d_energies,
                           double* d_dos, void* stream);
/* MeshBZ::compute_dos_band_at_band_auto (bz_mesh.cpp:889-916): range = band
 * extrema -/+ 0.1 eV, the mesh's reduce-BZ factor, spin degeneracy 2.
 * use_interp reproduces the reference's stub (all zeros). */
int uepm_bzmesh_dos_auto(uepm_bzmesh* mesh, int32_t band, int32_t n_points, int use_interp, int use_iw,
                         double* energies, double* dos);
/* MeshBZ::export_selected_bands_to_gmsh (bz_mesh.cpp:1426-1545): ASCII views
 * "band_<i>" / "grad_band_<i>" appended through uepm_msh_append_node_view
 * (model_msh_path supplies the mesh when out_path is new). */
int uepm_bzmesh_export_msh(const uepm_bzmesh* mesh, const char* out_path, int64_t n_valence, int64_t n_conduction,
                           int highest_valence_as_band0, const char* model_msh_path, int write_gradients);
/* MeshBZ::export_energies_and_gradients_to_vtk (bz_mesh.cpp:1647-1673). */
int uepm_bzmesh_export_vtk(const uepm_bzmesh* mesh, const char* path);
/* MeshBZ::read_mesh_bands_from_msh_file (bz_mesh.cpp:331-425), n < 0 keeps
 * every band of that kind. */
int uepm_bzmesh_read_bands_msh(uepm_bzmesh* mesh, const char* path, int32_t n_conduction, int32_t n_valence,
                               int auto_shift_conduction_band, int set_positive_valence_band);
/* MeshBZ::apply_scissor (bz_mesh.cpp:444-461). */
int uepm_bzmesh_apply_scissor(uepm_bzmesh* mesh, double value);

/* ---- dielectric function ------------------------------------------------------
 * DielectricFunction::compute_dielectric_function
 * (src/EPP/DielectricFunction.cpp:83-165) over the k-points it is given (the
 * reference's [offset, offset + nb_kpoints) slice): for each q (|q| <= 1e-15
 * becomes (1e-15, 1e-15, 1e-15)) and energy E,
 *   out[q][E] = sum_k sum_{c = 4..nb-1} sum_{v = 0..3} |x_c(k+q) . x_v(k)|^2
 *               [ (D - E)/((D - E)^2 + eta^2) + (D + E)/((D + E)^2 + eta^2) ],
 * D = E_c(k+q) - E_v(k), with the reference's bilinear overlap (Eigen's
 * adjoint().dot() conjugates twice).  Eigenpairs from the plan, which must
 * carry UEPM_VECTORS, no UEPM_SOC (SetMatrix(k, nonlocal) has SOC off), and
 * nb > 4.  k, q, energies, out: host; out nq x n_energies. */
int uepm_dielectric(uepm_plan* plan, const double* k, int64_t nk, const double* q, int64_t nq, const double* energies,
                    int32_t n_energies, double eta_smearing, double* out);
/* DielectricFunction::merge_results renormalisation (DielectricFunction.cpp:
 * 167-214) in place: eps = 1 + (2 pi / |q|^2) S / total_nk (|q|^2 = 1e-14 at
 * q = 0).  Host-only. */
int uepm_dielectric_merge(const double* q, int64_t nq, int32_t n_energies, int64_t total_nk, double* eps);
/* DielectricFunction::apply_kramers_kronig (DielectricFunction.cpp:216-250):
 * eps_imag = K (eps_real - 1), K_ij = -(2/pi) i / (j^2 - i^2), K_ii = 0, per q
 * row.  Host-only. */
int uepm_kramers_kronig(const double* eps_real, int64_t nq, int32_t n_energies, double* eps_imag);

/* ---- eigenstates on the mesh (BZ_States) ------------------------------------
 * BZ_States::compute_eigenstates and compute_shifted_eigenstates
 * (src/BZ_MESH/bz_states.cpp:24-85): eigenpairs at every mesh vertex,
 * k = position / (2 pi / a) with a the plan's lattice constant, or at
 * k + q_shift (q_shift in reduced units; NULL selects the unshifted set).
 * The plan must carry UEPM_VECTORS and no UEPM_SOC (the reference calls
 * SetMatrix(k, false), local and SOC off, and keeps nb eigenvector columns).
 * Eigenpairs stay on the plan's device inside the mesh object. */
int uepm_bzmesh_compute_states(uepm_bzmesh* mesh, uepm_plan* plan, const double* q_shift);
/* Host copies of a state set (shifted = 0: k, 1: k + q): evals nv x nb,
 * evecs nv x nb x dim complex (interleaved re, im), each may be NULL; nb and
 * dim may be NULL.  The reference keeps all dim eigenvalues per vertex; only
 * the nb of the plan are computed here. */
int uepm_bzmesh_states(const uepm_bzmesh* mesh, int shifted, double* evals, double* evecs, int32_t* nb, int32_t* dim);
/* BZ_States::compute_dielectric_function (bz_states.cpp:95-164): per
 * tetrahedron, |V|/4 times the sum over its four vertices of
 *   sum_{c = 4..nb-1} sum_{v = 0..3} |x_c(k+q) . x_v(k)|^2
 *     [ (D - E)/((D - E)^2 + eta^2) + (D + E)/((D + E)^2 + eta^2) ],
 * D = E_c(k+q) - E_v(k), summed over the tetrahedra into S(E); then
 * eps(E) = 1 + (q_e^2 / (eps_0 |q|^2)) / q_e * 2 / (2 pi)^3 * S(E) with q the
 * shift in 1/m (|q| = 0 gives the reference's IEEE infinities).  Needs both
 * state sets, nb > 4.  eps: n_energies; sums (S, may be NULL): n_energies;
 * total_volume (sum of |V|, may be NULL). */
int uepm_bzmesh_states_dielectric(uepm_bzmesh* mesh, const double* energies, int32_t n_energies, double eta_smearing,
                                  double* eps, double* sums, double* total_volume);
/* BZ_States::export_dielectric_function (bz_states.cpp:172-179):
 * "<prefix>_dielectric_function.csv", header "Energy (eV),Dielectric function",
 * rows "E,eps" in default stream formatting.  Host-only. */
int uepm_write_dielectric_csv(const char* prefix, const double* energies, const double* eps, int32_t n_energies);
/* BZ_States::export_full_eigenstates (bz_states.cpp:181-201): removes and
 * recreates <root>/eigenstates (root NULL or "": the working directory) and
 * writes per vertex k eigenvalues/eigenvalues_<k>.txt,
 * eigenvalues/shiftedeigenvalues_<k>.txt (one row of nb values) and
 * eigenvectors/eigenvectors_<k>.txt, eigenvectors/shiftedeigenvectors_<k>.txt
 * (dim rows x nb columns of "(re,im)"), in Eigen's default stream format.
 * The reference's eigenvalue rows hold all dim eigenvalues. */
int uepm_bzmesh_export_states(const uepm_bzmesh* mesh, const char* root);

/* BandStructure::AdjustValues (src/EPP/BandStructure.cpp:203-262) over
 * E (nk x nb) in place: the first level pair from level 2 up whose gap
 * exceeds 0.35 eV sets the valence maximum (fallback: level 3), which is
 * subtracted (the conduction minimum from levels >= 4, or 8 with soc, when
 * min_cond_to_zero).  band_gap may be NULL.  Host-only. */
int uepm_adjust_values(double* E, int64_t nk, int32_t nb, int soc, int min_cond_to_zero, double* band_gap);

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
