// uepm_abi.cpp — host side of the C ABI (include/uepm.h).
//
// Replaces the reference's per-thread Hamiltonian loop
// (BandStructure::Compute_parallel, src/EPP/BandStructure.cpp:151-201) by
// batched device stages; also restates the host-only helpers the path needs
// (basis generation BandStructure.cpp:26-57, k path SymmetryPoints.cpp:47-147,
// BZ sampling DielectricFunction.cpp:35-57).  No exception crosses the ABI:
// errors become status codes plus a per-thread message (uepm_last_error).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "uepm_internal.hpp"

namespace uepm_host {
std::string& last_error() {
    thread_local std::string msg;
    return msg;
}
}  // namespace uepm_host

using namespace uepm_host;

namespace {

// src/Common/physical_constants.hpp:18-41, same operation order.
constexpr double kPi = 3.14159265358979323846;
constexpr double kH = 6.62607015e-34;
constexpr double kQe = 1.602176634e-19;
constexpr double kHbar = kH / (2.0 * kPi);
constexpr double kMe = 9.1093837015e-31;
constexpr double kBohr = 5.29177210903e-11;

const char* kStageNames[kNumStages] = {"assemble_kernel", "tridiag_kernel", "tri_eig_kernel", "backtransform_kernel",
                                       "refine_kernel"};

}  // namespace

namespace uepm_host {

int64_t plan_chunk(const uepm_plan* p, int64_t nk) {
    const size_t mat_bytes = sizeof(double2) * (size_t)p->dim * p->dim;
    int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(4096, (int64_t)((4ull << 30) / mat_bytes)));
    if (p->cfg.max_chunk > 0) chunk = std::min<int64_t>(chunk, p->cfg.max_chunk);
    return std::max<int64_t>(1, std::min(chunk, nk));
}

namespace {
// Trailing size of the precision form's blocked reduction (n > 288) that the
// register grids take over; 0 when the configuration asks for the blocked kernel.
int mixed_c64_tail(const uepm_plan* p) {
    if (p->tri_reg || p->cfg.tridiag == 1 || p->dim <= uepm::tridiag_reg_max_n()) return 0;
    return uepm::tridiag_reg_c64_tail(p->dim);
}

struct Sizes {
    size_t H, d, e, tau, Vf, tauf, S32, Ev, X0, S, Z, W, X, kstat;
};
Sizes plan_sizes(const uepm_plan* p, int64_t chunk) {
    const int n = p->dim, nb = p->nb;
    const bool mixed = p->flags & UEPM_MIXED;
    const bool vec = p->flags & UEPM_VECTORS;
    const int nbr = mixed ? std::min(n, nb + kMixedExtra) : nb;
    const bool stage_vec = mixed || vec;
    Sizes z{};
    z.H = sizeof(double2) * (size_t)n * n * chunk;
    z.d = z.e = sizeof(double) * (size_t)n * chunk;
    z.tau = sizeof(double2) * (size_t)n * chunk;
    if (mixed) {
        z.Vf = sizeof(float2) * (size_t)n * n * chunk;
        z.tauf = sizeof(float2) * (size_t)n * chunk;
        z.S32 = p->tri_reg ? uepm::tridiag_reg_work_bytes(n) * chunk
                           : std::max(sizeof(float2) * (size_t)p->tp32.scratch_elems,
                                      uepm::tridiag_reg_work_bytes(mixed_c64_tail(p))) *
                                 chunk;
        z.Ev = sizeof(double) * (size_t)nbr * chunk;
        z.X0 = sizeof(double2) * (size_t)n * nbr * chunk;
    } else {
        // blocked-reduction scratch, then (after it) the FP64 register grids' hand-over blocks
        const int t = uepm::tridiag_reg_c128_tail(n);
        const size_t blocked = t < n ? sizeof(double2) * (size_t)p->tp.scratch_elems : 0;
        z.S = std::max(blocked, uepm::tridiag_reg_c128_work_bytes(t)) * chunk;
    }
    if (stage_vec) {
        z.Z = sizeof(double) * (size_t)n * nbr * chunk;
        z.W = sizeof(double) * 4 * (size_t)n * nbr * chunk;  // U factor (3 rows) + iterate
        if (!mixed) z.X = sizeof(double2) * (size_t)n * nb * chunk;
    }
    z.kstat = sizeof(int) * (size_t)chunk;
    return z;
}
}  // namespace

int plan_lanes(const uepm_plan* p, int64_t nk) { return (p->cfg.lanes == 1 || nk <= plan_chunk(p, nk)) ? 1 : 2; }

size_t plan_workspace_bytes(const uepm_plan* p, int64_t chunk, int lanes) {
    const Sizes z = plan_sizes(p, chunk);
    return (size_t)lanes *
           (z.H + z.d + z.e + z.tau + z.Vf + z.tauf + z.S32 + z.Ev + z.X0 + z.S + z.Z + z.W + z.X + z.kstat);
}

int plan_reserve(uepm_plan* p, int64_t chunk, int lanes) {
    const Sizes z = plan_sizes(p, chunk);
    for (int l = 0; l < lanes; ++l) {
        Lane& w = p->lane[l];
        if (l == 0 && p->pins > 0) {
            // a batch graph holds lane 0's buffers: they may be used, not reallocated
            const DeviceBuffer* b[] = {&w.H, &w.d, &w.e, &w.tau, &w.Vf, &w.tauf, &w.S32, &w.Ev, &w.X0,
                                       &w.S, &w.Z, &w.W, &w.X, &w.kstat};
            const size_t need[] = {z.H, z.d, z.e, z.tau, z.Vf, z.tauf, z.S32, z.Ev, z.X0, z.S, z.Z, z.W, z.X, z.kstat};
            for (int i = 0; i < 14; ++i)
                if (need[i] > b[i]->bytes)
                    return fail(UEPM_ERR_INVALID_ARG,
                                "the plan's workspace is held by a batch graph: solves of up to the reserved size "
                                "only, until uepm_batch_destroy");
        }
        HIP_OK(w.H.ensure(z.H));
        HIP_OK(w.d.ensure(z.d));
        HIP_OK(w.e.ensure(z.e));
        HIP_OK(w.tau.ensure(z.tau));
        if (z.Vf) HIP_OK(w.Vf.ensure(z.Vf));
        if (z.tauf) HIP_OK(w.tauf.ensure(z.tauf));
        if (z.S32) HIP_OK(w.S32.ensure(z.S32));
        if (z.Ev) HIP_OK(w.Ev.ensure(z.Ev));
        if (z.X0) HIP_OK(w.X0.ensure(z.X0));
        HIP_OK(w.kstat.ensure(z.kstat));
        if (z.S) HIP_OK(w.S.ensure(z.S));
        if (z.Z) HIP_OK(w.Z.ensure(z.Z));
        if (z.W) HIP_OK(w.W.ensure(z.W));
        if (z.X) HIP_OK(w.X.ensure(z.X));
    }
    return UEPM_OK;
}

size_t plan_carve(const uepm_plan* p, int64_t chunk, char* base, LaneView* v) {
    const Sizes z = plan_sizes(p, chunk);
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        if (!bytes) return nullptr;
        void* ptr = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return ptr;
    };
    LaneView t;
    t.H = static_cast<double2*>(take(z.H));
    t.d = static_cast<double*>(take(z.d));
    t.e = static_cast<double*>(take(z.e));
    t.tau = static_cast<double2*>(take(z.tau));
    t.Vf = static_cast<float2*>(take(z.Vf));
    t.tauf = static_cast<float2*>(take(z.tauf));
    t.S32 = static_cast<float2*>(take(z.S32));
    t.Ev = static_cast<double*>(take(z.Ev));
    t.X0 = static_cast<double2*>(take(z.X0));
    t.S = static_cast<double2*>(take(z.S));
    t.Z = static_cast<double*>(take(z.Z));
    t.W = static_cast<double*>(take(z.W));
    t.X = static_cast<double2*>(take(z.X));
    t.kstat = static_cast<int*>(take(z.kstat));
    if (base && v) *v = t;
    return off;
}

void plan_apply_config(uepm_plan* p) {
    const bool mixed = p->flags & UEPM_MIXED;
    const bool reg_ok = mixed && p->dim <= uepm::tridiag_reg_max_n();
    p->tri_reg = p->cfg.tridiag != 1 && reg_ok;
    const int bs = p->cfg.tri_block == 256 ? 256 : 512;
    p->tp = uepm::tri_plan(p->dim, (int)sizeof(double2), p->ctx->lds_limit, bs);
    p->tp32 = uepm::tri_plan(p->dim, (int)sizeof(float2), p->ctx->lds_limit, bs);
    p->bt_bands = p->cfg.bt_bands > 0 ? p->cfg.bt_bands : 16;
    const int nbr = mixed ? std::min(p->dim, p->nb + kMixedExtra) : p->nb;
    // automatic: the blocked MFMA back-transform above dim 192 (both forms) or, in the
    // precision form, past 16 Ritz vectors, where it is 2.5-3.6x faster; below, the
    // reflector-by-reflector kernel (no LDS) shares the chip better with the other
    // lane's register grids, and in FP64 it is as fast or faster
    const int lds = p->ctx->lds_limit;
    p->bt_wy = p->cfg.bt_bands == 0 && (mixed ? (p->dim > 192 || nbr > 16) && uepm::apply_q_wy_fits(p->dim, nbr, lds)
                                              : p->dim > 192 && uepm::apply_q_wy_f64_fits(p->dim, lds));
}

}  // namespace uepm_host

namespace uepm_host {

bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone;
}

int solve_args_ok(const uepm_plan* p, int64_t nk, const double* dk, const double* devals, const double* devecs,
                  const double* dgrads, const char* who) {
    if (!p) return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": null plan");
    if (nk < 0) return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": nk < 0");
    if (nk == 0) return UEPM_OK;
    if (!dk || !devals) return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": null k or eigenvalue buffer");
    if (devecs && !(p->flags & UEPM_VECTORS))
        return fail(UEPM_ERR_INVALID_ARG, "eigenvectors requested from a plan without UEPM_VECTORS");
    if (dgrads && !(p->flags & UEPM_GRADIENTS))
        return fail(UEPM_ERR_INVALID_ARG, "gradients requested from a plan without UEPM_GRADIENTS");
    return UEPM_OK;
}

int solve_chunks(uepm_plan* p, const double* dk, int64_t nk, double* devals, double* devecs, double* dgrads,
                 hipStream_t s, const LaneView* views, int lanes, int64_t chunk, bool no_alloc) {
    const int n = p->dim, nb = p->nb;
    const bool vec = p->flags & UEPM_VECTORS;
    const bool need_vec = vec && (devecs || dgrads);
    const bool mixed = p->flags & UEPM_MIXED;
    const int nbr = mixed ? std::min(n, nb + kMixedExtra) : nb;
    const bool stage_vec = mixed || need_vec;  // eigenvectors of T are needed
    const bool capturing = stream_capturing(s);
    const int nchunks = (int)((nk + chunk - 1) / chunk);
    // after the plan's previous solve, whichever stream it ran on (the
    // workspace and the status words are the plan's); the no-allocation path
    // creates no event, so it orders only after solves that created one
    if (!capturing) {
        if (p->lane_done[0]) HIP_OK(hipStreamWaitEvent(s, p->lane_done[0], 0));
        else if (!no_alloc) HIP_OK(hipEventCreateWithFlags(&p->lane_done[0], hipEventDisableTiming));
    }
    unsigned long long* status = p->status.as<unsigned long long>();
    HIP_OK(uepm::launch_status_reset(status, s));
    if (lanes > 1) {
        // the side stream starts after the caller's earlier work and the caller's
        // stream resumes after the side stream's last chunk (join below)
        if (!p->side) HIP_OK(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
        if (!p->fork) HIP_OK(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
        if (!p->join) HIP_OK(hipEventCreateWithFlags(&p->join, hipEventDisableTiming));
        HIP_OK(hipEventRecord(p->fork, s));
        HIP_OK(hipStreamWaitEvent(p->side, p->fork, 0));
    }
    const bool timing = p->timing && !no_alloc;
    if (timing) {
        const size_t need = (size_t)nchunks * kNumStages * 2;
        while (p->ev.size() < need) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            p->ev.push_back(e);
        }
    }
    auto rec = [&](int chunk_i, int stage, int end) -> hipError_t {
        if (!timing) return hipSuccess;
        return hipEventRecord(p->ev[((size_t)chunk_i * kNumStages + stage) * 2 + end],
                              chunk_i % lanes ? p->side : s);
    };
    bool used[kNumStages] = {true, true, true, stage_vec, mixed};
    p->stage_name[kTridiag] = mixed ? (p->tri_reg              ? "tridiag_reg_kernel<c64>"
                                       : mixed_c64_tail(p) > 0 ? "tridiag_blocked+reg_kernel<c64>"
                                                               : "tridiag_blocked_kernel<c64>")
                                    : uepm::tridiag_reg_c128_tail(n) == n ? "tridiag_reg_kernel<c128>"
                                    : uepm::tridiag_reg_c128_tail(n) > 0 ? "tridiag_blocked+reg_kernel<c128>"
                                                                         : "tridiag_blocked_kernel<c128>";
    p->stage_name[kTriEig] = "tri_eig_wave_kernel";
    p->stage_name[kBack] = mixed ? "cluster_orth+apply_q_kernel<c64>" : "cluster_orth+apply_q_kernel<c128>";
    p->stage_name[kRefine] = uepm::refine_kernel_name(nbr);
    // the chunks; on an error the side stream is still joined into the caller's
    auto run_chunks = [&]() -> int {
        for (int ci = 0; ci < nchunks; ++ci) {
            const int64_t c0 = (int64_t)ci * chunk;
            const LaneView& w = views[ci % lanes];
            const hipStream_t cs = ci % lanes ? p->side : s;
            const int64_t c = std::min(chunk, nk - c0);
            const double* kc = dk + 3 * c0;
            double2* Hc = w.H;
            int* ks = w.kstat;
            HIP_OK(rec(ci, kAssemble, 0));
            HIP_OK(uepm::launch_assemble(p->P, p->G.as<int32_t>(), kc, c, Hc, n, cs));
            HIP_OK(rec(ci, kAssemble, 1));
            if (mixed) {
                // precision form: complex64 reduction, FP64 tridiagonal eigen-stage,
                // complex64 back-transform, FP64 Rayleigh-Ritz on the FP64 H(k)
                float2* Vc = w.Vf;
                HIP_OK(rec(ci, kTridiag, 0));
                if (p->tri_reg)
                    HIP_OK(uepm::launch_tridiag_reg_c64(Hc, n, n, c, Vc, w.d, w.e,
                                                        w.tauf, w.S32, cs));
                else {
                    // n > 288: the blocked reduction for the leading steps, the register grids
                    // for the trailing block it leaves in place in Vc
                    const int t = mixed_c64_tail(p);
                    const int s0 = n - t;
                    HIP_OK(uepm::launch_tridiag_blocked_c64(p->tp32, Hc, n, c, w.S32, Vc, w.d,
                                                            w.e, w.tauf, cs,
                                                            t > 0 ? s0 : n - 1));
                    if (t > 0) {
                        const int64_t nn = (int64_t)n * n;
                        float2* B = Vc + s0 + (int64_t)s0 * n;
                        HIP_OK(uepm::launch_tridiag_reg_c64_block(B, nn, t, n, c, B, n, nn, w.d + s0,
                                                                  w.e + s0, w.tauf + s0, n,
                                                                  w.S32, cs));
                    }
                }
                HIP_OK(rec(ci, kTridiag, 1));
                HIP_OK(rec(ci, kTriEig, 0));
                // T carries complex64 rounding (~1e-7 ||T||): brackets of 1e-10 ||T|| are ample
                // for inverse iteration; the FP64 refinement sets the eigenvalues
                HIP_OK(uepm::launch_tri_eig(w.d, w.e, n, nbr, c, w.Ev,
                                            w.Z, w.W, kMixedInvSweeps, kMixedBisectTol, ks,
                                            nullptr, c0, cs));
                HIP_OK(rec(ci, kTriEig, 1));
                HIP_OK(rec(ci, kBack, 0));
                HIP_OK(uepm::launch_cluster_orth(w.d, w.e, w.Ev,
                                                 w.Z, n, nbr, c, cs));
                HIP_OK(uepm::launch_apply_q_c64(Vc, n, n, w.tauf, w.Z, nbr,
                                                w.X0, p->P, kc, p->G.as<int32_t>(), c,
                                                p->bt_wy ? 0 : p->bt_bands, p->ctx->lds_limit, cs));
                HIP_OK(rec(ci, kBack, 1));
                HIP_OK(rec(ci, kRefine, 0));
                HIP_OK(uepm::launch_refine(Hc, n, n, w.X0, nbr, nb, devals + c0 * nb,
                                           devecs ? reinterpret_cast<double2*>(devecs) + (size_t)c0 * nb * n : nullptr,
                                           dgrads ? dgrads + 3 * c0 * nb : nullptr, p->P, kc, p->G.as<int32_t>(), c, ks,
                                           status, c0, cs));
                HIP_OK(rec(ci, kRefine, 1));
                continue;
            }
            HIP_OK(rec(ci, kTridiag, 0));
            {
                // the blocked reduction for the leading steps, the FP64 register grids for the
                // trailing block of size t (the whole matrix when n <= 192); reflectors in place in H
                const int t = uepm::tridiag_reg_c128_tail(n);
                const int s0 = n - t;
                const int64_t nn = (int64_t)n * n;
                if (t < n)
                    HIP_OK(uepm::launch_tridiag_blocked_f64(p->tp, Hc, n, c, w.S, Hc, w.d,
                                                            w.e, w.tau, cs,
                                                            t > 0 ? s0 : n - 1));
                if (t > 0) {
                    double2* B = Hc + s0 + (int64_t)s0 * n;
                    HIP_OK(uepm::launch_tridiag_reg_c128(B, nn, t, n, t == n, c, B, n, nn, w.d + s0,
                                                         w.e + s0, w.tau + s0, n,
                                                         w.S, cs));
                }
            }
            HIP_OK(rec(ci, kTridiag, 1));
            HIP_OK(rec(ci, kTriEig, 0));
            HIP_OK(uepm::launch_tri_eig(w.d, w.e, n, nb, c, devals + c0 * nb,
                                        need_vec ? w.Z : nullptr, need_vec ? w.W : nullptr,
                                        need_vec ? kFp64InvSweeps : 0, 0.0, need_vec ? ks : nullptr,
                                        need_vec ? nullptr : status, c0, cs));
            HIP_OK(rec(ci, kTriEig, 1));
            if (need_vec) {
                double2* Xc = devecs ? reinterpret_cast<double2*>(devecs) + (size_t)c0 * nb * n : w.X;
                HIP_OK(rec(ci, kBack, 0));
                HIP_OK(uepm::launch_cluster_orth(w.d, w.e, devals + c0 * nb,
                                                 w.Z, n, nb, c, cs));
                HIP_OK(uepm::launch_apply_q_f64(Hc, n, n, w.tau, w.Z, nb, Xc,
                                                dgrads ? dgrads + 3 * c0 * nb : nullptr, p->P, kc, p->G.as<int32_t>(),
                                                c, ks, status, c0, p->bt_wy ? 0 : p->bt_bands,
                                                p->ctx->lds_limit, cs));
                HIP_OK(rec(ci, kBack, 1));
            }
        }
        return UEPM_OK;
    };
    const int rc = run_chunks();
    if (lanes > 1) {
        HIP_OK(hipEventRecord(p->join, p->side));
        HIP_OK(hipStreamWaitEvent(s, p->join, 0));
    }
    if (!capturing && p->lane_done[0]) HIP_OK(hipEventRecord(p->lane_done[0], s));
    if (rc != UEPM_OK) return rc;
    if (timing) {
        HIP_OK(hipStreamSynchronize(s));
        for (int st = 0; st < kNumStages; ++st) {
            p->stage_ms[st] = 0.0;
            p->stage_launches[st] = used[st] ? nchunks : 0;
            if (!used[st]) continue;
            for (int ci = 0; ci < nchunks; ++ci) {
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, p->ev[((size_t)ci * kNumStages + st) * 2],
                                           p->ev[((size_t)ci * kNumStages + st) * 2 + 1]));
                p->stage_ms[st] += ms;
            }
        }
    }
    return UEPM_OK;
}

}  // namespace uepm_host

namespace uepm_host {
int solve_region(uepm_plan* p, const double* dk, int64_t nk, double* devals, hipStream_t s, char* base,
                 size_t bytes) {
    const int v = solve_args_ok(p, nk, dk, devals, nullptr, nullptr, "server solve");
    if (v != UEPM_OK || nk == 0) return v;
    if (!base) return fail(UEPM_ERR_INVALID_ARG, "server solve: no workspace region");
    const size_t per_k = plan_carve(p, 1, nullptr, nullptr);
    if (per_k > bytes)
        return fail(UEPM_ERR_NO_MEMORY, "server solve: one k-point needs " + std::to_string(per_k) +
                                            " bytes of workspace, the region holds " + std::to_string(bytes));
    HIP_OK(hipSetDevice(p->ctx->device));
    int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(plan_chunk(p, nk), (int64_t)(bytes / per_k)));
    while (chunk > 1 && plan_carve(p, chunk, nullptr, nullptr) > bytes) --chunk;
    LaneView view;
    plan_carve(p, chunk, base, &view);
    return solve_chunks(p, dk, nk, devals, nullptr, nullptr, s, &view, 1, chunk, false);
}
}  // namespace uepm_host

extern "C" {

const char* uepm_last_error(void) { return last_error().c_str(); }

int uepm_ctx_create(int device, uepm_ctx** out) {
    if (!out) return fail(UEPM_ERR_INVALID_ARG, "uepm_ctx_create: null output");
    int n = 0;
    HIP_OK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(UEPM_ERR_INVALID_ARG, "uepm_ctx_create: no such device");
    auto* c = new uepm_ctx;
    c->device = device;
    hipError_t e = hipGetDeviceProperties(&c->prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(UEPM_ERR_DEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        c->lds_limit = lds;
    e = hipSetDevice(device);
    const int st = e == hipSuccess ? noalloc_setup(device)
                                   : fail(UEPM_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    if (st != UEPM_OK) {
        delete c;
        return st;
    }
    *out = c;
    return UEPM_OK;
}

int uepm_ctx_destroy(uepm_ctx* ctx) {
    delete ctx;
    return UEPM_OK;
}

// BandStructure::GenerateBasisVectors (src/EPP/BandStructure.cpp:26-57)
int uepm_basis_vectors(int nn, int32_t* G, int capacity, int* count) {
    static const unsigned G2[] = {
        0,   3,   4,   8,   11,  12,  16,  19,  20,  24,  27,  32,  35,  36,  40,  43,  44,  48,  51,  52,  56,  59,  67,  68,
        75,  76,  80,  83,  84,  88,  96,  99,  104, 107, 108, 115, 116, 120, 123, 128, 131, 132, 136, 139, 140, 144, 147, 152,
        155, 160, 163, 164, 168, 171, 172, 176, 179, 180, 184, 187, 192, 195, 196, 200, 203, 204, 208, 211, 212, 216, 219, 224,
        227, 228, 232, 236, 243, 244, 248, 251, 259, 260, 264, 267, 268, 272, 275, 276, 280, 283, 291, 296, 299, 300, 304, 307,
        308, 312, 315, 320, 323, 324, 331, 332, 339, 355, 356, 360, 363, 371, 376, 384, 387, 395, 420, 451};
    const int nG2 = (int)(sizeof(G2) / sizeof(G2[0]));
    if (nn < 2 || nn > nG2)
        return fail(UEPM_ERR_RUNTIME, "BandStructure::Initialize: GenerateBasisVectors failed (nearestNeighborsNumber must be between 2 and " +
                                          std::to_string(nG2) + ")");
    const unsigned cut = G2[nn - 1];
    const int size = (int)std::ceil(std::sqrt((double)cut));
    int cnt = 0;
    for (int i = -size; i <= size; ++i)
        for (int j = -size; j <= size; ++j)
            for (int k = -size; k <= size; ++k) {
                const int x = -i + j + k, y = i - j + k, z = i + j - k;
                if ((double)(x * x + y * y + z * z) <= cut) {
                    if (G && cnt < capacity) {
                        G[3 * cnt] = x;
                        G[3 * cnt + 1] = y;
                        G[3 * cnt + 2] = z;
                    }
                    ++cnt;
                }
            }
    if (count) *count = cnt;
    return UEPM_OK;
}

// SymmetryPoints::GeneratePoints (src/EPP/SymmetryPoints.cpp:47-147)
int uepm_kpath(const char* labels, unsigned nrPoints, double* out, unsigned* sympos, int* count) {
    if (!labels) return fail(UEPM_ERR_INVALID_ARG, "null path");
    const size_t M = std::strlen(labels);
    if (M < 2) return fail(UEPM_ERR_INVALID_ARG, "Path must contain at least two symmetry points.");
    if (nrPoints < M) return fail(UEPM_ERR_INVALID_ARG, "nrPoints must be at least the number of nodes in the path.");
    struct P3 {
        double x, y, z;
    };
    std::vector<P3> nodes;
    for (size_t s = 0; s < M; ++s) {
        switch (labels[s]) {
            case 'L': nodes.push_back({0.5, 0.5, 0.5}); break;
            case 'G': nodes.push_back({0., 0., 0.}); break;
            case 'X': nodes.push_back({1., 0., 0.}); break;
            case 'W': nodes.push_back({1., 0.5, 0.}); break;
            case 'K': nodes.push_back({0.75, 0.75, 0.}); break;
            case 'U': nodes.push_back({1., 0.25, 0.25}); break;
            default: return fail(UEPM_ERR_INVALID_ARG, std::string("Unknown symmetry point label: ") + labels[s]);
        }
    }
    std::vector<double> seg(M - 1);
    for (size_t i = 1; i < M; ++i) {
        const double dx = nodes[i].x - nodes[i - 1].x, dy = nodes[i].y - nodes[i - 1].y, dz = nodes[i].z - nodes[i - 1].z;
        seg[i - 1] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    const double total = std::accumulate(seg.begin(), seg.end(), 0.0);
    const unsigned R = nrPoints - (unsigned)M;
    int cnt = 0, ns = 0;
    auto push = [&](double x, double y, double z) {
        if (out) {
            out[3 * cnt] = x;
            out[3 * cnt + 1] = y;
            out[3 * cnt + 2] = z;
        }
        ++cnt;
    };
    auto mark = [&]() {
        if (sympos) sympos[ns] = (unsigned)cnt;
        ++ns;
    };
    if (total == 0.0) {
        for (size_t j = 0; j < M; ++j) {
            mark();
            push(nodes[j].x, nodes[j].y, nodes[j].z);
        }
        for (unsigned i = 0; i < R; ++i) push(nodes.back().x, nodes.back().y, nodes.back().z);
    } else {
        std::vector<unsigned> m(M - 1);
        std::vector<double> want(M - 1);
        std::vector<size_t> order(M - 1);
        for (size_t s = 0; s + 1 < M; ++s) {
            want[s] = (seg[s] / total) * (double)R;
            m[s] = (unsigned)std::floor(want[s]);
            order[s] = s;
        }
        const unsigned rem = R - std::accumulate(m.begin(), m.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](size_t p, size_t q) {
            return want[p] - std::floor(want[p]) > want[q] - std::floor(want[q]);
        });
        for (unsigned i = 0; i < rem; ++i) ++m[order[i]];
        mark();
        push(nodes[0].x, nodes[0].y, nodes[0].z);
        for (size_t s = 0; s + 1 < M; ++s) {
            const P3& A = nodes[s];
            const double dx = nodes[s + 1].x - A.x, dy = nodes[s + 1].y - A.y, dz = nodes[s + 1].z - A.z;
            const double denom = (double)(m[s] + 1);
            for (unsigned i = 1; i <= m[s]; ++i) {
                const double t = (double)i / denom;
                push(A.x + dx * t, A.y + dy * t, A.z + dz * t);
            }
            mark();
            push(nodes[s + 1].x, nodes[s + 1].y, nodes[s + 1].z);
        }
    }
    if (count) *count = cnt;
    return UEPM_OK;
}

int uepm_bz_sample(uint64_t seed, int64_t n, double* out) {
    if (n < 0 || (n > 0 && !out)) return fail(UEPM_ERR_INVALID_ARG, "uepm_bz_sample: bad arguments");
    uint64_t state = seed;
    auto draw = [&]() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    };
    int64_t got = 0;
    while (got < n) {
        const double x = draw(), y = draw(), z = draw();
        if (std::fabs(x) <= 1.0 && std::fabs(y) <= 1.0 && std::fabs(z) <= 1.0 &&
            std::fabs(x) + std::fabs(y) + std::fabs(z) <= 3.0 / 2.0) {
            out[3 * got] = x;
            out[3 * got + 1] = y;
            out[3 * got + 2] = z;
            ++got;
        }
    }
    return UEPM_OK;
}

int uepm_plan_create(uepm_ctx* ctx, const uepm_material* m, const int32_t* G, int N, unsigned flags, int nb,
                     uepm_plan** out) {
    if (!ctx || !m || !G || !out) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_create: null argument");
    if (N < 1) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_create: empty basis");
    const bool soc = flags & UEPM_SOC;
    const int dim = soc ? 2 * N : N;
    if (nb < 1 || nb > dim) return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_plan_create: nb must be in [1, dim]");
    if (dim > UEPM_MAX_DIM)
        return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_plan_create: matrix dimension " + std::to_string(dim) +
                                               " exceeds " + std::to_string(UEPM_MAX_DIM) +
                                               ", the largest the reductions take");
    if (!(m->lattice_constant_angstrom > 0.0)) return fail(UEPM_ERR_INVALID_ARG, "lattice constant must be positive");
    if (soc) {
        auto ok = [](int p) { return p == 2 || p == 3 || p == 4; };
        if (!ok(m->period_cation))
            return fail(UEPM_ERR_RUNTIME, "SpinOrbitCorrection::compute_lambda_1: Unsupported period for cation.");
        if (!ok(m->period_anion))
            return fail(UEPM_ERR_RUNTIME, "SpinOrbitCorrection::compute_lambda_2: Unsupported period for anion.");
    }
    if (flags & UEPM_GRADIENTS) flags |= UEPM_VECTORS;
    if ((flags & UEPM_MIXED) && std::min(dim, nb + kMixedExtra) > 28)
        return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_plan_create: the precision form refines at most 24 bands");
    HIP_OK(hipSetDevice(ctx->device));
    auto* p = new uepm_plan;
    p->ctx = ctx;
    p->flags = flags;
    p->N = N;
    p->dim = dim;
    p->nb = nb;
    uepm::DevParams& P = p->P;
    P.N = N;
    P.dim = dim;
    P.nl = (flags & UEPM_NONLOCAL) ? 1 : 0;
    P.soc = soc ? 1 : 0;
    P.nb = nb;
    const double a = 1.0e-10 * m->lattice_constant_angstrom;  // Material.h:108-111
    constexpr double one_eight = 1.0 / 8.0;
    P.a = a;
    P.tau = one_eight * a;
    P.tpa = 2.0 * kPi / a;
    const double diag_factor = (kHbar * kHbar) / (2.0 * kMe * kQe);  // Hamiltonian.cpp:68
    P.kin = P.tpa * P.tpa * diag_factor;
    for (int i = 0; i < 4; ++i) {
        P.VS[i] = m->V_S[i];
        P.VA[i] = m->V_A[i];
    }
    // Material.cpp:142-152, 170
    P.fourier = 2.0 * kPi / a;
    P.diag_factor = std::pow(kHbar, 2) / (2.0 * kMe * kQe);
    const double kF = std::pow(96 * kPi * kPi, 1.0 / 3.0) / a;
    P.kF2 = std::pow(kF, 2.0);
    const double omega = (1.0 / 8.0) * std::pow(a, 3.0);
    P.V_pre = 4.0 * kPi / omega;
    P.tau_over_a = P.tau / a;
    P.well = (m->nl_well_type == 0) ? 0 : 1;
    P.alpha0_a = m->alpha_0_anion;
    P.dfbeta_a = P.diag_factor * m->beta_0_anion;
    P.A2_a = m->A2_anion;
    P.R0_a = m->R0_anion;
    P.R2_a = m->R2_anion;
    P.gpre_a = 5.0 * std::pow(kPi, 1.5) * (std::pow(m->R2_anion, 3.0) / omega);
    P.alpha0_c = m->alpha_0_cation;
    P.dfbeta_c = P.diag_factor * m->beta_0_cation;
    P.A2_c = m->A2_cation;
    P.R0_c = m->R0_cation;
    P.R2_c = m->R2_cation;
    P.gpre_c = 5.0 * std::pow(kPi, 1.5) * (std::pow(m->R2_cation, 3.0) / omega);
    // F_0 contributes only when R0 != 0 (Material.cpp:159,180) and its
    // coefficient A_0 = alpha + diag*beta*(...) is not identically zero.
    P.need_l0_a = (m->R0_anion != 0.0) && (m->alpha_0_anion != 0.0 || m->beta_0_anion != 0.0);
    P.need_l0_c = (m->R0_cation != 0.0) && (m->alpha_0_cation != 0.0 || m->beta_0_cation != 0.0);
    P.need_l2_a = m->A2_anion != 0.0;
    P.need_l2_c = m->A2_cation != 0.0;
    P.same_r0 = P.need_l0_a && P.need_l0_c && m->R0_anion == m->R0_cation;
    // SpinOrbitFunctional.cpp
    P.soc_kfac = (2.0 * kPi) / a;
    P.period_a = m->period_anion;
    P.period_c = m->period_cation;
    P.xi_a = m->radial_extent_anion;
    P.xi_c = m->radial_extent_cation;
    P.mu = m->mu_soc;
    P.alpha_mu = m->alpha_soc * m->mu_soc;
    P.bohr = kBohr;
    // Hamiltonian.cpp:158-161
    P.grad_pref = 2.0 * diag_factor * (P.tpa * P.tpa);
    P.frac_to_phys = 1.0 / P.tpa;

    plan_apply_config(p);
    hipError_t e = p->G.ensure(sizeof(int32_t) * 3 * (size_t)N);
    if (e == hipSuccess) e = hipMemcpy(p->G.p, G, sizeof(int32_t) * 3 * (size_t)N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = p->status.ensure(2 * sizeof(unsigned long long));
    if (e == hipSuccess) {
        const unsigned long long none[2] = {0ull, ~0ull};
        e = hipMemcpy(p->status.p, none, sizeof(none), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        delete p;
        return fail(UEPM_ERR_DEVICE, std::string("uepm_plan_create: ") + hipGetErrorString(e));
    }
    *out = p;
    return UEPM_OK;
}

int uepm_plan_destroy(uepm_plan* plan) {
    if (plan && plan->pins > 0)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_destroy: the plan is used by a batch graph (uepm_batch_destroy first)");
    delete plan;
    return UEPM_OK;
}

int uepm_plan_dim(const uepm_plan* plan) { return plan ? plan->dim : -1; }

int uepm_plan_set_timing(uepm_plan* plan, int enable) {
    if (!plan) return fail(UEPM_ERR_INVALID_ARG, "null plan");
    plan->timing = enable;
    return UEPM_OK;
}

int uepm_plan_last_timing(const uepm_plan* plan, int stage, double* ms, int64_t* launches, const char** name) {
    if (!plan) return fail(UEPM_ERR_INVALID_ARG, "null plan");
    if (stage < 0 || stage >= kNumStages) return fail(UEPM_ERR_OUT_OF_RANGE, "no such stage");
    if (ms) *ms = plan->stage_ms[stage];
    if (launches) *launches = plan->stage_launches[stage];
    if (name) *name = plan->stage_name[stage][0] ? plan->stage_name[stage] : kStageNames[stage];
    return UEPM_OK;
}

int uepm_plan_get_config(const uepm_plan* p, uepm_config* cfg) {
    if (!p || !cfg) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_get_config: null argument");
    *cfg = p->cfg;
    return UEPM_OK;
}

int uepm_plan_set_config(uepm_plan* p, const uepm_config* cfg) {
    if (!p || !cfg) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_set_config: null argument");
    if (p->pins > 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_set_config: the plan is used by a batch graph");
    if (cfg->tridiag < 0 || cfg->tridiag > 2)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_config.tridiag must be 0, 1 or 2");
    if (cfg->tridiag == 2 && !((p->flags & UEPM_MIXED) && p->dim <= uepm::tridiag_reg_max_n()))
        return fail(UEPM_ERR_INVALID_ARG, "the register-resident reduction needs UEPM_MIXED and dim <= " +
                                              std::to_string(uepm::tridiag_reg_max_n()));
    if (cfg->tri_block != 0 && cfg->tri_block != 256 && cfg->tri_block != 512)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_config.tri_block must be 0, 256 or 512");
    if (cfg->bt_bands != 0 && cfg->bt_bands != 4 && cfg->bt_bands != 8 && cfg->bt_bands != 16)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_config.bt_bands must be 0, 4, 8 or 16");
    if (cfg->max_chunk < 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_config.max_chunk must be >= 0");
    if (cfg->lanes < 0 || cfg->lanes > 2) return fail(UEPM_ERR_INVALID_ARG, "uepm_config.lanes must be 0, 1 or 2");
    p->cfg = *cfg;
    plan_apply_config(p);
    return UEPM_OK;
}

int uepm_plan_workspace_bytes(const uepm_plan* p, int64_t nk, uint64_t* bytes) {
    if (!p || !bytes) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_workspace_bytes: null argument");
    if (nk < 0) return fail(UEPM_ERR_INVALID_ARG, "nk < 0");
    *bytes = nk == 0 ? 0 : (uint64_t)plan_workspace_bytes(p, plan_chunk(p, nk), plan_lanes(p, nk));
    return UEPM_OK;
}

int uepm_plan_reserve(uepm_plan* p, int64_t nk) {
    if (!p) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_reserve: null plan");
    if (nk < 0) return fail(UEPM_ERR_INVALID_ARG, "nk < 0");
    if (nk == 0) return UEPM_OK;
    HIP_OK(hipSetDevice(p->ctx->device));
    return plan_reserve(p, plan_chunk(p, nk), plan_lanes(p, nk));
}

int uepm_plan_tune(uepm_plan* p, const double* dk, int64_t nk, void* stream_, uepm_config* chosen) {
    if (!p || !dk || nk <= 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_tune: null plan, no k-points");
    if (p->pins > 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_tune: the plan is used by a batch graph");
    HIP_OK(hipSetDevice(p->ctx->device));
    hipStream_t s = (hipStream_t)stream_;
    // candidates: reduction variant x workgroup size, back-transform band groups
    std::vector<uepm_config> cand;
    const bool mixed = p->flags & UEPM_MIXED;
    const bool reg_ok = mixed && p->dim <= uepm::tridiag_reg_max_n();
    // back-transform: the blocked MFMA kernel (0: automatic, which is the blocked
    // kernel in the precision form and above dim 192 in FP64) or 8 / 16 bands per
    // workgroup of the reflector-by-reflector one
    std::vector<int> bts = {8, 16};
    if (mixed || (p->dim > 192 && uepm::apply_q_wy_f64_fits(p->dim, p->ctx->lds_limit))) bts.insert(bts.begin(), 0);
    for (int bt : bts) {
        if (reg_ok) cand.push_back({2, 0, bt, p->cfg.max_chunk, p->cfg.lanes});
        for (int tb : {256, 512}) {
            cand.push_back({1, tb, bt, p->cfg.max_chunk, p->cfg.lanes});
            // above the register grids: blocked leading steps, register-grid tail
            if (mixed && !reg_ok) cand.push_back({0, tb, bt, p->cfg.max_chunk, p->cfg.lanes});
        }
    }
    const uepm_config saved = p->cfg;
    const int saved_timing = p->timing;
    p->timing = 0;
    DeviceBuffer ev, vec, gr;
    int rc = UEPM_OK;
    if (ev.ensure(sizeof(double) * (size_t)nk * p->nb) != hipSuccess)
        rc = fail(UEPM_ERR_NO_MEMORY, "uepm_plan_tune: out of device memory");
    const bool want_vec = p->flags & UEPM_VECTORS;
    if (rc == UEPM_OK && want_vec &&
        vec.ensure(sizeof(double2) * (size_t)nk * p->nb * p->dim) != hipSuccess)
        rc = fail(UEPM_ERR_NO_MEMORY, "uepm_plan_tune: out of device memory");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == UEPM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(UEPM_ERR_DEVICE, "uepm_plan_tune: hipEventCreate failed");
    float best = 3.4e38f;
    uepm_config best_cfg = saved;
    for (size_t c = 0; rc == UEPM_OK && c < cand.size(); ++c) {
        p->cfg = cand[c];
        plan_apply_config(p);
        double* dv = want_vec ? vec.as<double>() : nullptr;
        rc = uepm_solve_device(p, dk, nk, ev.as<double>(), dv, nullptr, s);  // warm-up (and workspace)
        if (rc != UEPM_OK) break;
        float ms = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(UEPM_ERR_DEVICE, "hipEventRecord"); break; }
        for (int r = 0; r < 2 && rc == UEPM_OK; ++r) rc = uepm_solve_device(p, dk, nk, ev.as<double>(), dv, nullptr, s);
        if (rc != UEPM_OK) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(UEPM_ERR_DEVICE, "uepm_plan_tune: event timing failed");
            break;
        }
        if (ms < best) {
            best = ms;
            best_cfg = cand[c];
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    p->timing = saved_timing;
    p->cfg = rc == UEPM_OK ? best_cfg : saved;
    plan_apply_config(p);
    if (rc == UEPM_OK && chosen) *chosen = p->cfg;
    return rc;
}

int uepm_assemble_device(uepm_plan* p, const double* dk, int64_t nk, double* dH, void* stream) {
    if (!p || (nk > 0 && (!dk || !dH))) return fail(UEPM_ERR_INVALID_ARG, "uepm_assemble_device: null argument");
    if (nk < 0) return fail(UEPM_ERR_INVALID_ARG, "nk < 0");
    if (nk == 0) return UEPM_OK;
    HIP_OK(hipSetDevice(p->ctx->device));
    const int64_t chunk = 65535;
    for (int64_t c0 = 0; c0 < nk; c0 += chunk) {
        const int64_t c = std::min(chunk, nk - c0);
        HIP_OK(uepm::launch_assemble(p->P, p->G.as<int32_t>(), dk + 3 * c0, c,
                                     reinterpret_cast<double2*>(dH) + (size_t)c0 * p->dim * p->dim, p->dim,
                                     (hipStream_t)stream));
    }
    return UEPM_OK;
}

int uepm_solve_device(uepm_plan* p, const double* dk, int64_t nk, double* devals, double* devecs, double* dgrads,
                      void* stream_) {
    const int v = solve_args_ok(p, nk, dk, devals, devecs, dgrads, "uepm_solve_device");
    if (v != UEPM_OK || nk == 0) return v;
    HIP_OK(hipSetDevice(p->ctx->device));
    hipStream_t s = (hipStream_t)stream_;
    const int64_t chunk = plan_chunk(p, nk);
    const int lanes = stream_capturing(s) ? 1 : plan_lanes(p, nk);  // a stream under graph capture keeps one lane
    {
        const int st = plan_reserve(p, chunk, lanes);
        if (st != UEPM_OK) return st;
    }
    LaneView views[2];
    for (int l = 0; l < lanes; ++l) views[l] = lane_view(p->lane[l]);
    return solve_chunks(p, dk, nk, devals, devecs, dgrads, s, views, lanes, chunk, false);
}

int uepm_device_alloc(uepm_ctx* ctx, uint64_t bytes, void** ptr) {
    if (!ctx || !ptr) return fail(UEPM_ERR_INVALID_ARG, "uepm_device_alloc: null argument");
    HIP_OK(hipSetDevice(ctx->device));
    *ptr = nullptr;
    if (bytes == 0) return UEPM_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory) return fail(UEPM_ERR_NO_MEMORY, "uepm_device_alloc: out of device memory");
    HIP_OK(e);
    return UEPM_OK;
}

int uepm_device_free(uepm_ctx* ctx, void* ptr) {
    if (!ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_device_free: null context");
    if (!ptr) return UEPM_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipFree(ptr));
    return UEPM_OK;
}

int uepm_memcpy_htod(uepm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(UEPM_ERR_INVALID_ARG, "uepm_memcpy_htod: null argument");
    if (!bytes) return UEPM_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return UEPM_OK;
}

int uepm_memcpy_dtoh(uepm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(UEPM_ERR_INVALID_ARG, "uepm_memcpy_dtoh: null argument");
    if (!bytes) return UEPM_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return UEPM_OK;
}

int uepm_memset_device(uepm_ctx* ctx, void* dst, int value, uint64_t bytes) {
    if (!ctx || (bytes && !dst)) return fail(UEPM_ERR_INVALID_ARG, "uepm_memset_device: null argument");
    if (!bytes) return UEPM_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipMemset(dst, value, bytes));
    return UEPM_OK;
}

int uepm_memcpy_dtod(uepm_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream) {
    if (!ctx || (bytes && (!dst || !src))) return fail(UEPM_ERR_INVALID_ARG, "uepm_memcpy_dtod: null argument");
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_memcpy_dtod: pointers must be 16-byte aligned");
    if (!bytes) return UEPM_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(uepm::launch_copy_bytes(dst, src, bytes, ctx->prop.multiProcessorCount, static_cast<hipStream_t>(stream)));
    return UEPM_OK;
}

int uepm_stream_create(uepm_ctx* ctx, void** stream) {
    if (!ctx || !stream) return fail(UEPM_ERR_INVALID_ARG, "uepm_stream_create: null argument");
    HIP_OK(hipSetDevice(ctx->device));
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return UEPM_OK;
}

int uepm_stream_destroy(uepm_ctx* ctx, void* stream) {
    if (!ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_stream_destroy: null context");
    if (!stream) return UEPM_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipStreamDestroy((hipStream_t)stream));
    return UEPM_OK;
}

int uepm_stream_synchronize(uepm_ctx* ctx, void* stream) {
    if (!ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_stream_synchronize: null context");
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream));
    return UEPM_OK;
}

int uepm_device_synchronize(uepm_ctx* ctx) {
    if (!ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_device_synchronize: null context");
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipDeviceSynchronize());
    return UEPM_OK;
}

int uepm_solve(uepm_plan* p, const double* k, int64_t nk, double* evals, double* evecs, double* grads) {
    if (!p) return fail(UEPM_ERR_INVALID_ARG, "uepm_solve: null plan");
    if (nk < 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_solve: nk < 0");
    if (nk == 0) return UEPM_OK;
    if (!k || !evals) return fail(UEPM_ERR_INVALID_ARG, "uepm_solve: null k or eigenvalue buffer");
    for (int64_t i = 0; i < 3 * nk; ++i)
        if (!std::isfinite(k[i]))
            return fail(UEPM_ERR_INVALID_ARG, "uepm_solve: k-point " + std::to_string(i / 3) + " is not finite");
    HIP_OK(hipSetDevice(p->ctx->device));
    const int n = p->dim, nb = p->nb;
    DeviceBuffer dk, dev, dvec, dgr;
    HIP_OK(dk.ensure(sizeof(double) * 3 * nk));
    HIP_OK(dev.ensure(sizeof(double) * nb * nk));
    if (evecs) HIP_OK(dvec.ensure(sizeof(double2) * (size_t)n * nb * nk));
    if (grads) HIP_OK(dgr.ensure(sizeof(double) * 3 * nb * nk));
    HIP_OK(hipMemcpy(dk.p, k, sizeof(double) * 3 * nk, hipMemcpyHostToDevice));
    int r = uepm_solve_device(p, dk.as<double>(), nk, dev.as<double>(), evecs ? dvec.as<double>() : nullptr,
                              grads ? dgr.as<double>() : nullptr, nullptr);
    if (r != UEPM_OK) return r;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(evals, dev.p, sizeof(double) * nb * nk, hipMemcpyDeviceToHost));
    if (evecs) HIP_OK(hipMemcpy(evecs, dvec.p, sizeof(double2) * (size_t)n * nb * nk, hipMemcpyDeviceToHost));
    if (grads) HIP_OK(hipMemcpy(grads, dgr.p, sizeof(double) * 3 * nb * nk, hipMemcpyDeviceToHost));
    return uepm_plan_status(p, nullptr, nullptr, nullptr);
}

int uepm_plan_status(uepm_plan* p, void* stream, int64_t* n_failed, int64_t* first_failed) {
    if (!p) return fail(UEPM_ERR_INVALID_ARG, "uepm_plan_status: null plan");
    HIP_OK(hipSetDevice(p->ctx->device));
    unsigned long long h[2] = {0ull, ~0ull};
    HIP_OK(hipMemcpyAsync(h, p->status.p, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream));
    if (n_failed) *n_failed = (int64_t)h[0];
    if (first_failed) *first_failed = h[0] ? (int64_t)h[1] : -1;
    if (h[0])
        return fail(UEPM_ERR_NOT_CONVERGED, "Eigenvalue decomposition failed for " + std::to_string(h[0]) +
                                                " k-point(s), the first at index " + std::to_string(h[1]));
    return UEPM_OK;
}

}  // extern "C"
