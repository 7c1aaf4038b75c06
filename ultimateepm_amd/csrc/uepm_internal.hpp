// uepm_internal.hpp — definitions shared by the host-side sources of the C ABI
// (uepm_abi.cpp: plans and solves; runtime.cpp: tuning, batch graphs, serving).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/uepm.h"
#include "solver.hpp"

namespace uepm_host {

// per-thread message behind uepm_last_error()
std::string& last_error();

inline int fail(int code, const std::string& msg) {
    last_error() = msg;
    return code;
}

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return ::uepm_host::fail(UEPM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DeviceBuffer() { release(); }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

enum Stage { kAssemble = 0, kTridiag, kTriEig, kBack, kRefine, kNumStages };
// Extra Ritz vectors carried by the precision form beyond the nb requested,
// so that a degenerate pair straddling band nb is refined as a whole.
constexpr int kMixedExtra = 4;
// Bisection bracket (relative to ||T||_inf) of the precision form's eigen-stage.
constexpr double kMixedBisectTol = 1e-10;
// Inverse-iteration sweeps: the FP64 form's vectors are its output; the
// precision form's only seed the FP64 Rayleigh-Ritz step, whose values are
// second order in their error (two sweeps leave ~1e-5 in the vectors; one
// sweep leaves some k-points with two vectors of a close pair nearly parallel,
// and the Ritz step then misses a band by tens of meV)
constexpr int kFp64InvSweeps = 3;
constexpr int kMixedInvSweeps = 2;

// Workspace of one chunk in flight. A solve over two or more chunks runs two
// lanes: even chunks on the caller's stream, odd chunks on the plan's side
// stream, each with its own workspace.
struct Lane {
    DeviceBuffer H, d, e, tau, Z, W, X, S;
    DeviceBuffer Vf, tauf, S32, Ev, X0;  // precision form
    DeviceBuffer kstat;                  // per k-point failure flag of the chunk in flight
};

}  // namespace uepm_host

struct uepm_ctx {
    int device = 0;
    hipDeviceProp_t prop{};
    int lds_limit = 65536;  // dynamic LDS bytes one workgroup may use
};

struct uepm_plan {
    uepm_ctx* ctx = nullptr;
    uepm::DevParams P{};
    unsigned flags = 0;
    int N = 0, dim = 0, nb = 0;
    uepm_host::DeviceBuffer G;
    uepm_host::Lane lane[2];
    hipStream_t side = nullptr;  // the second lane's stream (created on first use)
    hipEvent_t fork = nullptr, join = nullptr;
    // the last solve's failure count and first failed k-point (device, two
    // words, cleared by each solve)
    uepm_host::DeviceBuffer status;
    // batch graphs holding raw pointers into lane 0's workspace: while > 0,
    // lane 0 is not reallocated and the plan is not destroyed
    int pins = 0;
    // completion of each lane's last chunk: the next solve's streams wait on
    // them, so solves of one plan on different streams do not overlap
    hipEvent_t lane_done[2] = {nullptr, nullptr};
    uepm::TriPlan tp{}, tp32{};
    uepm_config cfg{};    // requested configuration (0 fields = automatic)
    bool tri_reg = false; // register-resident complex64 reduction (precision form, dim <= 288)
    int bt_bands = 16;    // bands per workgroup of the reflector-by-reflector back-transform
    bool bt_wy = false;   // the blocked MFMA back-transform instead (complex64 or FP64 by form)
    int timing = 0;
    std::vector<hipEvent_t> ev;  // pairs per stage launch
    const char* stage_name[uepm_host::kNumStages] = {"assemble_kernel", "", "", "", ""};  // kernels of the last solve
    double stage_ms[uepm_host::kNumStages] = {0, 0, 0, 0, 0};
    int64_t stage_launches[uepm_host::kNumStages] = {0, 0, 0, 0, 0};
    ~uepm_plan() {
        for (auto e : ev) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (auto e : lane_done)
            if (e) (void)hipEventDestroy(e);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace uepm_host {
// One lane's workspace as raw device pointers: the plan's own buffers
// (lane_view), or a region carved out of a caller-independent scratch array
// (plan_carve).
struct LaneView {
    double2* H = nullptr;
    double* d = nullptr;
    double* e = nullptr;
    double2* tau = nullptr;
    float2* Vf = nullptr;
    float2* tauf = nullptr;
    float2* S32 = nullptr;
    double* Ev = nullptr;
    double2* X0 = nullptr;
    double2* S = nullptr;
    double* Z = nullptr;
    double* W = nullptr;
    double2* X = nullptr;
    int* kstat = nullptr;
};
inline LaneView lane_view(const Lane& w) {
    LaneView v;
    v.H = w.H.as<double2>();
    v.d = w.d.as<double>();
    v.e = w.e.as<double>();
    v.tau = w.tau.as<double2>();
    v.Vf = w.Vf.as<float2>();
    v.tauf = w.tauf.as<float2>();
    v.S32 = w.S32.as<float2>();
    v.Ev = w.Ev.as<double>();
    v.X0 = w.X0.as<double2>();
    v.S = w.S.as<double2>();
    v.Z = w.Z.as<double>();
    v.W = w.W.as<double>();
    v.X = w.X.as<double2>();
    v.kstat = w.kstat.as<int>();
    return v;
}
// Bytes one lane of `chunk` k-points takes when its buffers are laid out
// back to back at 256-byte boundaries from `base`; with base != nullptr the
// view's pointers are set to that layout.
size_t plan_carve(const uepm_plan* p, int64_t chunk, char* base, LaneView* view);
// hipStreamIsCapturing reports a capture in progress (or fails)
bool stream_capturing(hipStream_t s);
// argument checks shared by the device-buffer solve entry points
int solve_args_ok(const uepm_plan* p, int64_t nk, const double* dk, const double* devals, const double* devecs,
                  const double* dgrads, const char* who);
// The solve of nk k-points in chunks of `chunk` over the lane workspaces
// views[0..lanes): status reset, ordering after the plan's previous solve,
// the stages of every chunk (alternate chunks on the plan's side stream when
// lanes == 2), stage timing.  no_alloc: one lane, and no event, stream or
// device memory is created (stage timing is then not recorded).
int solve_chunks(uepm_plan* p, const double* dk, int64_t nk, double* devals, double* devecs, double* dgrads,
                 hipStream_t s, const LaneView* views, int lanes, int64_t chunk, bool no_alloc);
// Eigenvalues of nk k-points on stream s with the chunk workspace carved out of
// the device region [base, base + bytes): one chunk in flight, chunks of the
// plan's size or smaller so that the workspace fits.  Allocates no device memory.
int solve_region(uepm_plan* p, const double* dk, int64_t nk, double* devals, hipStream_t s, char* base,
                 size_t bytes);
// Looks up the no-allocation scratch array's address on `device` (current),
// once per device; called by uepm_ctx_create.
int noalloc_setup(int device);
// k-points per launch for a solve of nk points under the plan's configuration
int64_t plan_chunk(const uepm_plan* p, int64_t nk);
// lanes a solve of nk points runs (2 when it spans two or more chunks, unless
// the configuration asks for one)
int plan_lanes(const uepm_plan* p, int64_t nk);
// workspace of `lanes` chunks in flight (bytes) and its allocation
size_t plan_workspace_bytes(const uepm_plan* p, int64_t chunk, int lanes = 1);
int plan_reserve(uepm_plan* p, int64_t chunk, int lanes = 1);
// applies p->cfg to the derived launch parameters (validated by the caller)
void plan_apply_config(uepm_plan* p);
}  // namespace uepm_host
