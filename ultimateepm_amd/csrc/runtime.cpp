// runtime.cpp — batch graphs and the serving engine over solve plans.
//
// Batch (uepm_batch_*): the solves of many plans (materials, cutoffs, SOC on
// or off, so uneven matrix sizes) are captured once into a hipGraph and
// replayed with a single launch.  Plans are spread over concurrent branches
// (heaviest first); every plan's workspace is reserved before capture, so the
// graph allocates nothing.
//
// Server (uepm_server_*): continuous batching of band-structure requests.  A
// scheduler thread admits queued requests in arrival order whenever an
// in-flight slot is free; each admitted request runs on its slot's stream
// (copy in, solve, copy out, completion event), so requests that arrive while
// others run overlap on the device.  Each slot owns one workspace arena of
// budget / max_inflight bytes, allocated on the slot's first request and kept:
// a request's chunk workspace is carved out of it, chunks sized to fit.  Plans
// are pooled per (material, basis, flags, bands) and hold no workspace, so a
// request never allocates or frees device memory while others run (either
// would synchronise the device); idle plans are destroyed, oldest first, only
// when the pool exceeds kMaxPooledPlans.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "uepm_internal.hpp"

using namespace uepm_host;

// ============================================================== batch graphs
struct uepm_batch {
    uepm_ctx* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    // plans whose lane-0 workspace the graph's kernels address: pinned for the
    // graph's life (their buffers are not reallocated, the plans not destroyed)
    std::vector<uepm_plan*> pinned;
    ~uepm_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto s : streams) (void)hipStreamDestroy(s);
        for (auto* p : pinned) --p->pins;
    }
};

extern "C" {

int uepm_batch_create(uepm_ctx* ctx, const uepm_batch_item* items, int nitems, int branches, uepm_batch** out) {
    if (!ctx || !out || (nitems > 0 && !items) || nitems < 0)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_batch_create: null argument");
    if (branches < 1) branches = 1;
    for (int i = 0; i < nitems; ++i) {
        const uepm_batch_item& it = items[i];
        if (!it.plan || it.nk < 0 || (it.nk > 0 && (!it.d_k || !it.d_evals)))
            return fail(UEPM_ERR_INVALID_ARG, "uepm_batch_create: item " + std::to_string(i) + " is incomplete");
        if (it.plan->ctx != ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_batch_create: plan of another context");
        if (it.d_evecs && !(it.plan->flags & UEPM_VECTORS))
            return fail(UEPM_ERR_INVALID_ARG, "uepm_batch_create: eigenvectors from a plan without UEPM_VECTORS");
        if (it.d_grads && !(it.plan->flags & UEPM_GRADIENTS))
            return fail(UEPM_ERR_INVALID_ARG, "uepm_batch_create: gradients from a plan without UEPM_GRADIENTS");
    }
    HIP_OK(hipSetDevice(ctx->device));
    // jobs: the items of one plan run in order on one branch (a plan's workspace is not shared)
    std::map<uepm_plan*, std::vector<int>> jobs;
    std::map<uepm_plan*, int64_t> maxnk;
    for (int i = 0; i < nitems; ++i) {
        jobs[items[i].plan].push_back(i);
        maxnk[items[i].plan] = std::max(maxnk[items[i].plan], items[i].nk);
    }
    std::vector<std::pair<double, uepm_plan*>> order;
    for (auto& j : jobs) {
        double cost = 0.0;
        for (int i : j.second) cost += (double)items[i].nk * j.first->dim * j.first->dim * j.first->dim;
        order.push_back({cost, j.first});
        // solves under capture run one lane
        const int st = plan_reserve(j.first, plan_chunk(j.first, maxnk[j.first]), 1);
        if (st != UEPM_OK) return st;
    }
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    branches = std::max(1, std::min<int>(branches, (int)order.size()));
    std::vector<std::vector<uepm_plan*>> lanes(branches);
    std::vector<double> load(branches, 0.0);
    for (auto& o : order) {  // longest job to the least loaded branch
        const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lanes[b].push_back(o.second);
        load[b] += o.first;
    }
    auto* B = new uepm_batch;
    B->ctx = ctx;
    auto cleanup = [&](int code) {
        delete B;
        return code;
    };
    for (int b = 0; b <= branches; ++b) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(UEPM_ERR_DEVICE, "uepm_batch_create: hipStreamCreate failed"));
        B->streams.push_back(s);
    }
    for (int e = 0; e <= branches; ++e) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(UEPM_ERR_DEVICE, "uepm_batch_create: hipEventCreate failed"));
        B->events.push_back(ev);
    }
    hipStream_t origin = B->streams[0];
    if (hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed) != hipSuccess)
        return cleanup(fail(UEPM_ERR_DEVICE, "uepm_batch_create: hipStreamBeginCapture failed"));
    int st = UEPM_OK;
    if (hipEventRecord(B->events[0], origin) != hipSuccess) st = fail(UEPM_ERR_DEVICE, "hipEventRecord (fork)");
    for (int b = 0; b < branches && st == UEPM_OK; ++b) {
        hipStream_t s = B->streams[b + 1];
        if (hipStreamWaitEvent(s, B->events[0], 0) != hipSuccess) {
            st = fail(UEPM_ERR_DEVICE, "hipStreamWaitEvent (fork)");
            break;
        }
        for (uepm_plan* p : lanes[b]) {
            const int saved = p->timing;
            p->timing = 0;
            for (int i : jobs[p]) {
                const uepm_batch_item& it = items[i];
                st = uepm_solve_device(p, it.d_k, it.nk, it.d_evals, it.d_evecs, it.d_grads, s);
                if (st != UEPM_OK) break;
            }
            p->timing = saved;
            if (st != UEPM_OK) break;
        }
        if (st == UEPM_OK && (hipEventRecord(B->events[b + 1], s) != hipSuccess ||
                              hipStreamWaitEvent(origin, B->events[b + 1], 0) != hipSuccess))
            st = fail(UEPM_ERR_DEVICE, "hipEventRecord / hipStreamWaitEvent (join)");
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(origin, &g);
    if (st != UEPM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return cleanup(st);
    }
    if (ec != hipSuccess || !g)
        return cleanup(fail(UEPM_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec)));
    B->graph = g;
    const hipError_t ei = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess)
        return cleanup(fail(UEPM_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei)));
    for (auto& j : jobs) {
        ++j.first->pins;
        B->pinned.push_back(j.first);
    }
    *out = B;
    return UEPM_OK;
}

int uepm_batch_launch(uepm_batch* b, void* stream) {
    if (!b) return fail(UEPM_ERR_INVALID_ARG, "uepm_batch_launch: null batch");
    HIP_OK(hipSetDevice(b->ctx->device));
    HIP_OK(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return UEPM_OK;
}

int uepm_batch_destroy(uepm_batch* b) {
    delete b;
    return UEPM_OK;
}

}  // extern "C"

// ============================================================ serving engine
namespace {

using Clock = std::chrono::steady_clock;

struct PlanKey {
    std::string bytes;  // material struct, shells, flags, nb
    bool operator==(const PlanKey& o) const { return bytes == o.bytes; }
};
struct PlanKeyHash {
    size_t operator()(const PlanKey& k) const { return std::hash<std::string>()(k.bytes); }
};

struct PooledPlan {
    uepm_plan* plan = nullptr;
    PlanKey key;
    Clock::time_point last_used;
};
constexpr size_t kMaxPooledPlans = 512;  // a plan without workspace holds a few KB of device memory

struct Request {
    uint64_t id = 0;
    uepm_material mat{};
    int shells = 0;
    unsigned flags = 0;
    int nb = 0;
    std::vector<double> k;
    int64_t nk = 0;
    Clock::time_point t_submit, t_admit, t_done;
    int status = -1;  // -1 queued/running, else a UEPM status
    std::string error;
    std::vector<double> evals;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uepm_host::DeviceBuffer arena;  // the slot's chunk workspace (budget / max_inflight bytes)
    uepm_host::DeviceBuffer dk, dev;
    double* hk = nullptr;   // pinned staging
    double* hev = nullptr;
    unsigned long long* hstat = nullptr;  // the plan's status words after the solve (pinned)
    size_t hk_cap = 0, hev_cap = 0;
    std::shared_ptr<Request> req;
    PooledPlan* plan = nullptr;
};

}  // namespace

struct uepm_server {
    uepm_ctx* ctx = nullptr;
    uint64_t budget = 0;
    int max_inflight = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // owned by the worker thread
    std::vector<Slot> slots;
    std::list<PooledPlan> plans;  // every plan, idle or busy
    std::unordered_map<PooledPlan*, bool> busy;
    size_t used = 0;
    uepm_server_stats stats{};
};

namespace {

int grow_pinned(double*& p, size_t& cap, size_t need) {
    if (need <= cap) return UEPM_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    HIP_OK(hipHostMalloc((void**)&p, need * sizeof(double), hipHostMallocDefault));
    cap = need;
    return UEPM_OK;
}

// Workspace bytes of one slot's arena: an equal share of the budget.
uint64_t slot_share(const uepm_server* S) { return S->budget / (uint64_t)std::max(1, S->max_inflight); }

// An idle plan for the request's key (created on demand).  Returns nullptr with
// *status set when the request can never run: its plan fails, or one k-point's
// workspace exceeds a slot's share of the budget (UEPM_ERR_NO_MEMORY).
PooledPlan* acquire_plan(uepm_server* S, const Request& r, const PlanKey& key, int* status) {
    *status = UEPM_OK;
    for (auto& q : S->plans)
        if (!S->busy[&q] && q.key == key) return &q;
    int count = 0;
    if (uepm_basis_vectors(r.shells, nullptr, 0, &count) != UEPM_OK) {
        *status = UEPM_ERR_RUNTIME;
        return nullptr;
    }
    std::vector<int32_t> G(3 * (size_t)count);
    uepm_basis_vectors(r.shells, G.data(), count, &count);
    uepm_plan* plan = nullptr;
    const int st = uepm_plan_create(S->ctx, &r.mat, G.data(), count, r.flags, r.nb, &plan);
    if (st != UEPM_OK) {
        *status = st;
        return nullptr;
    }
    const size_t per_k = std::max<size_t>(1, plan_carve(plan, 1, nullptr, nullptr));
    const uint64_t share = slot_share(S);
    if (per_k > share) {
        uepm_plan_destroy(plan);
        *status = fail(UEPM_ERR_NO_MEMORY, "server: one k-point of this request needs " + std::to_string(per_k) +
                                               " bytes of workspace, more than a slot's share (" +
                                               std::to_string(share) + " bytes) of the budget");
        return nullptr;
    }
    uepm_config cfg{};
    cfg.max_chunk = (int32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, share / per_k));
    cfg.lanes = 1;  // the slots are the concurrency; one request, one chunk in flight
    uepm_plan_set_config(plan, &cfg);
    while (S->plans.size() >= kMaxPooledPlans) {
        PooledPlan* victim = nullptr;
        for (auto& q : S->plans)
            if (!S->busy[&q] && (!victim || q.last_used < victim->last_used)) victim = &q;
        if (!victim) break;
        uepm_plan_destroy(victim->plan);
        S->busy.erase(victim);
        S->plans.remove_if([&](const PooledPlan& q) { return &q == victim; });
        ++S->stats.evictions;
    }
    S->plans.push_back(PooledPlan{plan, key, Clock::now()});
    PooledPlan* pp = &S->plans.back();
    S->busy[pp] = false;
    return pp;
}

void finish(uepm_server* S, Slot& sl, int status, const std::string& err, bool solved = false) {
    auto r = sl.req;
    r->t_done = Clock::now();
    std::string msg = err;
    if (solved && sl.hstat[0] != 0) {  // the eigenvalues of the other k-points are still returned
        status = UEPM_ERR_NOT_CONVERGED;
        msg = "Eigenvalue decomposition failed for " + std::to_string(sl.hstat[0]) +
              " k-point(s), the first at index " + std::to_string(sl.hstat[1]);
    }
    if (status == UEPM_OK || solved) r->evals.assign(sl.hev, sl.hev + (size_t)r->nk * r->nb);
    {
        std::lock_guard<std::mutex> g(S->mu);
        r->status = status;
        r->error = msg;
        ++S->stats.completed;
        S->stats.kpoints += (uint64_t)r->nk;
    }
    if (sl.plan) {
        S->busy[sl.plan] = false;
        sl.plan->last_used = Clock::now();
    }
    sl.plan = nullptr;
    sl.req.reset();
    S->cv_done.notify_all();
}

void worker_main(uepm_server* S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        bool progress = false;
        int inflight = 0;
        for (auto& sl : S->slots) {
            if (!sl.req) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                finish(S, sl, UEPM_OK, "", true);
                progress = true;
            } else if (q != hipErrorNotReady) {
                finish(S, sl, UEPM_ERR_DEVICE, std::string("request failed on the device: ") + hipGetErrorString(q));
                progress = true;
            } else {
                ++inflight;
            }
        }
        for (;;) {
            Slot* free_slot = nullptr;
            for (auto& sl : S->slots)
                if (!sl.req) {
                    free_slot = &sl;
                    break;
                }
            std::shared_ptr<Request> r;
            {
                std::lock_guard<std::mutex> g(S->mu);
                if (!free_slot || S->queue.empty()) break;
                r = S->queue.front();
            }
            PlanKey key;
            key.bytes.assign(reinterpret_cast<const char*>(&r->mat), sizeof(uepm_material));
            key.bytes.append(reinterpret_cast<const char*>(&r->shells), sizeof(int));
            key.bytes.append(reinterpret_cast<const char*>(&r->flags), sizeof(unsigned));
            key.bytes.append(reinterpret_cast<const char*>(&r->nb), sizeof(int));
            int st = UEPM_OK;
            PooledPlan* pp = acquire_plan(S, *r, key, &st);
            {
                std::lock_guard<std::mutex> g(S->mu);
                S->queue.pop_front();
            }
            progress = true;
            Slot& sl = *free_slot;
            sl.req = r;
            r->t_admit = Clock::now();
            if (!pp) {
                finish(S, sl, st, last_error());
                continue;
            }
            S->busy[pp] = true;
            sl.plan = pp;
            // request buffers: sized for 4096 k-points and 32 bands at least, so that they
            // rarely grow (a device reallocation would synchronise the device)
            const size_t kcount = 3 * (size_t)r->nk, ecount = (size_t)r->nk * r->nb;
            const size_t kcap = std::max<size_t>(kcount, 3 * 4096), ecap = std::max<size_t>(ecount, 4096 * 32);
            st = grow_pinned(sl.hk, sl.hk_cap, kcap);
            if (st == UEPM_OK) st = grow_pinned(sl.hev, sl.hev_cap, ecap);
            if (st == UEPM_OK && (sl.dk.ensure(kcap * sizeof(double)) != hipSuccess ||
                                  sl.dev.ensure(ecap * sizeof(double)) != hipSuccess))
                st = fail(UEPM_ERR_NO_MEMORY, "server: request buffers");
            if (st == UEPM_OK && !sl.arena.p) {
                if (sl.arena.ensure(slot_share(S)) != hipSuccess)
                    st = fail(UEPM_ERR_NO_MEMORY, "server: the slot's workspace arena (" +
                                                      std::to_string(slot_share(S)) + " bytes)");
                else
                    S->used += sl.arena.bytes;
            }
            if (st == UEPM_OK) {
                std::memcpy(sl.hk, r->k.data(), kcount * sizeof(double));
                if (hipMemcpyAsync(sl.dk.p, sl.hk, kcount * sizeof(double), hipMemcpyHostToDevice, sl.stream) !=
                    hipSuccess)
                    st = fail(UEPM_ERR_DEVICE, "server: copy in");
            }
            if (st == UEPM_OK)
                st = solve_region(pp->plan, sl.dk.as<double>(), r->nk, sl.dev.as<double>(), sl.stream,
                                  sl.arena.as<char>(), sl.arena.bytes);
            if (st == UEPM_OK &&
                (hipMemcpyAsync(sl.hev, sl.dev.p, ecount * sizeof(double), hipMemcpyDeviceToHost, sl.stream) !=
                     hipSuccess ||
                 hipMemcpyAsync(sl.hstat, pp->plan->status.p, 2 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
                 hipEventRecord(sl.done, sl.stream) != hipSuccess))
                st = fail(UEPM_ERR_DEVICE, "server: copy out");
            if (st != UEPM_OK) {
                (void)hipStreamSynchronize(sl.stream);
                finish(S, sl, st, last_error());
            } else {
                ++inflight;
            }
        }
        std::unique_lock<std::mutex> lk(S->mu);
        S->stats.inflight = inflight;
        S->stats.queued = (int)S->queue.size();
        S->stats.used_bytes = S->used;
        if (S->stop && S->queue.empty() && inflight == 0) break;
        if (!progress) S->cv_work.wait_for(lk, std::chrono::microseconds(inflight ? 50 : 2000));
    }
}

}  // namespace

extern "C" {

int uepm_server_create(uepm_ctx* ctx, uint64_t budget, int max_inflight, uepm_server** out) {
    if (!ctx || !out) return fail(UEPM_ERR_INVALID_ARG, "uepm_server_create: null argument");
    if (budget == 0 || max_inflight < 1) return fail(UEPM_ERR_INVALID_ARG, "uepm_server_create: empty budget");
    HIP_OK(hipSetDevice(ctx->device));
    auto* S = new uepm_server;
    S->ctx = ctx;
    S->budget = budget;
    S->max_inflight = max_inflight;
    S->slots.resize(max_inflight);
    for (auto& sl : S->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&sl.hstat, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
            delete S;
            return fail(UEPM_ERR_DEVICE, "uepm_server_create: stream or event creation failed");
        }
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return UEPM_OK;
}

int uepm_server_submit(uepm_server* S, const uepm_request* q, uint64_t* ticket) {
    if (!S || !q || !ticket || !q->material || (q->nk > 0 && !q->k) || q->nk < 0)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_server_submit: null argument");
    if (q->flags & (UEPM_VECTORS | UEPM_GRADIENTS))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_server_submit: requests return eigenvalues only");
    for (int64_t i = 0; i < 3 * q->nk; ++i)
        if (!std::isfinite(q->k[i]))
            return fail(UEPM_ERR_INVALID_ARG, "uepm_server_submit: k-point " + std::to_string(i / 3) + " is not finite");
    auto r = std::make_shared<Request>();
    r->mat = *q->material;
    r->shells = q->nearest_neighbors;
    r->flags = q->flags;
    r->nb = q->nb;
    r->nk = q->nk;
    r->k.assign(q->k, q->k + 3 * (size_t)q->nk);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) return fail(UEPM_ERR_RUNTIME, "uepm_server_submit: server is shutting down");
        r->id = S->next_id++;
        *ticket = r->id;
        S->live[r->id] = r;
        ++S->stats.submitted;
        if (r->nk == 0) {
            r->status = UEPM_OK;
            r->t_admit = r->t_done = r->t_submit;
            ++S->stats.completed;
        } else {
            S->queue.push_back(r);
        }
    }
    S->cv_work.notify_one();
    return UEPM_OK;
}

int uepm_server_wait(uepm_server* S, uint64_t ticket, double* evals, double* latency_s, double* queue_s) {
    if (!S) return fail(UEPM_ERR_INVALID_ARG, "uepm_server_wait: null server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->live.find(ticket);
        if (it == S->live.end()) return fail(UEPM_ERR_INVALID_ARG, "uepm_server_wait: unknown ticket");
        r = it->second;
        S->cv_done.wait(lk, [&] { return r->status >= 0; });
        S->live.erase(it);
    }
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (queue_s) *queue_s = std::chrono::duration<double>(r->t_admit - r->t_submit).count();
    // a k-point whose solve failed leaves NaN eigenvalues; the others are returned
    if (evals && !r->evals.empty()) std::memcpy(evals, r->evals.data(), r->evals.size() * sizeof(double));
    if (r->status != UEPM_OK) return fail(r->status, r->error);
    return UEPM_OK;
}

int uepm_server_stats_get(uepm_server* S, uepm_server_stats* out) {
    if (!S || !out) return fail(UEPM_ERR_INVALID_ARG, "uepm_server_stats_get: null argument");
    std::lock_guard<std::mutex> g(S->mu);
    *out = S->stats;
    return UEPM_OK;
}

int uepm_server_destroy(uepm_server* S) {
    if (!S) return UEPM_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    for (auto& sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.hk) (void)hipHostFree(sl.hk);
        if (sl.hev) (void)hipHostFree(sl.hev);
        if (sl.hstat) (void)hipHostFree(sl.hstat);
    }
    for (auto& q : S->plans) uepm_plan_destroy(q.plan);
    delete S;
    return UEPM_OK;
}

}  // extern "C"
