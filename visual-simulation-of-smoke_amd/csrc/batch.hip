// Batched plumes (include/visual-simulation-of-smoke-3d-ext.h): independent
// contexts of uneven grid sizes advanced together.
//
// The fused path (every plume in FIXED_ITERATIONS mode, or every plume
// solving to a tolerance with Jacobi or no preconditioner; no network, no
// kernel timing): each kernel of the step is one launch over all plumes
// (batched.hpp: a block finds its plume in the launch's table and runs the
// kernel body with that plume's own block numbering), on one stream of the
// batch.  A plume therefore computes exactly what its own context's step
// would, bit for bit.
//   * Fixed mode: the whole step -- time step, forces, advection, right-hand
//     side, the K CG iterations, projection, scalars -- is one captured
//     hipGraph, replayed each step (the source emitter, whose step number
//     changes, is launched just before it).
//   * Tolerance mode: the solve runs as in run_tolerance_pcg for every plume
//     at once: the iteration passes stop per plume on the device (each
//     plume's own `done` flag, which its pass kernels test); between chunks
//     the host reads every plume's scalar block in one gathered copy; the
//     FP64 refinement passes run over the plumes whose solve is still open
//     (a per-plume gate array); one host thread drives it all.
// Other batches (multilevel or learned preconditioners, kernel timing, mixed
// modes) step each plume's context in turn from the calling thread, on the
// plume's lane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

#include "context.hpp"
#include "visual-simulation-of-smoke-3d-ext.h"

namespace {

using smoke::BatchedLaunch;
using smoke::Context;
using smoke::hip_check;
using smoke::HipError;
using smoke::PcgScalars;
using smoke::TableArena;

// Tables built per solve phase go through a pinned staging buffer and a
// device buffer, each in two halves used alternately: a half is rewritten only
// after the host copy out of it has completed (its event), and its device copy
// is ordered on the stream after every launch that read the previous tables.
struct PhaseTables {
    static constexpr size_t kCapacity = size_t(4) << 20;
    unsigned char* host[2] = {};
    unsigned char* dev[2] = {};
    hipEvent_t copied[2] = {};
    int half = 1;
    TableArena arena;

    void create() {
        for (int h = 0; h < 2; ++h) {
            void* p = nullptr;
            hip_check(hipHostMalloc(&p, kCapacity, hipHostMallocDefault), "hipHostMalloc");
            host[h] = static_cast<unsigned char*>(p);
            hip_check(hipMalloc(&p, kCapacity), "hipMalloc");
            dev[h] = static_cast<unsigned char*>(p);
            hip_check(hipEventCreateWithFlags(&copied[h], hipEventDisableTiming), "hipEventCreate");
        }
    }
    void destroy() {
        for (int h = 0; h < 2; ++h) {
            if (copied[h]) (void)hipEventSynchronize(copied[h]), (void)hipEventDestroy(copied[h]);
            if (host[h]) (void)hipHostFree(host[h]);
            if (dev[h]) (void)hipFree(dev[h]);
            host[h] = dev[h] = nullptr;
            copied[h] = nullptr;
        }
    }
    void begin() {
        half ^= 1;
        arena.clear();
    }
    // copies the phase's tables to the device; returns the device base
    const unsigned char* upload(hipStream_t s) {
        const size_t n = arena.image.size();
        if (n > kCapacity) throw HipError("batch phase tables exceed their buffer");
        hip_check(hipEventSynchronize(copied[half]), "phase tables reuse");
        if (n) {
            std::memcpy(host[half], arena.image.data(), n);
            hip_check(hipMemcpyAsync(dev[half], host[half], n, hipMemcpyHostToDevice, s), "phase tables");
        }
        hip_check(hipEventRecord(copied[half], s), "phase tables event");
        return dev[half];
    }
};

struct PlumeSignature {
    SmokeSimulationSolverOptions options;
    int pcg_shape[2];
    int advect_block[4];
    int iterations;
    bool operator==(const PlumeSignature& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

PlumeSignature signature_of(const Context& c) {
    PlumeSignature s;
    std::memset(&s, 0, sizeof(s));
    s.options = c.options;
    std::memcpy(s.pcg_shape, c.pcg_shape, sizeof(s.pcg_shape));
    std::memcpy(s.advect_block, c.advect_block, sizeof(s.advect_block));
    s.iterations = c.config.pressure_iterations;
    return s;
}

// Which path a batch takes: 0 sequential per-plume steps, 1 fused fixed,
// 2 fused tolerance.
int fused_mode(const std::vector<Context*>& P) {
    bool all_fixed = true, all_tol = true;
    for (const Context* c : P) {
        if (c->timing || c->learned) return 0;
        const bool fixed = c->options.pressure_mode == SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS;
        const bool tol_plain = c->options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE &&
                               (c->options.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_JACOBI ||
                                c->options.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_NONE);
        all_fixed = all_fixed && fixed;
        all_tol = all_tol && tol_plain;
    }
    return all_fixed ? 1 : (all_tol ? 2 : 0);
}

struct FusedPlan {
    int mode = 0;
    std::vector<Context*> ctx;
    std::vector<PlumeSignature> sig;
    TableArena stat;                 // tables of the static launches
    unsigned char* stat_dev = nullptr;
    smoke::BatchedStages stages;
    smoke::BatchedSolve solve;
    PcgScalars* gathered_dev = nullptr;   // one scalar block per plume
    PcgScalars* gathered_host = nullptr;  // pinned, three slots of `count` blocks
    hipEvent_t gathered_event[3] = {};
    PhaseTables phase;
    // fixed mode: the iteration passes of iteration k run for the plumes with
    // more than k iterations (gate arrays in the static tables, one per
    // distinct iteration count), and the whole step is one graph
    std::vector<std::pair<int, size_t>> fixed_gates;   // (iterations, table offset): plumes with more than that many
    int max_iterations = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t launches_per_step = 0;       // non-solve stages + right-hand side
    uint64_t launches_per_iteration = 0;

    void release() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        for (hipEvent_t& e : gathered_event)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        if (gathered_host) (void)hipHostFree(gathered_host);
        if (gathered_dev) (void)hipFree(gathered_dev);
        if (stat_dev) (void)hipFree(stat_dev);
        gathered_host = nullptr;
        gathered_dev = nullptr;
        stat_dev = nullptr;
        phase.destroy();
    }
};

}  // namespace

struct SmokeSimulationBatch_t {
    int device = 0;
    std::vector<hipStream_t> lanes;
    std::vector<SmokeSimulationContext> plumes;
    std::vector<int32_t> lane_of;
    std::vector<std::vector<uint32_t>> order;   // plume indices per lane, ascending
    hipStream_t solve = nullptr;                 // the fused path's stream
    std::vector<hipEvent_t> lane_ready;
    hipEvent_t solved = nullptr;
    std::unique_ptr<FusedPlan> plan;
    bool fusion = true;
    SmokeSimulationBatchInfo info{};
};

namespace {

constexpr int kDefaultLanes = 16;
constexpr int kMaxLanes = 16;

// Longest processing time first: plumes by decreasing cells (stable), each to
// the least loaded lane (lowest index on ties).
void assign_lanes(const uint64_t* cells, uint32_t count, int lanes, int32_t* out) {
    std::vector<uint32_t> idx(count);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return cells[a] > cells[b]; });
    std::vector<uint64_t> load(size_t(lanes), 0);
    for (uint32_t i : idx) {
        int best = 0;
        for (int l = 1; l < lanes; ++l)
            if (load[size_t(l)] < load[size_t(best)]) best = l;
        out[i] = best;
        load[size_t(best)] += cells[i];
    }
}

void release(SmokeSimulationBatch_t& b) {
    (void)hipSetDevice(b.device);
    for (hipStream_t s : b.lanes)
        if (s) (void)hipStreamSynchronize(s);
    if (b.solve) (void)hipStreamSynchronize(b.solve);
    if (b.plan) b.plan->release();
    b.plan.reset();
    for (SmokeSimulationContext c : b.plumes)
        if (c) smoke_simulation_destroy_context_cuda(c);
    for (hipEvent_t e : b.lane_ready)
        if (e) (void)hipEventDestroy(e);
    if (b.solved) (void)hipEventDestroy(b.solved);
    if (b.solve) (void)hipStreamDestroy(b.solve);
    for (hipStream_t s : b.lanes)
        if (s) (void)hipStreamDestroy(s);
    b.lane_ready.clear();
    b.solved = nullptr;
    b.solve = nullptr;
    b.plumes.clear();
    b.lanes.clear();
}

void run_all(const std::vector<BatchedLaunch>& L, const unsigned char* dev, const int* gate, hipStream_t s) {
    for (const BatchedLaunch& l : L) smoke::run_launch(l, dev, gate, s);
}

uint64_t count_launches(const std::vector<BatchedLaunch>& L) { return L.size(); }

// The static launches of a fused step in stream order, up to the solve
// (first) or after it (second).
void enqueue_before_solve(FusedPlan& F, hipStream_t s) {
    run_all(F.stages.time_step, F.stat_dev, nullptr, s);
    run_all(F.stages.forces, F.stat_dev, nullptr, s);
    run_all(F.stages.advect, F.stat_dev, nullptr, s);
    run_all(F.solve.rhs, F.stat_dev, nullptr, s);
}

void enqueue_after_solve(FusedPlan& F, hipStream_t s) {
    run_all(F.stages.project, F.stat_dev, nullptr, s);
    run_all(F.stages.scalars, F.stat_dev, nullptr, s);
}

void enqueue_fixed_step(FusedPlan& F, hipStream_t s) {
    enqueue_before_solve(F, s);
    run_all(F.solve.fixed_start, F.stat_dev, nullptr, s);
    for (int k = 0; k < F.max_iterations; ++k) {
        // plumes with more than k iterations
        const int* gate = nullptr;
        for (const auto& g : F.fixed_gates)
            if (k >= g.first) gate = reinterpret_cast<const int*>(F.stat_dev + g.second);
        run_all(F.solve.iteration, F.stat_dev, gate, s);
    }
    enqueue_after_solve(F, s);
}

std::unique_ptr<FusedPlan> build_plan(SmokeSimulationBatch_t& b, int mode) {
    auto F = std::make_unique<FusedPlan>();
    F->mode = mode;
    for (SmokeSimulationContext c : b.plumes) F->ctx.push_back(static_cast<Context*>(c));
    for (Context* c : F->ctx) F->sig.push_back(signature_of(*c));
    const size_t n = F->ctx.size();
    try {
        void* p = nullptr;
        hip_check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(PcgScalars)), "hipMalloc");
        F->gathered_dev = static_cast<PcgScalars*>(p);
        hip_check(hipHostMalloc(&p, 3 * std::max<size_t>(n, 1) * sizeof(PcgScalars), hipHostMallocDefault), "hipHostMalloc");
        F->gathered_host = static_cast<PcgScalars*>(p);
        for (hipEvent_t& e : F->gathered_event) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        F->phase.create();
        smoke::build_batched_stages(F->ctx, F->stat, F->stages);
        smoke::build_batched_solve(F->ctx, F->gathered_dev, F->stat, F->solve);
        if (mode == 1) {
            std::vector<int> counts;
            for (Context* c : F->ctx) counts.push_back(c->config.pressure_iterations);
            std::vector<int> distinct = counts;
            std::sort(distinct.begin(), distinct.end());
            distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
            F->max_iterations = distinct.back();
            // iteration k >= t (t a plume's count) runs only the plumes with more than t
            for (size_t d = 0; d + 1 < distinct.size(); ++d) {
                std::vector<int> gate(n);
                for (size_t q = 0; q < n; ++q) gate[q] = counts[q] > distinct[d] ? 1 : 0;
                F->fixed_gates.emplace_back(distinct[d], F->stat.add(gate.data(), n * sizeof(int)));
            }
        }
        hip_check(hipMalloc(&p, std::max<size_t>(F->stat.image.size(), 1)), "hipMalloc");
        F->stat_dev = static_cast<unsigned char*>(p);
        hip_check(hipMemcpy(F->stat_dev, F->stat.image.data(), F->stat.image.size(), hipMemcpyHostToDevice), "batch tables");
        F->launches_per_step = count_launches(F->stages.time_step) + count_launches(F->stages.forces) + count_launches(F->stages.advect) +
                               count_launches(F->solve.rhs) + count_launches(F->stages.project) + count_launches(F->stages.scalars);
        F->launches_per_iteration = count_launches(F->solve.iteration);
        if (mode == 1) {
            hip_check(hipStreamSynchronize(b.solve), "sync");
            hip_check(hipStreamBeginCapture(b.solve, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_fixed_step(*F, b.solve);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b.solve, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(b.solve, &F->graph), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&F->graph_exec, F->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        }
    } catch (...) {
        F->release();
        throw;
    }
    return F;
}

// One gathered copy of every plume's scalar block into pinned slot `slot`.
void enqueue_gather(FusedPlan& F, int slot, hipStream_t s) {
    run_all(F.solve.gather, F.stat_dev, nullptr, s);
    const size_t n = F.ctx.size();
    hip_check(hipMemcpyAsync(F.gathered_host + slot * n, F.gathered_dev, n * sizeof(PcgScalars), hipMemcpyDeviceToHost, s), "gather scalars");
    hip_check(hipEventRecord(F.gathered_event[slot], s), "gather event");
}

const PcgScalars* wait_gather(FusedPlan& F, int slot) {
    hip_check(hipEventSynchronize(F.gathered_event[slot]), "gather wait");
    return F.gathered_host + slot * F.ctx.size();
}

// The tolerance solve of every plume (the logic of run_tolerance_pcg).
void fused_tolerance_solve(FusedPlan& F, hipStream_t s) {
    const size_t n = F.ctx.size();
    std::vector<Context*>& P = F.ctx;
    // warm start and the first FP64 residual
    F.phase.begin();
    std::vector<BatchedLaunch> L;
    smoke::build_batched_warm_start(P, F.phase.arena, L);
    smoke::build_batched_residual(P, false, F.phase.arena, L);
    run_all(L, F.phase.upload(s), nullptr, s);
    enqueue_gather(F, 2, s);
    const PcgScalars* first = wait_gather(F, 2);
    std::vector<PcgScalars> host(first, first + n);
    std::vector<int> active(n, 0), passes(n, 0);
    std::vector<int> zero(n, 0);
    bool any_zero = false;
    int chunk = 1 << 30;
    for (size_t p = 0; p < n; ++p) {
        const auto& o = P[p]->options;
        chunk = std::min(chunk, std::max(1, o.check_interval));
        if (host[p].bb == 0.0) {
            // b = 0: the solution is x = 0 exactly
            zero[p] = 1;
            any_zero = true;
            host[p].rr_true = 0.0;
            host[p].done = 1;
        }
        active[p] = host[p].done ? 0 : 1;
    }
    auto upload_gate = [&](const std::vector<int>& gate) -> const int* {
        F.phase.begin();
        const size_t off = F.phase.arena.add(gate.data(), n * sizeof(int));
        return reinterpret_cast<const int*>(F.phase.upload(s) + off);
    };
    if (any_zero) {
        F.phase.begin();
        const size_t off = F.phase.arena.add(zero.data(), n * sizeof(int));
        std::vector<BatchedLaunch> Z;
        smoke::build_batched_zero_x64(P, F.phase.arena, Z);
        const unsigned char* base = F.phase.upload(s);
        run_all(Z, base, reinterpret_cast<const int*>(base + off), s);
    }
    auto any = [&](const std::vector<int>& v) { return std::any_of(v.begin(), v.end(), [](int x) { return x != 0; }); };
    while (any(active)) {
        const int* gate = upload_gate(active);
        // inner solve: chunks of iterations over the open plumes, at most two
        // in flight; each plume's passes stop on its own done flag
        auto enqueue_chunk = [&](int slot) {
            for (int k = 0; k < chunk; ++k) run_all(F.solve.iteration, F.stat_dev, gate, s);
            enqueue_gather(F, slot, s);
        };
        enqueue_chunk(0);
        enqueue_chunk(1);
        int slot = 0;
        for (;;) {
            const PcgScalars* st = wait_gather(F, slot);
            bool all_done = true;
            for (size_t p = 0; p < n && all_done; ++p)
                if (active[p] && !st[p].done) all_done = false;
            if (all_done) break;
            enqueue_chunk(slot);
            slot ^= 1;
        }
        // FP64 refinement of the open plumes: x' = x + e, true residual
        F.phase.begin();
        const size_t off = F.phase.arena.add(active.data(), n * sizeof(int));
        std::vector<BatchedLaunch> R;
        smoke::build_batched_residual(P, true, F.phase.arena, R);
        const unsigned char* base = F.phase.upload(s);
        run_all(R, base, reinterpret_cast<const int*>(base + off), s);
        for (size_t p = 0; p < n; ++p)
            if (active[p]) P[p]->x64_current ^= 1;
        enqueue_gather(F, 2, s);
        const PcgScalars* st = wait_gather(F, 2);
        for (size_t p = 0; p < n; ++p) {
            if (!active[p]) continue;
            host[p] = st[p];
            passes[p] += 1;
            const auto& o = P[p]->options;
            if (host[p].done || passes[p] > o.max_refinements || host[p].iteration >= o.max_iterations) active[p] = 0;
        }
    }
    F.phase.begin();
    std::vector<BatchedLaunch> X;
    smoke::build_batched_pressure_from_x64(P, F.phase.arena, X);
    run_all(X, F.phase.upload(s), nullptr, s);
    for (size_t p = 0; p < n; ++p) {
        Context& c = *P[p];
        const auto& o = c.options;
        const PcgScalars& h = host[p];
        c.solves_done = o.warm_start == 2 ? std::min(c.solves_done + 1, 2) : 1;
        c.stats.iterations = h.iteration;
        c.stats.refinements = passes[p] > 0 ? passes[p] - 1 : 0;
        c.stats.rhs_norm = std::sqrt(h.bb);
        c.stats.relative_residual = h.bb > 0 ? std::sqrt(h.rr_true / h.bb) : 0.0;
        c.stats.converged = h.bb > 0 ? (std::sqrt(h.rr_true) <= o.relative_tolerance * std::sqrt(h.bb)) : 1;
        c.stats.learned_iterations = 0;
        c.stats.learned_fallback = 0;
    }
}

SmokeSimulationResult guarded(SmokeSimulationBatch_t& b, auto&& fn) {
    try {
        if (hipSetDevice(b.device) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        fn();
        return SMOKE_SIMULATION_RESULT_OK;
    } catch (const std::bad_alloc&) {
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    } catch (...) {
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
}

void fused_step(SmokeSimulationBatch_t& b, const SmokeSimulationPlumeSource* sources, uint32_t step) {
    FusedPlan& F = *b.plan;
    hipStream_t s = b.solve;
    // the plumes' own streams (updates, views) are ordered around the batch's
    for (size_t l = 0; l < b.lanes.size(); ++l) {
        hip_check(hipEventRecord(b.lane_ready[l], b.lanes[l]), "lane event");
        hip_check(hipStreamWaitEvent(s, b.lane_ready[l], 0), "lane wait");
    }
    if (sources) {
        F.phase.begin();
        std::vector<BatchedLaunch> S;
        smoke::build_batched_sources(F.ctx, sources, step, F.phase.arena, S);
        run_all(S, F.phase.upload(s), nullptr, s);
    }
    if (F.mode == 1) {
        hip_check(hipGraphLaunch(F.graph_exec, s), "hipGraphLaunch");
        for (Context* c : F.ctx) c->stats.iterations = c->config.pressure_iterations;
    } else {
        enqueue_before_solve(F, s);
        fused_tolerance_solve(F, s);
        enqueue_after_solve(F, s);
    }
    for (Context* c : F.ctx) {
        c->stats.steps += 1;
        c->diagnostics_dirty = true;
        c->step_info_pending = false;   // this step's dt was not copied out: the stats read it from the device
    }
    hip_check(hipEventRecord(b.solved, s), "batch event");
    for (hipStream_t lane : b.lanes) hip_check(hipStreamWaitEvent(lane, b.solved, 0), "lane wait");
}

// the plan for the plumes' current options (rebuilt when any of them changed)
int prepare_plan(SmokeSimulationBatch_t& b) {
    std::vector<Context*> P;
    for (SmokeSimulationContext c : b.plumes) P.push_back(static_cast<Context*>(c));
    const int mode = b.fusion ? fused_mode(P) : 0;
    bool same = b.plan && b.plan->mode == mode;
    for (size_t p = 0; same && p < P.size(); ++p) same = b.plan->sig[p] == signature_of(*P[p]);
    if (same) return mode;
    if (b.plan) {
        hip_check(hipStreamSynchronize(b.solve), "sync");
        b.plan->release();
        b.plan.reset();
    }
    if (mode != 0) b.plan = build_plan(b, mode);
    return mode;
}

}  // namespace

extern "C" {

SmokeSimulationResult smoke_simulation_ext_batch_lanes(const uint64_t* cells, uint32_t count, int32_t lanes, int32_t* out_lane) {
    if ((count > 0 && (!cells || !out_lane)) || lanes < 1 || lanes > kMaxLanes) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    try {
        assign_lanes(cells, count, lanes, out_lane);
    } catch (const std::bad_alloc&) {
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    }
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_create_batch(const SmokeSimulationBatchCreateDesc* desc, SmokeSimulationBatch* out_batch) {
    if (!desc || !out_batch || !desc->plumes || desc->count == 0 || desc->lanes < 0 || desc->lanes > kMaxLanes)
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    *out_batch = nullptr;
    std::unique_ptr<SmokeSimulationBatch_t> b{new (std::nothrow) SmokeSimulationBatch_t{}};
    if (!b) return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    SmokeSimulationResult result = SMOKE_SIMULATION_RESULT_OK;
    try {
        if (hipGetDevice(&b->device) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        const uint32_t n = desc->count;
        const int lanes = int(std::min<uint32_t>(n, uint32_t(desc->lanes ? desc->lanes : kDefaultLanes)));
        std::vector<uint64_t> cells(n);
        for (uint32_t i = 0; i < n; ++i) {
            const SmokeSimulationConfig& cfg = desc->plumes[i].config;
            cells[i] = uint64_t(cfg.nx) * cfg.ny * cfg.nz;
        }
        b->lane_of.resize(n);
        assign_lanes(cells.data(), n, lanes, b->lane_of.data());
        b->order.resize(size_t(lanes));
        for (uint32_t i = 0; i < n; ++i) b->order[size_t(b->lane_of[i])].push_back(i);
        for (int l = 0; l < lanes; ++l) {
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
                release(*b);
                return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
            }
            b->lanes.push_back(s);
        }
        if (hipStreamCreateWithFlags(&b->solve, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->solved, hipEventDisableTiming) != hipSuccess) {
            release(*b);
            return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        }
        b->lane_ready.assign(size_t(lanes), nullptr);
        for (int l = 0; l < lanes; ++l) {
            if (hipEventCreateWithFlags(&b->lane_ready[size_t(l)], hipEventDisableTiming) != hipSuccess) {
                release(*b);
                return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
            }
        }
        b->plumes.assign(n, nullptr);
        for (uint32_t i = 0; i < n; ++i) {
            SmokeSimulationContextCreateDesc d = desc->plumes[i];
            d.stream = b->lanes[size_t(b->lane_of[i])];
            result = smoke_simulation_create_context_cuda(&d, &b->plumes[i]);
            if (result != SMOKE_SIMULATION_RESULT_OK) break;
        }
    } catch (const std::bad_alloc&) {
        result = SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    }
    if (result != SMOKE_SIMULATION_RESULT_OK) {
        release(*b);
        return result;
    }
    *out_batch = b.release();
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_destroy_batch(SmokeSimulationBatch batch) {
    if (!batch) return SMOKE_SIMULATION_RESULT_OK;
    release(*batch);
    delete batch;
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_batch_context(SmokeSimulationBatch batch, uint32_t index, SmokeSimulationContext* out_context,
                                                         int32_t* out_lane) {
    if (!batch || index >= batch->plumes.size()) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (out_context) *out_context = batch->plumes[index];
    if (out_lane) *out_lane = batch->lane_of[index];
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_step_batch(SmokeSimulationBatch batch, const SmokeSimulationPlumeSource* sources, uint32_t step) {
    if (!batch) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SmokeSimulationBatch_t& b = *batch;
    int mode = 0;
    SmokeSimulationResult res = guarded(b, [&] { mode = prepare_plan(b); });
    if (res != SMOKE_SIMULATION_RESULT_OK) return res;
    if (mode != 0) {
        res = guarded(b, [&] { fused_step(b, sources, step); });
        if (res == SMOKE_SIMULATION_RESULT_OK) {
            b.info.fused = mode;
            b.info.steps += 1;
            b.info.launches_per_step = b.plan->launches_per_step;
            b.info.launches_per_iteration = b.plan->launches_per_iteration;
            b.info.graph = b.plan->graph_exec ? 1 : 0;
        }
        return res;
    }
    // every other batch: each plume's own step in turn, on its lane.  A
    // network's activations belong to the network: one network attached to
    // plumes on two lanes would be written from two streams, so such a batch
    // is refused.
    {
        std::vector<std::vector<const void*>> nets(b.lanes.size());
        for (size_t l = 0; l < b.lanes.size(); ++l)
            for (uint32_t p : b.order[l])
                if (const void* net = static_cast<const Context*>(b.plumes[p])->learned) nets[l].push_back(net);
        for (size_t l = 0; l < nets.size(); ++l)
            for (size_t m = l + 1; m < nets.size(); ++m)
                for (const void* net : nets[l])
                    if (std::find(nets[m].begin(), nets[m].end(), net) != nets[m].end()) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
    b.info.fused = 0;
    b.info.steps += 1;
    size_t rounds = 0;
    for (const auto& o : b.order) rounds = std::max(rounds, o.size());
    for (size_t r = 0; r < rounds; ++r)
        for (size_t l = 0; l < b.lanes.size(); ++l) {
            if (r >= b.order[l].size()) continue;
            const uint32_t p = b.order[l][r];
            if (sources) {
                res = smoke_simulation_ext_plume_sources(b.plumes[p], &sources[p], step);
                if (res != SMOKE_SIMULATION_RESULT_OK) return res;
            }
            res = smoke_simulation_step_cuda(b.plumes[p]);
            if (res != SMOKE_SIMULATION_RESULT_OK) return res;
        }
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_synchronize_batch(SmokeSimulationBatch batch) {
    if (!batch) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (hipSetDevice(batch->device) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    for (hipStream_t s : batch->lanes)
        if (hipStreamSynchronize(s) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (batch->solve && hipStreamSynchronize(batch->solve) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_set_batch_fusion(SmokeSimulationBatch batch, int32_t enable) {
    if (!batch) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    batch->fusion = enable != 0;
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_get_batch_info(SmokeSimulationBatch batch, SmokeSimulationBatchInfo* out_info) {
    if (!batch || !out_info) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    *out_info = batch->info;
    return SMOKE_SIMULATION_RESULT_OK;
}

}  // extern "C"
