// Batched plumes (include/visual-simulation-of-smoke-3d-ext.h): independent
// contexts of uneven grid sizes on a few HIP streams ("lanes").  Small grids
// leave most of the 256 CUs idle in any one launch; stepping plumes on
// separate lanes lets the hardware overlap their kernels.  Each plume is an
// ordinary context, so its arithmetic -- and its bits -- are those of a
// context created alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <new>
#include <numeric>
#include <set>
#include <thread>
#include <vector>

#include "context.hpp"
#include "visual-simulation-of-smoke-3d-ext.h"

struct SmokeSimulationBatch_t {
    int device = 0;
    std::vector<hipStream_t> lanes;
    std::vector<SmokeSimulationContext> plumes;
    std::vector<int32_t> lane_of;
    std::vector<std::vector<uint32_t>> order;   // plume indices per lane, ascending
    // batched fixed-iteration solves: a stream for the iterations of every
    // plume, the events joining it with the lanes, and per group of plumes
    // with the same launch shapes and iteration count its table in device
    // memory (kept while the groups stay the same)
    hipStream_t solve = nullptr;
    std::vector<hipEvent_t> lane_ready;
    hipEvent_t solved = nullptr;
    struct SolveGroup {
        std::vector<uint32_t> plumes;
        int shape0 = 0, shape1 = 0, iterations = 0;
        void* table = nullptr;
    };
    std::vector<SolveGroup> groups;
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

void release_groups(SmokeSimulationBatch_t& b) {
    for (auto& g : b.groups)
        if (g.table) (void)hipFree(g.table);
    b.groups.clear();
}

void release(SmokeSimulationBatch_t& b) {
    hipSetDevice(b.device);
    for (hipStream_t s : b.lanes)
        if (s) hipStreamSynchronize(s);
    if (b.solve) hipStreamSynchronize(b.solve);
    for (SmokeSimulationContext c : b.plumes)
        if (c) smoke_simulation_destroy_context_cuda(c);
    release_groups(b);
    for (hipEvent_t e : b.lane_ready)
        if (e) hipEventDestroy(e);
    if (b.solved) hipEventDestroy(b.solved);
    if (b.solve) hipStreamDestroy(b.solve);
    for (hipStream_t s : b.lanes)
        if (s) hipStreamDestroy(s);
    b.lane_ready.clear();
    b.solved = nullptr;
    b.solve = nullptr;
    b.plumes.clear();
    b.lanes.clear();
}

// Runs `fn` on a plume's context the way the ABI entry points do: its device
// selected, exceptions turned into result codes.
template <class F>
SmokeSimulationResult guarded(SmokeSimulationBatch_t& b, uint32_t p, F&& fn) {
    try {
        if (hipSetDevice(b.device) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        fn(static_cast<smoke::Context&>(*b.plumes[p]));
        return SMOKE_SIMULATION_RESULT_OK;
    } catch (const std::bad_alloc&) {
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    } catch (...) {
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
}

// The groups of the batched solve for the plumes' current options: plumes
// with the same stencil launch shapes and iteration count, in index order.
// Tables are rebuilt only when the grouping changed (after the solve stream
// has finished with the old ones).
SmokeSimulationResult prepare_groups(SmokeSimulationBatch_t& b) {
    std::vector<SmokeSimulationBatch_t::SolveGroup> want;
    for (uint32_t p = 0; p < b.plumes.size(); ++p) {
        const smoke::Context& c = *b.plumes[p];
        auto it = std::find_if(want.begin(), want.end(), [&](const SmokeSimulationBatch_t::SolveGroup& g) {
            return g.shape0 == c.pcg_shape[0] && g.shape1 == c.pcg_shape[1] && g.iterations == c.config.pressure_iterations;
        });
        if (it == want.end()) {
            want.emplace_back();
            it = want.end() - 1;
            it->shape0 = c.pcg_shape[0];
            it->shape1 = c.pcg_shape[1];
            it->iterations = c.config.pressure_iterations;
        }
        it->plumes.push_back(p);
    }
    bool same = want.size() == b.groups.size();
    for (size_t g = 0; same && g < want.size(); ++g)
        same = want[g].plumes == b.groups[g].plumes && want[g].shape0 == b.groups[g].shape0 && want[g].shape1 == b.groups[g].shape1 &&
               want[g].iterations == b.groups[g].iterations;
    if (same) return SMOKE_SIMULATION_RESULT_OK;
    if (hipStreamSynchronize(b.solve) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    release_groups(b);
    try {
        for (auto& g : want) {
            std::vector<smoke::Context*> ctx;
            for (uint32_t p : g.plumes) ctx.push_back(b.plumes[p]);
            std::vector<unsigned char> image;
            smoke::fill_batched_table(ctx, image);
            if (hipMalloc(&g.table, image.size()) != hipSuccess) {
                g.table = nullptr;
                b.groups = std::move(want);
                release_groups(b);
                return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
            }
            if (hipMemcpy(g.table, image.data(), image.size(), hipMemcpyHostToDevice) != hipSuccess) {
                b.groups = std::move(want);
                release_groups(b);
                return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
            }
        }
    } catch (const std::bad_alloc&) {
        b.groups = std::move(want);
        release_groups(b);
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    }
    b.groups = std::move(want);
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult step_plume(SmokeSimulationBatch_t& b, uint32_t p, const SmokeSimulationPlumeSource* sources, uint32_t step) {
    if (sources) {
        const SmokeSimulationResult r = smoke_simulation_ext_plume_sources(b.plumes[p], &sources[p], step);
        if (r != SMOKE_SIMULATION_RESULT_OK) return r;
    }
    return smoke_simulation_step_cuda(b.plumes[p]);
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
    const size_t lanes = b.lanes.size();
    // a network's activations are per network: one network on two lanes
    // would be written by two streams at once
    bool host_waits = false;
    std::vector<std::set<const void*>> nets(lanes);
    for (size_t l = 0; l < lanes; ++l) {
        for (uint32_t p : b.order[l]) {
            const smoke::Context& c = *b.plumes[p];
            if (c.options.pressure_mode != SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS || !c.graph_exec || c.timing) host_waits = true;
            if (c.learned) nets[l].insert(c.learned);
        }
    }
    for (size_t l = 0; l < lanes; ++l)
        for (size_t m = l + 1; m < lanes; ++m)
            for (const void* n : nets[l])
                if (nets[m].count(n)) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (!host_waits) {
        // Every plume is a fixed-iteration step: the lanes run each plume's
        // step up to its CG iterations, the solve stream runs the iterations
        // of all plumes as one launch per pass and group, and the lanes
        // finish the steps.  Round r of each phase enqueues the r-th plume
        // of every lane, so that every lane has work queued early.
        SmokeSimulationResult res = prepare_groups(b);
        if (res != SMOKE_SIMULATION_RESULT_OK) return res;
        size_t rounds = 0;
        for (const auto& o : b.order) rounds = std::max(rounds, o.size());
        for (size_t r = 0; r < rounds; ++r)
            for (size_t l = 0; l < lanes; ++l) {
                if (r >= b.order[l].size()) continue;
                const uint32_t p = b.order[l][r];
                if (sources) {
                    res = smoke_simulation_ext_plume_sources(b.plumes[p], &sources[p], step);
                    if (res != SMOKE_SIMULATION_RESULT_OK) return res;
                }
                res = guarded(b, p, [](smoke::Context& c) { smoke::enqueue_fixed_step_before_iterations(c); });
                if (res != SMOKE_SIMULATION_RESULT_OK) return res;
            }
        for (size_t l = 0; l < lanes; ++l)
            if (hipEventRecord(b.lane_ready[l], b.lanes[l]) != hipSuccess || hipStreamWaitEvent(b.solve, b.lane_ready[l], 0) != hipSuccess)
                return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        for (const auto& g : b.groups) {
            std::vector<smoke::Context*> ctx;
            for (uint32_t p : g.plumes) ctx.push_back(b.plumes[p]);
            try {
                smoke::enqueue_batched_iterations(ctx, g.table, b.solve);
            } catch (...) {
                return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
            }
        }
        if (hipEventRecord(b.solved, b.solve) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        for (size_t l = 0; l < lanes; ++l)
            if (hipStreamWaitEvent(b.lanes[l], b.solved, 0) != hipSuccess) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        for (size_t r = 0; r < rounds; ++r)
            for (size_t l = 0; l < lanes; ++l) {
                if (r >= b.order[l].size()) continue;
                res = guarded(b, b.order[l][r], [](smoke::Context& c) { smoke::enqueue_fixed_step_after_iterations(c); });
                if (res != SMOKE_SIMULATION_RESULT_OK) return res;
            }
        return SMOKE_SIMULATION_RESULT_OK;
    }
    std::vector<SmokeSimulationResult> results(lanes, SMOKE_SIMULATION_RESULT_OK);
    std::vector<std::thread> workers;
    try {
        workers.reserve(lanes);
        for (size_t l = 0; l < lanes; ++l) {
            workers.emplace_back([&b, &results, l, sources, step] {
                if (hipSetDevice(b.device) != hipSuccess) {
                    results[l] = SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
                    return;
                }
                for (uint32_t p : b.order[l]) {
                    results[l] = step_plume(b, p, sources, step);
                    if (results[l] != SMOKE_SIMULATION_RESULT_OK) return;
                }
            });
        }
        for (std::thread& t : workers) t.join();
    } catch (...) {
        // a worker that could not start: wait for the ones that did
        for (std::thread& t : workers)
            if (t.joinable()) t.join();
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
    for (SmokeSimulationResult r : results)
        if (r != SMOKE_SIMULATION_RESULT_OK) return r;
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

}  // extern "C"
