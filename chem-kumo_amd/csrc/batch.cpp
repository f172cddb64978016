// Batched evaluation of many (molecule, grid) items: items are spread over
// several executors (streams with their own workspaces, longest items first
// onto the least-loaded stream) and evaluated with run_range in capture mode,
// so that no host synchronisation separates them. With a HIP graph the whole
// multi-stream sequence is recorded once (fork/join events tie the streams to
// the capture) and replayed by a single graph launch per run.
#include "runtime.hpp"

#include <algorithm>
#include <memory>
#include <numeric>

struct kumo_hip_batch {
    struct Item {
        kumo_hip_molecule* mol = nullptr;
        kumo_hip_grid grid{};
        kumo::Prepared pr;
        kumo::PointSource ps{};
        std::vector<double*> out;
        int exec = 0;
    };
    kumo_hip_context* ctx = nullptr;
    kumo_hip_options opt{};
    std::vector<Item> items;
    std::vector<std::unique_ptr<kumo::Executor>> execs;
    std::deque<std::vector<char>> graph_keep, run_keep;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> fork, join;
    double last_ms = 0.0;
    bool warmed = false;
    // Launch statistics of the last timed run (use_graph == 2), summed over
    // the executors.
    double kernel_ms = 0.0, flops = 0.0, rescue_ms = 0.0, fields_ms = 0.0;
    long long launches = 0, rescue_points = 0, rescue_flagged = 0, fields_bytes = 0;
    // Algorithmic FLOPs and rescued points of one run, read back in the first
    // (uncaptured) run: they do not change between runs.
    double warm_flops = 0.0;
    long long warm_rescue_points = 0, warm_rescue_flagged = 0;

    ~kumo_hip_batch() {
        if (!execs.empty() && execs[0]->stream) (void)hipStreamSynchronize(execs[0]->stream);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : fork)
            (void)hipEventDestroy(e);
        for (hipEvent_t e : join)
            (void)hipEventDestroy(e);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        for (auto& e : execs)
            e->destroy();
    }

    // Issues every item on its executor; the executors' streams are forked
    // from and joined back to stream 0. Serial: each executor's stream starts
    // when the previous one's work has ended (items issued executor by
    // executor), so no two launches overlap.
    void issue(bool capture, std::deque<std::vector<char>>* keep, bool serial = false) {
        hipStream_t s0 = execs[0]->stream;
        for (auto& e : execs) {
            e->capture = capture;
            e->keep = keep;
            e->reset_stats();
        }
        const kumo::Advance none = [](std::size_t) {};
        const auto run_item = [&](Item& it) {
            kumo::run_range(*execs[static_cast<std::size_t>(it.exec)], it.mol, it.pr, it.ps, opt, it.out.data(), 0,
                            it.pr.n, none);
        };
        if (serial) {
            for (std::size_t k = 0; k < execs.size(); ++k) {
                if (k > 0) {
                    KUMO_HIP(hipEventRecord(fork[k], execs[k - 1]->stream));
                    KUMO_HIP(hipStreamWaitEvent(execs[k]->stream, fork[k], 0));
                }
                for (auto& it : items)
                    if (static_cast<std::size_t>(it.exec) == k) run_item(it);
            }
            if (execs.size() > 1) {
                KUMO_HIP(hipEventRecord(join[execs.size() - 1], execs.back()->stream));
                KUMO_HIP(hipStreamWaitEvent(s0, join[execs.size() - 1], 0));
            }
        } else {
            for (std::size_t k = 1; k < execs.size(); ++k) {
                KUMO_HIP(hipEventRecord(fork[k], s0));
                KUMO_HIP(hipStreamWaitEvent(execs[k]->stream, fork[k], 0));
            }
            for (auto& it : items)
                run_item(it);
            for (std::size_t k = 1; k < execs.size(); ++k) {
                KUMO_HIP(hipEventRecord(join[k], execs[k]->stream));
                KUMO_HIP(hipStreamWaitEvent(s0, join[k], 0));
            }
        }
        for (auto& e : execs) {
            e->capture = false;
            e->keep = nullptr;
        }
    }
};

namespace {

template <typename F> int guarded(kumo_hip_context* ctx, F&& f) {
    try {
        f();
        return KUMO_HIP_OK;
    } catch (const kumo::Error& e) {
        if (ctx) ctx->last_error = e.what();
        return e.status;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return KUMO_HIP_RUNTIME_ERROR;
    }
}

} // namespace

extern "C" {

int kumo_hip_batch_create(kumo_hip_context* ctx, size_t nitems, const kumo_hip_molecule* const* molecules,
                          const kumo_hip_grid* grids, const kumo_hip_property* properties, size_t nproperties,
                          const kumo_hip_options* options, int32_t streams, double* const* outputs,
                          kumo_hip_batch** batch) {
    if (!ctx || !batch || (nitems && (!molecules || !grids || !outputs)) || (nproperties && !properties))
        return KUMO_HIP_INVALID_ARGUMENT;
    *batch = nullptr;
    return guarded(ctx, [&] {
        KUMO_HIP(hipSetDevice(ctx->device));
        auto b = std::make_unique<kumo_hip_batch>();
        b->ctx = ctx;
        kumo_hip_default_options(&b->opt);
        if (options) b->opt = *options;
        b->opt.progress = nullptr;
        b->opt.device_outputs = 1;
        const int nstreams = streams > 0 ? std::min<int>(streams, 16) : 4;
        std::vector<double> load(static_cast<std::size_t>(nstreams), 0.0);
        std::size_t nch = 0;
        for (std::size_t i = 0; i < nitems; ++i) {
            if (!molecules[i]) kumo::invalid("batch item without a molecule");
            kumo_hip_batch::Item it;
            it.mol = const_cast<kumo_hip_molecule*>(molecules[i]);
            it.grid = grids[i];
            kumo::PointsArg pts;
            pts.grid = &it.grid;
            it.pr = kumo::prepare(it.mol, pts, properties, nproperties, b->opt, true);
            nch = it.pr.nout;
            it.ps = kumo::make_source(&it.grid, nullptr);
            it.out.assign(outputs + i * nch, outputs + (i + 1) * nch);
            b->items.push_back(std::move(it));
        }
        // Longest items first onto the least-loaded stream (work ~ points x AOs).
        std::vector<std::size_t> order(b->items.size());
        std::iota(order.begin(), order.end(), 0);
        const auto work = [&](std::size_t i) {
            return static_cast<double>(b->items[i].pr.n) * static_cast<double>(b->items[i].mol->n);
        };
        std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) { return work(x) > work(y); });
        for (std::size_t i : order) {
            const auto k = static_cast<std::size_t>(std::min_element(load.begin(), load.end()) - load.begin());
            b->items[i].exec = static_cast<int>(k);
            load[k] += work(i);
        }
        for (int k = 0; k < nstreams; ++k) {
            auto e = std::make_unique<kumo::Executor>();
            e->create(ctx->device, ctx->boys_table.as<double>());
            e->exact_esp = true;
            b->execs.push_back(std::move(e));
        }
        b->fork.assign(static_cast<std::size_t>(nstreams), nullptr);
        b->join.assign(static_cast<std::size_t>(nstreams), nullptr);
        for (int k = 0; k < nstreams; ++k) {
            KUMO_HIP(hipEventCreateWithFlags(&b->fork[static_cast<std::size_t>(k)], hipEventDisableTiming));
            KUMO_HIP(hipEventCreateWithFlags(&b->join[static_cast<std::size_t>(k)], hipEventDisableTiming));
        }
        KUMO_HIP(hipEventCreate(&b->t0));
        KUMO_HIP(hipEventCreate(&b->t1));
        *batch = b.release();
    });
}

int kumo_hip_batch_run(kumo_hip_batch* b, int32_t use_graph) {
    if (!b) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(b->ctx, [&] {
        KUMO_HIP(hipSetDevice(b->ctx->device));
        hipStream_t s0 = b->execs[0]->stream;
        if (!b->warmed) {
            // Sizes every workspace and builds the molecules' lazily made
            // tables outside any capture.
            b->issue(false, nullptr);
            for (auto& e : b->execs) {
                KUMO_HIP(hipStreamSynchronize(e->stream));
                b->warm_flops += e->last_flops;
                b->warm_rescue_points += e->last_rescue_points;
                b->warm_rescue_flagged += e->last_rescue_flagged;
            }
            b->warmed = true;
        }
        if (use_graph == 1 && !b->graph_exec) {
            KUMO_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
            try {
                b->issue(true, &b->graph_keep);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s0, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            KUMO_HIP(hipStreamEndCapture(s0, &b->graph));
            KUMO_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        }
        if (use_graph == 2 || use_graph == 3) {
            // Timed run: every launch issued again without a graph but, as
            // in the graph, with no host synchronisation, so the streams'
            // launches overlap as they do on replay (3: the streams chained,
            // no overlap); each contraction, rescue and field launch is
            // bracketed by events on its stream. The FLOPs and rescued points
            // are those read back in the first run.
            b->run_keep.clear();
            for (auto& e : b->execs)
                e->events_in_capture = true;
            KUMO_HIP(hipEventRecord(b->t0, s0));
            try {
                b->issue(true, &b->run_keep, use_graph == 3);
            } catch (...) {
                for (auto& e : b->execs)
                    e->events_in_capture = false;
                throw;
            }
            KUMO_HIP(hipEventRecord(b->t1, s0));
            KUMO_HIP(hipEventSynchronize(b->t1));
            b->kernel_ms = b->rescue_ms = b->fields_ms = 0.0;
            b->launches = b->fields_bytes = 0;
            b->flops = b->warm_flops;
            b->rescue_points = b->warm_rescue_points;
            b->rescue_flagged = b->warm_rescue_flagged;
            for (auto& e : b->execs) {
                e->events_in_capture = false;
                KUMO_HIP(hipStreamSynchronize(e->stream));
                e->collect_timing();
                b->kernel_ms += e->last_kernel_ms;
                b->launches += e->last_launches;
                b->rescue_ms += e->last_rescue_ms;
                b->fields_ms += e->last_fields_ms;
                b->fields_bytes += e->last_fields_bytes;
            }
            float ms = 0.f;
            KUMO_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
            b->last_ms = ms;
            return;
        }
        KUMO_HIP(hipEventRecord(b->t0, s0));
        if (use_graph) {
            KUMO_HIP(hipGraphLaunch(b->graph_exec, s0));
        } else {
            b->run_keep.clear();
            b->issue(true, &b->run_keep);
        }
        KUMO_HIP(hipEventRecord(b->t1, s0));
        KUMO_HIP(hipEventSynchronize(b->t1));
        float ms = 0.f;
        KUMO_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
        b->last_ms = ms;
    });
}

int kumo_hip_batch_last_ms(const kumo_hip_batch* b, double* ms) {
    if (!b || !ms) return KUMO_HIP_INVALID_ARGUMENT;
    *ms = b->last_ms;
    return KUMO_HIP_OK;
}

int kumo_hip_batch_kernel_stats(const kumo_hip_batch* b, double* kernel_ms, int64_t* launches, double* flops,
                                double* rescue_ms, int64_t* rescue_points, double* fields_ms, int64_t* fields_bytes,
                                int64_t* rescue_flagged) {
    if (!b) return KUMO_HIP_INVALID_ARGUMENT;
    if (kernel_ms) *kernel_ms = b->kernel_ms;
    if (launches) *launches = b->launches;
    if (flops) *flops = b->flops;
    if (rescue_ms) *rescue_ms = b->rescue_ms;
    if (rescue_points) *rescue_points = b->rescue_points;
    if (fields_ms) *fields_ms = b->fields_ms;
    if (fields_bytes) *fields_bytes = b->fields_bytes;
    if (rescue_flagged) *rescue_flagged = b->rescue_flagged;
    return KUMO_HIP_OK;
}

int kumo_hip_batch_release(kumo_hip_batch* b) {
    delete b;
    return KUMO_HIP_OK;
}

} // extern "C"
