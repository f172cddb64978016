// Serving loop (kumo_hip_server_*): continuous batching of requests over
// resident molecules.
//
// Submission validates and plans a request on the caller's thread and queues
// it. Admission is in arrival order against an HBM budget: an admitted
// request owns device memory for its outputs (and listed points) from a pool
// that is never returned to the driver while the server lives, so steady
// serving allocates nothing. Workers (one HIP stream and workspace set each)
// repeatedly take the next block-aligned range of the admitted request with
// the fewest points left and evaluate it with run_range; ranges of different
// requests therefore overlap on the device, and a short request admitted
// while a cube is in flight is served within about one range time. The
// worker that finishes a request's last range copies its results to the
// host, releases its budget and admits waiting requests. Every screening
// block is evaluated by the same kernels whatever range or request mix it
// falls in, so results equal those of a lone evaluation.
#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    kumo_hip_molecule* mol = nullptr;
    kumo_hip_grid grid{};
    bool is_grid = false;
    std::vector<double> points;
    kumo::Prepared pr;
    kumo::PointSource ps{};
    kumo_hip_options opt{};
    std::vector<double*> host_out;
    void* dmem = nullptr;
    std::size_t bytes = 0;
    std::vector<double*> dout;
    long long next = 0;
    int inflight = 0;
    bool admitted = false, ready = false, done = false, failed = false;
    int status = KUMO_HIP_OK;
    std::string error;
    Clock::time_point t_submit, t_start, t_done;
    bool started = false;
};

// Device memory reused across requests: blocks are kept by size and handed
// out again for any request that fits in at most twice their size.
class Pool {
public:
    void* get(std::size_t bytes) {
        std::lock_guard<std::mutex> g(m_);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            sizes_[p] = it->first;
            free_.erase(it);
            return p;
        }
        void* p = nullptr;
        KUMO_HIP(hipMalloc(&p, bytes));
        sizes_[p] = bytes;
        return p;
    }
    void put(void* p) {
        std::lock_guard<std::mutex> g(m_);
        free_.emplace(sizes_.at(p), p);
    }
    void clear() {
        std::lock_guard<std::mutex> g(m_);
        for (auto& kv : free_)
            (void)hipFree(kv.second);
        free_.clear();
        sizes_.clear();
    }

private:
    std::mutex m_;
    std::multimap<std::size_t, void*> free_;
    std::unordered_map<void*, std::size_t> sizes_;
};

} // namespace

struct kumo_hip_server {
    kumo_hip_context* ctx = nullptr;
    std::vector<std::unique_ptr<kumo::Executor>> execs;
    std::vector<std::thread> workers;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::vector<std::shared_ptr<Request>> active;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
    uint64_t next_id = 1;
    std::size_t budget = 8ull << 30, used = 0;
    long long chunk = 131072;
    int preparing = 0;
    bool stop = false;
    Pool pool;

    // Moves waiting requests into the budget, in arrival order (m held).
    // Returns those whose device memory the caller must set up.
    std::vector<std::shared_ptr<Request>> admit_locked() {
        std::vector<std::shared_ptr<Request>> out;
        while (!pending.empty()) {
            auto& r = pending.front();
            const bool alone = active.empty() && preparing == 0 && out.empty();
            if (!alone && used + r->bytes > budget) break;
            used += r->bytes;
            r->admitted = true;
            out.push_back(r);
            pending.pop_front();
        }
        preparing += static_cast<int>(out.size());
        return out;
    }

    // Device memory and point upload of newly admitted requests (m not held).
    void prepare(std::vector<std::shared_ptr<Request>> list, hipStream_t stream) {
        while (!list.empty()) {
            prepare_once(list, stream);
            std::lock_guard<std::mutex> g(m);
            list = admit_locked(); // budget freed by requests that ended here
        }
    }

    void prepare_once(std::vector<std::shared_ptr<Request>>& list, hipStream_t stream) {
        for (auto& r : list) {
            try {
                r->dmem = pool.get(r->bytes);
                char* p = static_cast<char*>(r->dmem);
                const long long n = r->pr.n;
                r->dout.resize(r->pr.nout);
                for (std::size_t c = 0; c < r->pr.nout; ++c) {
                    r->dout[c] = reinterpret_cast<double*>(p);
                    p += static_cast<std::size_t>(n) * sizeof(double);
                }
                const double* dpts = nullptr;
                if (!r->is_grid && n > 0) {
                    KUMO_HIP(hipMemcpyAsync(p, r->points.data(), r->points.size() * sizeof(double),
                                            hipMemcpyHostToDevice, stream));
                    KUMO_HIP(hipStreamSynchronize(stream));
                    dpts = reinterpret_cast<const double*>(p);
                }
                r->ps = kumo::make_source(r->is_grid ? &r->grid : nullptr, dpts);
            } catch (const kumo::Error& e) {
                fail(r, e.status, e.what());
            }
        }
        std::lock_guard<std::mutex> g(m);
        for (auto& r : list) {
            --preparing;
            if (r->failed) {
                finish_locked(r);
                continue;
            }
            r->ready = true;
            if (r->pr.n == 0 || r->pr.nout == 0) {
                finish_locked(r);
                continue;
            }
            active.push_back(r);
        }
        cv_work.notify_all();
        cv_done.notify_all();
    }

    void fail(const std::shared_ptr<Request>& r, int status, const std::string& what) {
        std::lock_guard<std::mutex> g(m);
        if (!r->failed) {
            r->failed = true;
            r->status = status;
            r->error = what;
        }
        r->next = r->pr.n; // no further ranges
    }

    // Request done (m held): budget released, waiters woken.
    void finish_locked(const std::shared_ptr<Request>& r) {
        r->done = true;
        r->t_done = Clock::now();
        if (!r->started) r->t_start = r->t_done;
        if (r->admitted) used -= r->bytes;
        if (r->dmem) pool.put(r->dmem);
        r->dmem = nullptr;
        active.erase(std::remove(active.begin(), active.end(), r), active.end());
    }

    void worker(std::size_t k) {
        (void)hipSetDevice(ctx->device);
        kumo::Executor& ex = *execs[k];
        const kumo::Advance none = [](std::size_t) {};
        for (;;) {
            std::shared_ptr<Request> r;
            long long q0 = 0, q1 = 0;
            {
                std::unique_lock<std::mutex> lock(m);
                const auto has_work = [&] {
                    for (auto& a : active)
                        if (a->next < a->pr.n) return true;
                    return false;
                };
                cv_work.wait(lock, [&] { return stop || has_work(); });
                if (!has_work()) return; // stop requested and nothing left
                for (auto& a : active)
                    if (a->next < a->pr.n && (!r || a->pr.n - a->next < r->pr.n - r->next)) r = a;
                const long long bs = r->opt.block_size;
                const long long span = std::max(bs, chunk / bs * bs);
                q0 = r->next;
                q1 = std::min(r->pr.n, q0 + span);
                r->next = q1;
                ++r->inflight;
                if (!r->started) {
                    r->started = true;
                    r->t_start = Clock::now();
                }
            }
            try {
                kumo::run_range(ex, r->mol, r->pr, r->ps, r->opt, r->dout.data(), q0, q1, none);
            } catch (const kumo::Error& e) {
                fail(r, e.status, e.what());
            } catch (const std::exception& e) {
                fail(r, KUMO_HIP_RUNTIME_ERROR, e.what());
            }
            bool last;
            {
                std::lock_guard<std::mutex> g(m);
                --r->inflight;
                last = r->next >= r->pr.n && r->inflight == 0;
            }
            if (!last) continue;
            if (!r->failed) {
                try {
                    for (std::size_t c = 0; c < r->pr.nout; ++c)
                        KUMO_HIP(hipMemcpyAsync(r->host_out[c], r->dout[c],
                                                static_cast<std::size_t>(r->pr.n) * sizeof(double),
                                                hipMemcpyDeviceToHost, ex.stream));
                    KUMO_HIP(hipStreamSynchronize(ex.stream));
                } catch (const kumo::Error& e) {
                    fail(r, e.status, e.what());
                }
            }
            std::vector<std::shared_ptr<Request>> admitted;
            {
                std::lock_guard<std::mutex> g(m);
                finish_locked(r);
                admitted = admit_locked();
            }
            cv_done.notify_all();
            if (!admitted.empty()) prepare(admitted, ex.stream);
            cv_work.notify_all();
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

int kumo_hip_server_create(kumo_hip_context* ctx, const kumo_hip_server_options* options, kumo_hip_server** out) {
    if (!ctx || !out) return KUMO_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(ctx, [&] {
        KUMO_HIP(hipSetDevice(ctx->device));
        auto s = std::make_unique<kumo_hip_server>();
        s->ctx = ctx;
        int nw = 4;
        if (options) {
            if (options->workers < 0 || options->workers > 16) kumo::invalid("server workers must be in [0, 16]");
            if (options->hbm_budget < 0 || options->chunk_points < 0) kumo::invalid("invalid server options");
            if (options->workers) nw = options->workers;
            if (options->hbm_budget) s->budget = static_cast<std::size_t>(options->hbm_budget);
            if (options->chunk_points) s->chunk = options->chunk_points;
        }
        for (int k = 0; k < nw; ++k) {
            auto e = std::make_unique<kumo::Executor>();
            e->create(ctx->device, ctx->boys_table.as<double>());
            s->execs.push_back(std::move(e));
        }
        for (int k = 0; k < nw; ++k)
            s->workers.emplace_back([srv = s.get(), k] { srv->worker(static_cast<std::size_t>(k)); });
        *out = s.release();
    });
}

int kumo_hip_server_submit(kumo_hip_server* s, const kumo_hip_molecule* mol, const kumo_hip_grid* grid,
                           const double* points, size_t npoints, const kumo_hip_property* properties,
                           size_t nproperties, const kumo_hip_options* options, double* const* outputs,
                           uint64_t* request) {
    if (!s || !mol || !request || (!grid && npoints && !points) || (nproperties && (!properties || !outputs)))
        return KUMO_HIP_INVALID_ARGUMENT;
    *request = 0;
    return guarded(s->ctx, [&] {
        auto r = std::make_shared<Request>();
        r->t_submit = Clock::now();
        r->mol = const_cast<kumo_hip_molecule*>(mol);
        kumo_hip_default_options(&r->opt);
        if (options) r->opt = *options;
        r->opt.progress = nullptr;
        r->opt.device_outputs = 1;
        kumo::PointsArg pts;
        if (grid) {
            r->is_grid = true;
            r->grid = *grid;
            pts.grid = &r->grid;
        } else {
            r->points.assign(points, points + 3 * npoints);
            pts.points = r->points.data();
            pts.n = static_cast<long long>(npoints);
        }
        {
            kumo_hip_options check = r->opt;
            check.device_outputs = 0; // validate listed points on the host copy
            r->pr = kumo::prepare(r->mol, pts, properties, nproperties, check, false);
        }
        r->host_out.assign(outputs, outputs + r->pr.nout);
        const std::size_t n = static_cast<std::size_t>(r->pr.n);
        r->bytes = std::max<std::size_t>(256, (r->pr.nout * n + (r->is_grid ? 0 : 3 * n)) * sizeof(double));
        std::vector<std::shared_ptr<Request>> admitted;
        {
            std::lock_guard<std::mutex> g(s->m);
            if (s->stop) kumo::runtime("server is shutting down");
            r->id = s->next_id++;
            s->requests[r->id] = r;
            s->pending.push_back(r);
            admitted = s->admit_locked();
        }
        *request = r->id;
        if (!admitted.empty()) {
            // Upload on the context's stream (submission is the caller's thread).
            std::lock_guard<std::mutex> g(s->ctx->mutex);
            s->prepare(admitted, s->ctx->exec.stream);
        }
    });
}

int kumo_hip_server_query(kumo_hip_server* s, uint64_t request, int32_t* done) {
    if (!s || !done) return KUMO_HIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->m);
    auto it = s->requests.find(request);
    if (it == s->requests.end()) return KUMO_HIP_INVALID_ARGUMENT;
    *done = it->second->done ? 1 : 0;
    return KUMO_HIP_OK;
}

int kumo_hip_server_wait(kumo_hip_server* s, uint64_t request, kumo_hip_request_info* info) {
    if (!s) return KUMO_HIP_INVALID_ARGUMENT;
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lock(s->m);
        auto it = s->requests.find(request);
        if (it == s->requests.end()) return KUMO_HIP_INVALID_ARGUMENT;
        r = it->second;
        s->cv_done.wait(lock, [&] { return r->done; });
        s->requests.erase(it);
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = r->status;
        info->backend_used = r->pr.backend;
        info->queue_ms = ms_between(r->t_submit, r->t_start);
        info->run_ms = ms_between(r->t_start, r->t_done);
        info->latency_ms = ms_between(r->t_submit, r->t_done);
        std::strncpy(info->error, r->error.c_str(), sizeof(info->error) - 1);
    }
    if (r->status != KUMO_HIP_OK) {
        std::lock_guard<std::mutex> g(s->ctx->mutex);
        s->ctx->last_error = r->error;
    }
    return r->status;
}

int kumo_hip_server_release(kumo_hip_server* s) {
    if (!s) return KUMO_HIP_OK;
    {
        std::unique_lock<std::mutex> lock(s->m);
        s->cv_done.wait(lock, [&] {
            for (auto& kv : s->requests)
                if (!kv.second->done) return false;
            return s->pending.empty() && s->preparing == 0;
        });
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers)
        t.join();
    for (auto& e : s->execs)
        e->destroy();
    s->pool.clear();
    delete s;
    return KUMO_HIP_OK;
}

} // extern "C"
