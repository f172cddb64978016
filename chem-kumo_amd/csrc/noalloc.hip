// The allocation-free evaluation entry point (kumo_hip_evaluate_points_noalloc).
//
// Callers that cannot allocate device memory (a HIP graph being captured, a
// serving loop that must not call hipMalloc) evaluate point-path properties
// into their own device buffers on their own stream. Every workspace of the
// point pass (screening tables, channel moments, error scales, rescue plan and
// moments, field tasks) is a piece of one file-scope device array, `scratch`,
// of kNoAllocScratchBytes; evaluations larger than one batch of it are cut
// into batches of whole screening blocks (results do not depend on the batch
// size). Each GPU holds its own copy of the array from the first use of the
// library on it until the process ends; calls on one GPU share it, so calls
// that overlap on one GPU (several host threads, or several streams) race.
// No workspace needs to start at zero: every kernel of the pass writes what
// it later reads in the same call.
#include "runtime.hpp"

#include <array>
#include <mutex>

namespace {

constexpr std::size_t kNoAllocScratchBytes = std::size_t{64} << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
std::array<void*, kMaxDevices> g_scratch{}; // device address of `scratch` per device

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

namespace kumo {

// The device address of `scratch` on `device` (current), looked up once per
// device. The array's name in host code is only a handle for this lookup.
void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) invalid("device index beyond the allocation-free entry point's table");
    std::lock_guard<std::mutex> lock(g_mutex);
    void*& ptr = g_scratch[static_cast<std::size_t>(device)];
    if (!ptr) {
        void* p = nullptr;
        KUMO_HIP(hipSetDevice(device));
        KUMO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) runtime("allocation-free scratch has no device address");
        ptr = p;
    }
    return ptr;
}

} // namespace kumo

extern "C" {

int kumo_hip_evaluate_points_noalloc(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points,
                                     size_t npoints, const kumo_hip_property* properties, size_t nproperties,
                                     const kumo_hip_options* options, double* const* outputs, void* stream,
                                     int32_t* backend_used) {
    if (!ctx || !mol || (npoints && (!points || !outputs)) || (nproperties && !properties))
        return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        using namespace kumo;
        void* base = noalloc_scratch(ctx->device);
        KUMO_HIP(hipSetDevice(ctx->device));
        kumo_hip_options opt;
        kumo_hip_default_options(&opt);
        if (options) opt = *options;
        opt.progress = nullptr;
        opt.device_outputs = 1;
        auto* m = const_cast<kumo_hip_molecule*>(mol);
        PointsArg pts;
        pts.points = points;
        pts.n = static_cast<long long>(npoints);
        const Prepared pr = prepare(m, pts, properties, nproperties, opt, true);
        if (pr.orbital_pass || pr.plan.esp() || pr.plan.orbital_energies || pr.grid_density_passes ||
            pr.grid_orbital_passes)
            invalid("the allocation-free entry point evaluates density and derived-field properties only "
                    "(no MO amplitudes, electrostatic potential or orbital-energy fields)");
        if (backend_used) *backend_used = pr.backend;
        if (pr.n == 0 || pr.needed.empty()) return;
        std::size_t fixed = 0, per_block = 0;
        point_pass_bytes(m, pr, opt, fixed, per_block);
        if (fixed + per_block > kNoAllocScratchBytes)
            invalid("one screening block of this molecule needs " + std::to_string(fixed + per_block) +
                    " bytes, more than the " + std::to_string(kNoAllocScratchBytes >> 20) +
                    " MiB scratch of the allocation-free entry point");
        Executor ex;
        ex.stream = stream ? static_cast<hipStream_t>(stream) : ctx->exec.stream;
        ex.boys_table = ctx->boys_table.as<double>();
        ex.capture = true; // no host synchronisation, events or readback
        ex.no_alloc = true;
        ex.max_batch_blocks = static_cast<long long>((kNoAllocScratchBytes - fixed) / per_block);
        Arena arena{static_cast<char*>(base), kNoAllocScratchBytes, 0};
        ex.use_arena(&arena);
        const PointSource ps = make_source(nullptr, points);
        const Advance none = [](std::size_t) {};
        run_range(ex, m, pr, ps, opt, outputs, 0, pr.n, none);
        // the stream is the caller's (or the context's): not destroyed here
        ex.stream = nullptr;
    });
}

} // extern "C"
