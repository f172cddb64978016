// extern "C" entry points of include/kumo_hip.h. Every call catches all C++
// exceptions and maps them to a status code; the message is kept on the
// context for kumo_hip_last_error().
#include "runtime.hpp"

#include <cstring>
#include <new>

namespace {

template <typename F> int guarded(kumo_hip_context* ctx, F&& f) {
    try {
        f();
        return KUMO_HIP_OK;
    } catch (const kumo::Error& e) {
        if (ctx) ctx->last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host memory exhausted";
        return KUMO_HIP_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return KUMO_HIP_RUNTIME_ERROR;
    }
}

} // namespace

extern "C" {

void kumo_hip_default_options(kumo_hip_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->threads = 0;
    o->block_size = 128;
    o->backend = KUMO_HIP_AUTO;
    o->precision = KUMO_HIP_FP64;
    o->screening_threshold = 1e-12;
    o->potential_tolerance = 1e-8;
    o->fukui_delta = 0.1;
}

int kumo_hip_init(int device, kumo_hip_context** out) {
    if (!out) return KUMO_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    auto* ctx = new (std::nothrow) kumo_hip_context();
    if (!ctx) return KUMO_HIP_RUNTIME_ERROR;
    const int rc = guarded(ctx, [&] {
        int count = 0;
        KUMO_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) kumo::invalid("no such HIP device");
        KUMO_HIP(hipSetDevice(device));
        ctx->device = device;
        ctx->exec.create(device, nullptr);
        kumo::noalloc_scratch(device); // the allocation-free entry point's scratch on this device
        const auto table = kumo::boys_device_table();
        ctx->boys_table = kumo::upload(table.data(), table.size(), ctx->exec.stream);
        KUMO_HIP(hipStreamSynchronize(ctx->exec.stream));
        ctx->exec.boys_table = ctx->boys_table.as<double>();
    });
    if (rc != KUMO_HIP_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return KUMO_HIP_OK;
}

int kumo_hip_release(kumo_hip_context* ctx) {
    if (!ctx) return KUMO_HIP_OK;
    ctx->exec.destroy();
    ctx->cube.release();
    ctx->ws_out.release();
    ctx->boys_table.release();
    delete ctx;
    return KUMO_HIP_OK;
}

int kumo_hip_last_error(const kumo_hip_context* ctx, char* buffer, size_t length) {
    if (!ctx || !buffer || length == 0) return KUMO_HIP_INVALID_ARGUMENT;
    std::strncpy(buffer, ctx->last_error.c_str(), length - 1);
    buffer[length - 1] = '\0';
    return KUMO_HIP_OK;
}

int kumo_hip_molecule_from_fchk(kumo_hip_context* ctx, const char* path, kumo_hip_molecule** mol) {
    if (!ctx || !path || !mol) return KUMO_HIP_INVALID_ARGUMENT;
    *mol = nullptr;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        KUMO_HIP(hipSetDevice(ctx->device));
        *mol = kumo::load_fchk(ctx, path).release();
    });
}

int kumo_hip_molecule_create(kumo_hip_context* ctx, const kumo_hip_molecule_desc* desc, kumo_hip_molecule** mol) {
    if (!ctx || !desc || !mol) return KUMO_HIP_INVALID_ARGUMENT;
    *mol = nullptr;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        KUMO_HIP(hipSetDevice(ctx->device));
        *mol = kumo::load_molecule(ctx, *desc).release();
    });
}

int kumo_hip_molecule_info_get(const kumo_hip_molecule* m, kumo_hip_molecule_info* info) {
    if (!m || !info) return KUMO_HIP_INVALID_ARGUMENT;
    info->natoms = static_cast<int32_t>(m->atoms.size());
    info->nbasis = m->n;
    info->nblocks = static_cast<int32_t>(m->basis.blocks.size());
    info->max_l = m->basis.max_l;
    info->nalpha = m->nalpha;
    info->nbeta = m->nbeta;
    info->nmo_alpha = m->nmo_a;
    info->nmo_beta = m->nmo_b;
    info->unrestricted = m->unrestricted ? 1 : 0;
    info->ao_mo_eligible = m->eligible ? 1 : 0;
    info->total_relative_error = m->total_rel;
    info->spin_relative_error = m->spin_rel;
    info->device_bytes = m->device_bytes;
    return KUMO_HIP_OK;
}

int kumo_hip_molecule_release(kumo_hip_molecule* m) {
    if (!m) return KUMO_HIP_OK;
    if (m->ctx) (void)hipStreamSynchronize(m->ctx->exec.stream);
    delete m;
    return KUMO_HIP_OK;
}

int kumo_hip_molecule_blocks(const kumo_hip_molecule* m, int32_t* l, int32_t* offset, int32_t* count, double* cutoff_r2) {
    if (!m) return KUMO_HIP_INVALID_ARGUMENT;
    for (std::size_t i = 0; i < m->basis.blocks.size(); ++i) {
        const auto& b = m->basis.blocks[i];
        if (l) l[i] = b.l;
        if (offset) offset[i] = b.offset;
        if (count) count[i] = b.count;
        if (cutoff_r2) cutoff_r2[i] = b.cutoff_r2;
    }
    return KUMO_HIP_OK;
}

int kumo_hip_molecule_atoms(const kumo_hip_molecule* m, int32_t* z, double* charges, double* xyz) {
    if (!m) return KUMO_HIP_INVALID_ARGUMENT;
    for (std::size_t a = 0; a < m->atoms.size(); ++a) {
        if (z) z[a] = m->atoms[a].z;
        if (charges) charges[a] = m->atoms[a].charge;
        if (xyz)
            for (int k = 0; k < 3; ++k)
                xyz[3 * a + k] = m->atoms[a].xyz[k];
    }
    return KUMO_HIP_OK;
}

int kumo_hip_evaluate_grid(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const kumo_hip_grid* grid,
                           const kumo_hip_property* properties, size_t nproperties, const kumo_hip_options* options,
                           double* const* outputs, int32_t* backend_used) {
    if (!ctx || !mol || !grid || (nproperties && (!properties || !outputs))) return KUMO_HIP_INVALID_ARGUMENT;
    kumo_hip_options def;
    kumo_hip_default_options(&def);
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        KUMO_HIP(hipSetDevice(ctx->device));
        kumo::PointsArg pts;
        pts.grid = grid;
        kumo::evaluate(ctx, mol, pts, properties, nproperties, options ? *options : def, outputs, backend_used);
    });
}

int kumo_hip_evaluate_points(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points, size_t npoints,
                             const kumo_hip_property* properties, size_t nproperties, const kumo_hip_options* options,
                             double* const* outputs, int32_t* backend_used) {
    if (!ctx || !mol || (npoints && !points) || (nproperties && (!properties || !outputs)))
        return KUMO_HIP_INVALID_ARGUMENT;
    kumo_hip_options def;
    kumo_hip_default_options(&def);
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        KUMO_HIP(hipSetDevice(ctx->device));
        kumo::PointsArg pts;
        pts.points = points;
        pts.n = static_cast<long long>(npoints);
        kumo::evaluate(ctx, mol, pts, properties, nproperties, options ? *options : def, outputs, backend_used);
    });
}

int kumo_hip_ao_eval(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points, size_t npoints,
                     int32_t order, double* out) {
    if (!ctx || !mol || (npoints && (!points || !out))) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        if (order < 0 || order > 2) kumo::invalid("AO derivative order must be 0, 1 or 2");
        const long long n = static_cast<long long>(npoints);
        if (n == 0) return;
        const int panels = order == 0 ? 1 : (order == 1 ? 4 : 10);
        kumo::DeviceBuffer dp = kumo::upload(points, 3 * static_cast<std::size_t>(n), ctx->exec.stream);
        kumo::DeviceBuffer dout(static_cast<std::size_t>(mol->n) * panels * n * sizeof(double));
        kumo::PointSource ps{};
        ps.ny = ps.nz = 1;
        ps.points = dp.as<double>();
        kumo::launch_ao_eval(mol->view(), ps, n, order, dout.as<double>(), ctx->exec.stream);
        KUMO_HIP(hipMemcpyAsync(out, dout.get(), dout.bytes(), hipMemcpyDeviceToHost, ctx->exec.stream));
        KUMO_HIP(hipStreamSynchronize(ctx->exec.stream));
    });
}

int kumo_hip_screening_masks(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const kumo_hip_grid* grid,
                             const double* points, size_t npoints, int32_t block_size, int32_t ao_order,
                             double threshold, uint8_t* kept) {
    if (!ctx || !mol || !kept || (!grid && npoints && !points)) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        kumo::PointsArg pts;
        pts.grid = grid;
        pts.points = points;
        pts.n = static_cast<long long>(npoints);
        kumo::screening_masks(ctx, mol, pts, block_size, ao_order, threshold, kept);
    });
}

int kumo_hip_boys(kumo_hip_context* ctx, const double* T, size_t n, int32_t m, double* out) {
    if (!ctx || (n && (!T || !out))) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        if (m < 0 || m > 32) kumo::invalid("unsupported Boys function order");
        for (size_t i = 0; i < n; ++i)
            if (!(T[i] >= 0.0) || !std::isfinite(T[i]))
                kumo::invalid("Boys function argument must be finite and nonnegative");
        if (n == 0) return;
        kumo::DeviceBuffer dt = kumo::upload(T, n, ctx->exec.stream);
        kumo::DeviceBuffer dout(n * static_cast<std::size_t>(m + 1) * sizeof(double));
        kumo::launch_boys(ctx->boys_table.as<double>(), dt.as<double>(), static_cast<long long>(n), m, dout.as<double>(),
                          ctx->exec.stream);
        KUMO_HIP(hipMemcpyAsync(out, dout.get(), dout.bytes(), hipMemcpyDeviceToHost, ctx->exec.stream));
        KUMO_HIP(hipStreamSynchronize(ctx->exec.stream));
    });
}

int kumo_hip_boys_esp(kumo_hip_context* ctx, const double* T, size_t n, int32_t top, double* out) {
    if (!ctx || (n && (!T || !out))) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        if (top < 0 || top > 2 * kumo::kMaxL) kumo::invalid("the pair-loop Boys function has orders 0 to 10");
        for (size_t i = 0; i < n; ++i)
            if (!(T[i] >= 0.0) || !std::isfinite(T[i]))
                kumo::invalid("Boys function argument must be finite and nonnegative");
        if (n == 0) return;
        kumo::DeviceBuffer dt = kumo::upload(T, n, ctx->exec.stream);
        kumo::DeviceBuffer dout(n * static_cast<std::size_t>(top + 1) * sizeof(double));
        kumo::launch_boys_esp(ctx->boys_table.as<double>(), dt.as<double>(), static_cast<long long>(n), top,
                              dout.as<double>(), ctx->exec.stream);
        KUMO_HIP(hipMemcpyAsync(out, dout.get(), dout.bytes(), hipMemcpyDeviceToHost, ctx->exec.stream));
        KUMO_HIP(hipStreamSynchronize(ctx->exec.stream));
    });
}

int kumo_hip_last_kernel_stats(const kumo_hip_context* ctx, double* kernel_ms, int64_t* launches, double* flops,
                               double* padded_flops) {
    if (!ctx) return KUMO_HIP_INVALID_ARGUMENT;
    if (kernel_ms) *kernel_ms = ctx->exec.last_kernel_ms;
    if (launches) *launches = ctx->exec.last_launches;
    if (flops) *flops = ctx->exec.last_flops;
    if (padded_flops) *padded_flops = ctx->exec.last_flops;
    return KUMO_HIP_OK;
}

int kumo_hip_last_field_stats(const kumo_hip_context* ctx, double* fields_ms, int64_t* points, int64_t* bytes) {
    if (!ctx) return KUMO_HIP_INVALID_ARGUMENT;
    if (fields_ms) *fields_ms = ctx->exec.last_fields_ms;
    if (points) *points = ctx->exec.last_fields_points;
    if (bytes) *bytes = ctx->exec.last_fields_bytes;
    return KUMO_HIP_OK;
}

int kumo_hip_write_cube(kumo_hip_context* ctx, const char* path, const char* comment, const kumo_hip_grid* grid,
                        int32_t natoms, const int32_t* atomic_numbers, const double* nuclear_charges,
                        const double* coordinates, const double* values, int32_t values_on_device) {
    if (!ctx || !grid) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        kumo::write_field_cube(ctx, path, comment, *grid, {natoms, atomic_numbers, nuclear_charges, coordinates},
                               values, values_on_device != 0, 0);
    });
}

int kumo_hip_write_orbital_cube(kumo_hip_context* ctx, const char* path, const char* comment,
                                const kumo_hip_grid* grid, int32_t natoms, const int32_t* atomic_numbers,
                                const double* nuclear_charges, const double* coordinates, const double* values,
                                int32_t values_on_device, size_t orbital_index) {
    if (!ctx || !grid) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        // Field::write_orbital_cube (src/cube.cpp:257-262).
        if (orbital_index == 0 || orbital_index > 9999) kumo::invalid("invalid orbital index for Cube");
        kumo::write_field_cube(ctx, path, comment, *grid, {natoms, atomic_numbers, nuclear_charges, coordinates},
                               values, values_on_device != 0, orbital_index);
    });
}

int kumo_hip_format_scientific(kumo_hip_context* ctx, const double* values, size_t n, int32_t precision,
                               int32_t exact_only, char* text) {
    if (!ctx) return KUMO_HIP_INVALID_ARGUMENT;
    return guarded(ctx, [&] { kumo::format_scientific(ctx, values, n, precision, exact_only != 0, text); });
}

int kumo_hip_last_cube_stats(const kumo_hip_context* ctx, double* total_ms, double* kernel_ms, double* copy_ms,
                             int64_t* values, int64_t* bytes) {
    if (!ctx) return KUMO_HIP_INVALID_ARGUMENT;
    if (total_ms) *total_ms = ctx->cube.last_ms;
    if (kernel_ms) *kernel_ms = ctx->cube.last_kernel_ms;
    if (copy_ms) *copy_ms = ctx->cube.last_copy_ms;
    if (values) *values = ctx->cube.last_values;
    if (bytes) *bytes = ctx->cube.last_bytes;
    return KUMO_HIP_OK;
}

int kumo_hip_last_rescue_stats(const kumo_hip_context* ctx, double* rescue_ms, int64_t* rescue_points,
                               int64_t* rescue_flagged) {
    if (!ctx) return KUMO_HIP_INVALID_ARGUMENT;
    if (rescue_ms) *rescue_ms = ctx->exec.last_rescue_ms;
    if (rescue_points) *rescue_points = ctx->exec.last_rescue_points;
    if (rescue_flagged) *rescue_flagged = ctx->exec.last_rescue_flagged;
    return KUMO_HIP_OK;
}

} // extern "C"
