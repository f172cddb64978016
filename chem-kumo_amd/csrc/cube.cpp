// Gaussian Cube output of a field: Field::write_cube and
// Field::write_orbital_cube (src/cube.cpp:70-111, 253-262).
//
// The header (a few lines) is formatted here with snprintf in the formats of
// the reference's iostream settings (std::fixed, setprecision(6), setw(5) and
// setw(13), which libstdc++ writes through vsnprintf). The samples, millions
// of them, are formatted on the GPU where they lie (k_cube.hip) in chunks of
// kChunkValues: chunk c is formatted and copied into one pinned host buffer
// while the host writes chunk c - 1 from the other to the file.
#include "runtime.hpp"

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fcntl.h>
#include <string>
#include <unistd.h>

namespace kumo {

namespace {

// First header line of the reference's Cube files ("chem-kumo "
// CHEM_KUMO_VERSION, meson.build:4), so that the files are byte-identical.
constexpr const char* kProgramLine = "chem-kumo 2.1.0\n";
constexpr long long kChunkValues = 1LL << 21;

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

void put_all(int fd, const char* p, std::size_t n, const char* failure) {
    while (n) {
        const ssize_t w = ::write(fd, p, n);
        if (w < 0) {
            if (errno == EINTR) continue;
            runtime(failure);
        }
        p += w;
        n -= static_cast<std::size_t>(w);
    }
}

void append(std::string& s, const char* fmt, auto... args) {
    char buf[96];
    const int n = std::snprintf(buf, sizeof buf, fmt, args...);
    s.append(buf, static_cast<std::size_t>(n));
}

std::string cube_header(const char* comment, const kumo_hip_grid& g, const CubeAtoms& atoms,
                        std::size_t orbital_index) {
    std::string h = kProgramLine;
    h += comment;
    h += '\n';
    const long long natoms = atoms.natoms;
    append(h, "%5lld", orbital_index ? -natoms : natoms);
    for (int k = 0; k < 3; ++k) append(h, "%13.6f", g.origin[k]);
    h += '\n';
    for (int a = 0; a < 3; ++a) {
        append(h, "%5d", g.counts[a]);
        for (int k = 0; k < 3; ++k) append(h, "%13.6f", g.steps[3 * a + k]);
        h += '\n';
    }
    for (int i = 0; i < atoms.natoms; ++i) {
        append(h, "%5d", atoms.z[i]);
        append(h, "%13.6f", atoms.charges ? atoms.charges[i] : static_cast<double>(atoms.z[i]));
        for (int k = 0; k < 3; ++k) append(h, "%13.6f", atoms.xyz[3 * i + k]);
        h += '\n';
    }
    if (orbital_index) append(h, "%5d%5zu\n", 1, orbital_index);
    return h;
}

// Device copy of the samples (host samples are uploaded) after checking that
// every one is finite, as the reference does before it opens the file.
const double* finite_samples(kumo_hip_context* ctx, const double* values, long long n, bool on_device) {
    CubeBuffers& cb = ctx->cube;
    hipStream_t stream = ctx->exec.stream;
    const double* dv = values;
    if (!on_device && n) {
        cb.values.resize(static_cast<std::size_t>(n) * sizeof(double));
        KUMO_HIP(hipMemcpyAsync(cb.values.get(), values, static_cast<std::size_t>(n) * sizeof(double),
                                hipMemcpyHostToDevice, stream));
        dv = cb.values.as<double>();
    }
    cb.flag.resize(sizeof(int));
    launch_count_nonfinite(dv, n, cb.flag.as<int>(), stream);
    int bad = 0;
    KUMO_HIP(hipMemcpyAsync(&bad, cb.flag.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
    KUMO_HIP(hipStreamSynchronize(stream));
    if (bad) invalid("nonfinite Cube sample");
    return dv;
}

void ensure_host(CubeBuffers& cb, std::size_t bytes) {
    if (cb.host_bytes >= bytes) return;
    for (char*& h : cb.host) {
        if (h) (void)hipHostFree(h);
        h = nullptr;
    }
    cb.host_bytes = 0;
    for (char*& h : cb.host) KUMO_HIP(hipHostMalloc(reinterpret_cast<void**>(&h), bytes, hipHostMallocDefault));
    cb.host_bytes = bytes;
    for (int s = 0; s < 2; ++s)
        for (hipEvent_t* e : {&cb.start[s], &cb.kernel[s], &cb.done[s]})
            if (!*e) KUMO_HIP(hipEventCreate(e));
}

} // namespace

void CubeBuffers::release() {
    for (char*& h : host) {
        if (h) (void)hipHostFree(h);
        h = nullptr;
    }
    host_bytes = 0;
    for (hipEvent_t* e : {&start[0], &start[1], &kernel[0], &kernel[1], &done[0], &done[1]}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    values.release();
    text.release();
    flag.release();
}

void write_field_cube(kumo_hip_context* ctx, const char* path, const char* comment, const kumo_hip_grid& grid,
                      const CubeAtoms& atoms, const double* values, bool values_on_device, std::size_t orbital_index) {
    const auto t0 = std::chrono::steady_clock::now();
    const bool orbital = orbital_index != 0;
    if (!path) invalid("null Cube path");
    const long long n = validated_point_count(grid);
    if (atoms.natoms < 0) invalid("negative atom count");
    if (atoms.natoms > 0 && (!atoms.z || !atoms.xyz)) invalid("null atom arrays");
    if (!values) invalid("null Cube samples");
    if (!comment) comment = orbital ? "Molecular orbital amplitude" : "Electron density";
    const char* open_failure = orbital ? "cannot open orbital cube output" : "cannot open cube output";
    const char* write_failure = orbital ? "orbital cube write failed" : "cube write failed";

    std::lock_guard<std::mutex> lock(ctx->mutex);
    KUMO_HIP(hipSetDevice(ctx->device));
    CubeBuffers& cb = ctx->cube;
    hipStream_t stream = ctx->exec.stream;
    const double* dv = finite_samples(ctx, values, n, values_on_device);

    Fd file;
    file.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0666);
    if (file.fd < 0) runtime(open_failure);
    const std::string header = cube_header(comment, grid, atoms, orbital_index);
    put_all(file.fd, header.data(), header.size(), write_failure);

    const long long nz = grid.counts[2];
    const std::size_t chunk_bytes = static_cast<std::size_t>(14 * std::min(n, kChunkValues)) + 64;
    cb.text.resize(chunk_bytes);
    ensure_host(cb, chunk_bytes);
    const long long nchunks = (n + kChunkValues - 1) / kChunkValues;
    std::size_t bytes[2] = {0, 0};
    double kernel_ms = 0.0, copy_ms = 0.0;
    long long total = 0;
    // Waits for chunk c's copy and writes it to the file.
    auto drain = [&](long long c) {
        const int s = static_cast<int>(c & 1);
        KUMO_HIP(hipEventSynchronize(cb.done[s]));
        float k = 0.0f, m = 0.0f;
        KUMO_HIP(hipEventElapsedTime(&k, cb.start[s], cb.kernel[s]));
        KUMO_HIP(hipEventElapsedTime(&m, cb.kernel[s], cb.done[s]));
        kernel_ms += k;
        copy_ms += m;
        put_all(file.fd, cb.host[s], bytes[s], write_failure);
        total += static_cast<long long>(bytes[s]);
    };
    try {
        for (long long c = 0; c < nchunks; ++c) {
            const int s = static_cast<int>(c & 1);
            const long long first = c * kChunkValues;
            const long long count = std::min(kChunkValues, n - first);
            const long long t_first = cube_text_offset(first, nz);
            bytes[s] = static_cast<std::size_t>(cube_text_offset(first + count, nz) - t_first);
            KUMO_HIP(hipEventRecord(cb.start[s], stream));
            launch_cube_text(dv, first, count, nz, t_first, cb.text.as<unsigned char>(), stream);
            KUMO_HIP(hipEventRecord(cb.kernel[s], stream));
            KUMO_HIP(hipMemcpyAsync(cb.host[s], cb.text.get(), bytes[s], hipMemcpyDeviceToHost, stream));
            KUMO_HIP(hipEventRecord(cb.done[s], stream));
            if (c > 0) drain(c - 1);
        }
        if (nchunks > 0) drain(nchunks - 1);
    } catch (...) {
        (void)hipStreamSynchronize(stream); // no copy into the pinned buffers outlives the call
        throw;
    }
    if (::close(file.fd) != 0) {
        file.fd = -1;
        runtime(write_failure);
    }
    file.fd = -1;
    cb.last_kernel_ms = kernel_ms;
    cb.last_copy_ms = copy_ms;
    cb.last_values = n;
    cb.last_bytes = static_cast<long long>(header.size()) + total;
    cb.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void format_scientific(kumo_hip_context* ctx, const double* values, std::size_t n, int precision, bool exact_only,
                       char* text) {
    if (precision < 0 || precision > 12) invalid("precision must be in [0, 12]");
    if (n && (!values || !text)) invalid("null arrays");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    KUMO_HIP(hipSetDevice(ctx->device));
    hipStream_t stream = ctx->exec.stream;
    const double* dv = finite_samples(ctx, values, static_cast<long long>(n), false);
    if (!n) return;
    const std::size_t bytes = n * static_cast<std::size_t>(precision + 8);
    DeviceBuffer out(bytes);
    launch_sci_slots(dv, static_cast<long long>(n), precision, exact_only ? 1 : 0, out.as<char>(), stream);
    KUMO_HIP(hipMemcpyAsync(text, out.get(), bytes, hipMemcpyDeviceToHost, stream));
    KUMO_HIP(hipStreamSynchronize(stream));
}

} // namespace kumo
