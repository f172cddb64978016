// Context (one device), executors (a stream with its reusable workspaces) and
// molecule residency: the basis tables, density matrices and MO coefficients
// stay in HBM for the molecule's lifetime, so repeated evaluations upload
// nothing but their point list.
#pragma once

#include "basis.hpp"
#include "common.hpp"
#include "far_field.hpp"
#include "kernels.hpp"
#include "plan.hpp"

#include "../../include/kumo_hip.h"

#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace kumo {

// A stream and the workspaces of the evaluations that run on it. One
// evaluation (or evaluation range) at a time per executor.
struct Executor {
    hipStream_t stream = nullptr;
    const double* boys_table = nullptr;
    DeviceBuffer ws_tables, ws_moments, ws_phi, ws_w, ws_points, ws_tasks, ws_misc, ws_rescue, ws_rmom, ws_rscale,
        ws_oe, ws_oe_tasks, ws_esp, ws_far;
    // Events bracketing the contraction launches (index 0), the FP64 rescue
    // launches (index 1) and the field reductions (index 2).
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events[3];
    std::size_t events_used[3] = {0, 0, 0};
    double last_kernel_ms = 0.0, last_flops = 0.0, last_rescue_ms = 0.0, last_fields_ms = 0.0;
    long long last_launches = 0, last_rescue_points = 0, last_rescue_flagged = 0, last_fields_points = 0,
              last_fields_bytes = 0;
    // Pinned host memory for per-batch statistics copied back without
    // stopping the stream (kept AO counts, rescue tiles).
    int32_t* hstats = nullptr;
    std::size_t hstats_count = 0;
    int32_t* host_stats(std::size_t count);
    // Capture mode (a batch recorded into a HIP graph): no host
    // synchronisation, no timing events, no readback; small host arrays
    // copied to the device are kept in `keep` so that graph replays read them.
    bool capture = false;
    std::deque<std::vector<char>>* keep = nullptr;
    // Timing events even in capture mode (an eager issue with no host
    // synchronisation, whose launches overlap across streams as in a graph).
    bool events_in_capture = false;
    // Allocation-free mode: every workspace is a piece of a fixed arena,
    // batches hold at most max_batch_blocks screening blocks, and small host
    // tables reach the device as kernel arguments (capture mode is implied).
    bool no_alloc = false;
    long long max_batch_blocks = 0;
    // Exact ESP sums only (batches: the same values eager and graph-captured).
    bool exact_esp = false;
    void use_arena(Arena* arena) {
        for (DeviceBuffer* b : {&ws_tables, &ws_moments, &ws_phi, &ws_w, &ws_points, &ws_tasks, &ws_misc, &ws_rescue,
                                &ws_rmom, &ws_rscale, &ws_oe, &ws_oe_tasks, &ws_esp, &ws_far})
            b->use_arena(arena);
    }
    void create(int device, const double* table);
    void destroy();
    void reset_stats();
    void collect_timing();
};

struct AtomRec {
    int z;
    double charge;
    double xyz[3];
};

// One MO slot of the orbital-energy operand table.
struct OeSlot {
    int kind;          // 0 padding, 1 occupied, 2 virtual
    int channel;       // 0 alpha, 1 beta
    int mo;            // zero-based MO index in its channel
    double occupation; // n_i of an occupied MO (2 doubly, 1 singly occupied)
};

} // namespace kumo

namespace kumo {
// Buffers of the Cube writer: the device copy of host samples, the text of
// one chunk in HBM and two pinned host buffers (one being written to the
// file while the next chunk is formatted and copied).
struct CubeBuffers {
    DeviceBuffer values, text, flag;
    char* host[2] = {nullptr, nullptr};
    std::size_t host_bytes = 0;
    // Per host buffer: before the text kernel, after it, after the copy.
    hipEvent_t start[2] = {nullptr, nullptr};
    hipEvent_t kernel[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    double last_ms = 0.0, last_kernel_ms = 0.0, last_copy_ms = 0.0;
    long long last_bytes = 0, last_values = 0;
    void release();
};
} // namespace kumo

struct kumo_hip_context {
    int device = 0;
    std::string last_error;
    std::mutex mutex; // one evaluation at a time on the default executor
    kumo::DeviceBuffer boys_table;
    kumo::Executor exec;
    kumo::DeviceBuffer ws_out;
    kumo::CubeBuffers cube;
};

struct kumo_hip_molecule {
    kumo_hip_context* ctx = nullptr;
    std::vector<kumo::AtomRec> atoms;
    kumo::PreparedBasis basis;
    int n = 0;
    std::vector<double> P, S;   // row-major n x n
    std::vector<double> Ca, Cb; // n x nmo (AO rows)
    std::vector<double> Ea, Eb;
    int nmo_a = 0, nmo_b = 0, nalpha = 0, nbeta = 0;
    bool unrestricted = false, eligible = false;
    double total_rel = 0.0, spin_rel = 0.0;
    // Device residency
    kumo::DeviceBuffer d_shells, d_exps, d_weights, d_tdense, d_radw, d_tg, d_gk, d_lorder, d_rec;
    // Per screening threshold: squared radius per block beyond which no AO of
    // the block passes it (kept until the molecule is released).
    std::map<double, kumo::DeviceBuffer> d_screen_r2;
    // Autotuned AO->MO tile (points per workgroup) per (precision, order, Hessian).
    std::mutex tune_mutex;
    std::map<int, int> tuned_tile;
    kumo::DeviceBuffer d_p[4]; // total, spin, alpha, beta (alpha/beta on first use)
    kumo::DeviceBuffer d_ca, d_cb; // full MO matrices
    kumo::DeviceBuffer d_op64, d_op32, d_opbf, d_opnorm; // MO operand table (FP64, FP32, 3 bf16 planes)
    int slots = 0, alpha_slots = 0, operand_rows = 0;
    int max_chunks = 0; // contraction chunks of any screening block (bound)
    bool esp_ready = false;
    kumo::DeviceBuffer d_esp_pairs, d_esp_coef, d_esp_atoms;
    int esp_npairs = 0;
    std::vector<kumo::EspGroup> esp_groups; // pair ranges by L
    std::vector<kumo::EspPairDev> esp_pairs_host; // the pairs and coefficients, for far-field plans
    std::vector<double> esp_coef_host;
    // esp_ref_order[k]: the pair that is the k-th in the reference's order
    // (block pairs a <= b, then primitive pairs), which the far field's
    // pair budget follows.
    std::vector<int32_t> esp_ref_order;
    // Far-field plans per (target radius, tolerance), resident until the
    // molecule is released.
    struct FarPlanResident {
        kumo::DeviceBuffer level_d2, clusters, cluster_pairs, cluster_of, moments, ref_order;
        kumo::FarPlanDev dev{};
    };
    std::map<std::pair<double, double>, FarPlanResident> far_plans;
    // Orbital-energy operand (OrbitalEnergyFields, src/orbital_energy_fields.hpp:98-192;
    // built on first use): every MO of the FCHK in FP64, slots [0,
    // oe_occ_slots) the occupied MOs, then the virtual ones, each range
    // padded to 16.
    bool oe_ready = false;
    kumo::DeviceBuffer d_oe_op;
    int oe_slots = 0, oe_occ_slots = 0;
    std::vector<kumo::OeSlot> oe_slot;
    long long device_bytes = 0;
    std::mutex lazy; // guards the data built on first use

    kumo::MolView view() const {
        return {d_shells.as<kumo::DevShell>(), static_cast<int>(basis.shells.size()), n, d_exps.as<double>(),
                d_weights.as<double>(), d_tdense.as<double>(), d_radw.as<double>(), d_tg.as<double>(),
                d_gk.as<double>(), d_lorder.as<int32_t>(), d_rec.as<kumo::ShellRec>()};
    }
    const double* p_matrix(int component, hipStream_t stream);
    const double* screen_r2(double threshold, hipStream_t stream);
    // The squared screening radii for `threshold` if they are resident, else null.
    const double* screen_r2_resident(double threshold);
    void ensure_esp(hipStream_t stream);
    // The far-field plan for target clusters of `radius` bohr (enlarged for
    // rounding as the reference's cluster_plan does) and `tolerance` > 0.
    const kumo::FarPlanDev& far_plan(double radius, double tolerance, hipStream_t stream);
    void ensure_orbital_energy(hipStream_t stream);
};

namespace kumo {

// Device address of the allocation-free entry point's scratch array on
// `device`, looked up (hipGetSymbolAddress) on the first call per device.
void* noalloc_scratch(int device);

std::unique_ptr<kumo_hip_molecule> load_molecule(kumo_hip_context* ctx, const kumo_hip_molecule_desc& d);
std::unique_ptr<kumo_hip_molecule> load_fchk(kumo_hip_context* ctx, const std::string& path);
std::vector<double> boys_table_host();
std::vector<double> boys_device_table();

struct PointsArg {
    const kumo_hip_grid* grid = nullptr;
    const double* points = nullptr; // host or device
    long long n = 0;
};

// A validated request: plan, backend, output layout and pass structure.
struct Prepared {
    Plan plan;
    int backend = KUMO_HIP_AO_P;
    bool grid = false;
    long long n = 0;
    std::vector<kumo_hip_property> props;
    std::vector<std::size_t> first_channel;
    std::size_t nout = 0;
    std::vector<int> needed;
    bool orbital_pass = false;
    int grid_density_passes = 0, grid_orbital_passes = 0, passes = 0;
};

using Advance = std::function<void(std::size_t)>;

long long validated_point_count(const kumo_hip_grid& g);
PointSource make_source(const kumo_hip_grid* grid, const double* device_points);
Prepared prepare(const kumo_hip_molecule* mol, const PointsArg& pts, const kumo_hip_property* props,
                 std::size_t nprops, const kumo_hip_options& opt, bool points_on_device);
// Evaluates points [q0, q1) (q0 a multiple of the block size) of a prepared
// request into the device outputs dout[channel][point] (full-length arrays).
void run_range(Executor& ex, kumo_hip_molecule* mol, const Prepared& pr, const PointSource& ps,
               const kumo_hip_options& opt, double* const* dout, long long q0, long long q1, const Advance& advance);
// Device bytes the point pass of run_range takes from its workspaces: `fixed`
// whatever the batch, plus `per_block` for each screening block of a batch.
void point_pass_bytes(const kumo_hip_molecule* mol, const Prepared& pr, const kumo_hip_options& opt,
                      std::size_t& fixed, std::size_t& per_block);

void evaluate(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const PointsArg& pts,
              const kumo_hip_property* props, std::size_t nprops, const kumo_hip_options& opt, double* const* outputs,
              int32_t* backend_used);

void screening_masks(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const PointsArg& pts, int block_size,
                     int order, double threshold, uint8_t* kept);

// Field::write_cube / Field::write_orbital_cube (src/cube.cpp:70-111, 253-262);
// orbital_index 0 writes a density Cube.
struct CubeAtoms {
    int natoms = 0;
    const int32_t* z = nullptr;
    const double* charges = nullptr;
    const double* xyz = nullptr;
};
void write_field_cube(kumo_hip_context* ctx, const char* path, const char* comment, const kumo_hip_grid& grid,
                      const CubeAtoms& atoms, const double* values, bool values_on_device, std::size_t orbital_index);
// printf("%*.*e", precision + 8, precision, v) of host values into host text.
void format_scientific(kumo_hip_context* ctx, const double* values, std::size_t n, int precision, bool exact_only,
                       char* text);

} // namespace kumo
