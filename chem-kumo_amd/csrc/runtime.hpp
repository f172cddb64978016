// Context (one device, one stream, reusable workspaces) and molecule
// residency: the basis tables, density matrices and MO coefficients stay in
// HBM for the molecule's lifetime, so repeated evaluations upload nothing but
// their point list.
#pragma once

#include "basis.hpp"
#include "common.hpp"
#include "kernels.hpp"

#include "../../include/kumo_hip.h"

#include <memory>
#include <mutex>
#include <string>
#include <vector>

struct kumo_hip_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    std::mutex mutex; // one evaluation at a time per context
    kumo::DeviceBuffer boys_table;
    // Workspaces, grown on demand.
    kumo::DeviceBuffer ws_tables, ws_moments, ws_comp, ws_phi, ws_w, ws_out, ws_points, ws_tasks, ws_misc;
    // Timing of the dominant contraction kernel of the last evaluation.
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::size_t events_used = 0;
    double last_kernel_ms = 0.0;
    long long last_launches = 0;
    double last_flops = 0.0, last_padded_flops = 0.0;
};

namespace kumo {

struct AtomRec {
    int z;
    double charge;
    double xyz[3];
};

} // namespace kumo

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
    kumo::DeviceBuffer d_shells, d_exps, d_weights, d_tdense;
    kumo::DeviceBuffer d_p[4]; // total, spin, alpha, beta (alpha/beta on first use)
    kumo::DeviceBuffer d_ca, d_cb; // full MO matrices
    kumo::DeviceBuffer d_mo64, d_mo32, d_mobf;
    int slots = 0, alpha_slots = 0;
    bool esp_ready = false;
    kumo::DeviceBuffer d_esp_pairs, d_esp_coef, d_esp_atoms;
    int esp_npairs = 0;
    long long device_bytes = 0;

    kumo::MolView view() const {
        return {d_shells.as<kumo::DevShell>(), static_cast<int>(basis.shells.size()), n, d_exps.as<double>(),
                d_weights.as<double>(), d_tdense.as<double>()};
    }
    const double* p_matrix(int component);
    void ensure_esp();
};

namespace kumo {

std::unique_ptr<kumo_hip_molecule> load_molecule(kumo_hip_context* ctx, const kumo_hip_molecule_desc& d);
std::unique_ptr<kumo_hip_molecule> load_fchk(kumo_hip_context* ctx, const std::string& path);
std::vector<double> boys_table_host();

struct PointsArg {
    const kumo_hip_grid* grid = nullptr;
    const double* points = nullptr; // host or device (see options.device_outputs)
    long long n = 0;
};

void evaluate(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const PointsArg& pts,
              const kumo_hip_property* props, std::size_t nprops, const kumo_hip_options& opt, double* const* outputs,
              int32_t* backend_used);

void screening_masks(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const PointsArg& pts, int block_size,
                     int order, double threshold, uint8_t* kept);

} // namespace kumo
