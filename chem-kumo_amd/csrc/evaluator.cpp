// One evaluation on the device: the passes of Evaluator::Run
// (src/evaluator.cpp:556-802) in the reference order — point path, ESP,
// orbital pass, grid path — each over batches of whole screening blocks.
// Every screening block is evaluated by the same kernels whatever batch it
// falls in, so the results do not depend on the batch size, on the range a
// caller evaluates or on other work on the device; a serving loop can
// therefore cut a request into block-aligned ranges and run them on several
// streams.
#include "plan.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace kumo {
namespace {


void validate_options(const kumo_hip_options& o, const kumo_hip_molecule* mol) {
    if (o.block_size < 1 || o.block_size > 65536) invalid("evaluation block_size must be in [1, 65536]");
    if (o.threads < 0) invalid("invalid evaluation options");
    if (!std::isfinite(o.screening_threshold) || o.screening_threshold < 0.0)
        invalid("screening threshold must be finite and nonnegative");
    if (!std::isfinite(o.potential_tolerance) || o.potential_tolerance < 0.0)
        invalid("potential tolerance must be finite and nonnegative");
    if (!std::isfinite(o.fukui_delta) || o.fukui_delta <= 0.0) invalid("fukui_delta must be finite and positive");
    if (o.backend != KUMO_HIP_AUTO && o.backend != KUMO_HIP_AO_P && o.backend != KUMO_HIP_AO_MO)
        invalid("unknown evaluation backend");
    if (o.precision < KUMO_HIP_FP64 || o.precision > KUMO_HIP_BF16X6) invalid("unknown contraction precision");
    if (o.backend == KUMO_HIP_AO_MO && !mol->eligible)
        invalid("AO_MO backend requested but orbital representation is not eligible");
    if (o.threads > 1) invalid("explicit multi-thread control is not implemented in the API adapter");
    if (o.tile_points != 0 && o.tile_points != 16 && o.tile_points != 32 && o.tile_points != 64)
        invalid("tile_points must be 0, 16, 32 or 64");
    if (o.esp_far_field < KUMO_HIP_ESP_AUTO || o.esp_far_field > KUMO_HIP_ESP_EXACT)
        invalid("unknown electrostatic potential method");
}

// The checks of OrbitalEnergyFields (src/orbital_energy_fields.hpp:117-190,
// 229-233), in the reference's order, made before any pass
// (src/evaluator.cpp:579-596). `what` names the most restrictive group.
void check_orbital_energy(const kumo_hip_molecule* mol, const Plan& plan, double delta) {
    const std::string what = plan.orbital_weighted_fukui ? "the orbital-weighted Fukui function or dual descriptor"
                             : plan.virtual_orbitals     ? "the local electron affinity and related fields"
                                                         : "the average local ionization energy";
    if (plan.virtual_orbitals && !mol->unrestricted && mol->nalpha != mol->nbeta)
        invalid(what + " is not defined for a restricted open-shell calculation");
    if (plan.orbital_weighted_fukui && (mol->unrestricted || mol->nalpha != mol->nbeta))
        invalid(what + " is defined only for a restricted closed-shell calculation");
    if (plan.orbital_weighted_fukui && (!std::isfinite(delta) || delta <= 0.0))
        invalid(what + ": the Fukui delta must be finite and positive");
    if (!mol->eligible)
        invalid(what + " needs occupied MOs that reproduce the FCHK density matrices (an SCF calculation)");
    const auto& eb = mol->unrestricted ? mol->Eb : mol->Ea;
    const auto size = [](const std::vector<double>& v) { return static_cast<long long>(v.size()); };
    if (size(mol->Ea) < mol->nmo_a || (mol->unrestricted && size(mol->Eb) < mol->nmo_b) ||
        size(mol->Ea) < mol->nalpha || size(eb) < mol->nbeta)
        invalid(what + " needs the FCHK orbital energies");
    const int nvir = mol->unrestricted ? std::max(0, mol->nmo_a - mol->nalpha) + std::max(0, mol->nmo_b - mol->nbeta)
                                       : std::max(0, mol->nmo_a - std::max(mol->nalpha, mol->nbeta));
    if (plan.virtual_orbitals && nvir == 0) invalid(what + " needs virtual orbitals, and the FCHK has none");
    if (plan.orbital_weighted_fukui && (mol->nalpha == 0 || mol->nalpha >= mol->nmo_a))
        invalid(what + " needs occupied and virtual orbitals, and the FCHK lacks one kind");
}

// Normalized orbital-weighted Fukui weights of energies [first, last)
// (orbital_weights, src/orbital_energy_fields.hpp:75-96): Gaussians of the
// distance from mu, exponents shifted by their largest value.
std::vector<double> fukui_weights(const std::vector<double>& e, int first, int last, double mu, double delta) {
    std::vector<double> x(static_cast<std::size_t>(last - first)), w(x.size());
    for (int i = first; i < last; ++i) {
        const double t = (mu - e[static_cast<std::size_t>(i)]) / delta;
        x[static_cast<std::size_t>(i - first)] = -t * t;
    }
    const double top = *std::max_element(x.begin(), x.end());
    double sum = 0.0;
    for (std::size_t i = 0; i < x.size(); ++i) {
        w[i] = std::exp(x[i] - top);
        sum += w[i];
    }
    for (double& v : w)
        v /= sum;
    return w;
}

// Weights [kOeSums][slots] of the MO slots of the orbital-energy operand in
// the orbital sums (columns and energies of src/orbital_energy_fields.hpp:
// 155-192, 219-251): occupation n_i and n_i |e_i| for occupied MOs, 1, e_i
// and (e_i < 0) w e_i for virtual MOs, the Fukui weights g_i.
std::vector<double> orbital_energy_weights(const kumo_hip_molecule* mol, const Plan& plan, double delta) {
    const int ns = mol->oe_slots;
    std::vector<double> w(static_cast<std::size_t>(kOeSums) * static_cast<std::size_t>(ns), 0.0);
    const auto at = [&](int sum, int slot) -> double& {
        return w[static_cast<std::size_t>(sum) * static_cast<std::size_t>(ns) + static_cast<std::size_t>(slot)];
    };
    const double attach = mol->unrestricted ? 1.0 : 2.0;
    for (int q = 0; q < ns; ++q) {
        const OeSlot& sl = mol->oe_slot[static_cast<std::size_t>(q)];
        if (sl.kind == 0) continue;
        const double e = (sl.channel == 1 ? mol->Eb : mol->Ea)[static_cast<std::size_t>(sl.mo)];
        if (sl.kind == 1) {
            at(kOeOccDensity, q) = sl.occupation;
            at(kOeOccEnergy, q) = sl.occupation * std::abs(e);
        } else {
            at(kOeVirDensity, q) = 1.0;
            at(kOeVirEnergy, q) = e;
            if (e < 0.0) at(kOeAttachment, q) = attach * e;
        }
    }
    if (plan.orbital_weighted_fukui) { // restricted closed shell: alpha MOs only
        const int homo = mol->nalpha, count = mol->nmo_a;
        const auto& E = mol->Ea;
        const double mu = 0.5 * (E[static_cast<std::size_t>(homo - 1)] + E[static_cast<std::size_t>(homo)]);
        const auto plus = fukui_weights(E, homo, count, mu, delta);
        const auto minus = fukui_weights(E, 0, homo, mu, delta);
        for (int q = 0; q < ns; ++q) {
            const OeSlot& sl = mol->oe_slot[static_cast<std::size_t>(q)];
            if (sl.kind == 1) at(kOeFukuiMinus, q) = minus[static_cast<std::size_t>(sl.mo)];
            if (sl.kind == 2) at(kOeFukuiPlus, q) = plus[static_cast<std::size_t>(sl.mo - homo)];
        }
    }
    return w;
}

struct TableSet {
    KeptTables t;
};

// Tables of nblk screening blocks; fold_per_block > 0 adds the fold
// workspace (that many folded rows per block, fold_slots wide).
TableSet tables(Executor& ex, int nblk, int nao, int nsh, int max_chunks, int fold_per_block = 0,
                int fold_slots = 0) {
    const std::size_t b = static_cast<std::size_t>(nblk), ns = static_cast<std::size_t>(nsh);
    const std::size_t mc = static_cast<std::size_t>(max_chunks);
    const std::size_t fr = b * static_cast<std::size_t>(std::max(0, fold_per_block));
    const std::size_t sizes[] = {
        b * nao,                // kept
        4 * b * nao,            // kept_list
        4 * b,                  // nkept
        4 * b * ns,             // kept_shell
        4 * b * ns,             // kept_bits
        4 * b,                  // nkshell
        4 * b * ns,             // ent
        4 * b * ns,             // ent_bits
        b * ns,                 // ent_row
        4 * b * (mc + 1),       // chunk_begin
        8 * b * mc * kChunkRows, // chunk_rows
        4 * b,                  // nchunks
        8 * fr * static_cast<std::size_t>(std::max(0, fold_slots)), // fold_rows
        6 * fr * static_cast<std::size_t>(std::max(0, fold_slots)), // fold_lo
        8 * fr,                 // fold_norm
        4 * fr,                 // fold_desc
        fr ? 4 * b : 0,         // fold_n
    };
    const auto al = [](std::size_t x) { return (x + 255) / 256 * 256; };
    std::size_t total = 0;
    for (std::size_t x : sizes)
        total += al(x);
    ex.ws_tables.resize(total);
    char* p = ex.ws_tables.as<char>();
    void* ptr[17];
    for (int i = 0; i < 17; ++i) {
        ptr[i] = p;
        p += al(sizes[i]);
    }
    TableSet s;
    s.t.kept = static_cast<uint8_t*>(ptr[0]);
    s.t.kept_list = static_cast<int32_t*>(ptr[1]);
    s.t.nkept = static_cast<int32_t*>(ptr[2]);
    s.t.kept_shell = static_cast<int32_t*>(ptr[3]);
    s.t.kept_bits = static_cast<uint32_t*>(ptr[4]);
    s.t.nkshell = static_cast<int32_t*>(ptr[5]);
    s.t.ent = static_cast<int32_t*>(ptr[6]);
    s.t.ent_bits = static_cast<uint32_t*>(ptr[7]);
    s.t.ent_row = static_cast<uint8_t*>(ptr[8]);
    s.t.chunk_begin = static_cast<int32_t*>(ptr[9]);
    s.t.chunk_rows = static_cast<int32_t*>(ptr[10]);
    s.t.nchunks = static_cast<int32_t*>(ptr[11]);
    s.t.fold_rows = static_cast<double*>(ptr[12]);
    s.t.fold_lo = ptr[13];
    s.t.fold_norm = static_cast<float2*>(ptr[14]);
    s.t.fold_desc = static_cast<int32_t*>(ptr[15]);
    s.t.fold_n = static_cast<int32_t*>(ptr[16]);
    s.t.fold_per_block = fr ? fold_per_block : 0;
    s.t.fold_slots = fold_slots;
    s.t.fold_cap = static_cast<int>(fr);
    s.t.max_chunks = max_chunks;
    s.t.lists = 1;
    return s;
}

template <typename T> T* grow(DeviceBuffer& b, std::size_t count) {
    b.resize(std::max<std::size_t>(1, count) * sizeof(T));
    return b.as<T>();
}

void sync(Executor& ex) {
    if (!ex.capture) KUMO_HIP(hipStreamSynchronize(ex.stream));
}

template <typename T> void upload_async(Executor& ex, T* dst, const T* src, std::size_t count) {
    const void* from = src;
    if (ex.keep) {
        ex.keep->emplace_back(reinterpret_cast<const char*>(src), reinterpret_cast<const char*>(src + count));
        from = ex.keep->back().data();
    }
    KUMO_HIP(hipMemcpyAsync(dst, from, count * sizeof(T), hipMemcpyHostToDevice, ex.stream));
}

void record(Executor& ex, bool start, int kind = 0);

hipEvent_t next_event(Executor& ex, bool start, int kind = 0) {
    auto& ev = ex.events[kind];
    std::size_t& used = ex.events_used[kind];
    if (used == ev.size()) {
        hipEvent_t a, b;
        KUMO_HIP(hipEventCreate(&a));
        KUMO_HIP(hipEventCreate(&b));
        ev.push_back({a, b});
    }
    auto& pr = ev[used];
    if (!start) ++used;
    return start ? pr.first : pr.second;
}

void record(Executor& ex, bool start, int kind) {
    if (!ex.capture || ex.events_in_capture) KUMO_HIP(hipEventRecord(next_event(ex, start, kind), ex.stream));
}

// Relative error of the channel moments assumed by the ELF rescue, per
// contraction precision. On a 640-AO UHF molecule over 4.1e6 points the
// BF16X6 estimate with 5e-7 flags 14 % of the points and leaves at most 8e-7
// ELF error on the others (BF16X3 moments are about 100x less accurate).
double rescue_eps(int precision) {
    switch (precision) {
    case KUMO_HIP_FP32: return 1e-6;
    case KUMO_HIP_BF16X3: return 2e-4;
    default: return 5e-7;
    }
}
// Largest ELF error the model may leave to the contraction precision: a third
// of the 1e-5 the mixed precisions are held to (the model overestimates: at
// 1e-6 the largest error left on the 640-AO cube was 1.8e-7).
constexpr double kElfRescueTol = 3e-6;

// Density rescue: unit of the contraction's rounding for the precisions
// whose operands carry 24 bits (BF16X3, with 16-bit operands, has no density
// rescue). The kernels accumulate, per point and channel, the scale
// s_c = sqrt(sum_k Phi_k^2 ||C_k,c||^2): the bf16 MFMA rounds each 8-term K
// group to about 24 bits with random sign, so the error of Psi_i is about
// 2^-24 sqrt(sum_k (C_ik Phi_k)^2) and that of rho_c about
// 2 eps sqrt(rho_c) s_c (the sum over k with absolute values, the worst case,
// overestimated it by up to the square root of the kept rows). Calibrated on
// the 1.66e8 points of the 256-molecule batch against the FP64 path:
// eps = 1.0 x 2^-24 leaves no unflagged point with a relative density error
// above 7.1e-7 (0.7 x 2^-24: 8.3e-7; 0.4 x 2^-24: 1.5e-6).
constexpr double kRhoEpsBf16x6 = 1.0 / 16777216.0, kRhoEpsFp32 = 2.25 / 16777216.0;
double rho_rescue_eps(int precision) {
    switch (precision) {
    case KUMO_HIP_FP32: return kRhoEpsFp32;
    case KUMO_HIP_BF16X6: return kRhoEpsBf16x6;
    default: return 0.0;
    }
}
constexpr double kRhoRescueTol = 5e-7; // largest relative density error left to the contraction

// ELF rescue from per-point error scales (BF16X6): the unit e of the error
// model, calibrated. The model bounds each channel's errors of rho, |grad rho|
// and tau from the density scale s_c and the gradient scale
// w_c = sqrt(sum_k |grad Phi_k|^2 ||C_k,c||^2) (Cauchy-Schwarz over the MOs)
// and carries the intervals through D = tau - |grad rho|^2 / 8 rho to ELF;
// the bounds compound, so e is fitted: over the 256-molecule batch (1.66e8
// points, with the density rescue at 1.0 x 2^-24) e = 3e-10 leaves no ELF
// error above 2.3e-6 where rho >= 1e-14 and flags, with the density, 7.6 % of
// the points (e = 1e-9: 1.1e-6 and 8.9 %; e = 1e-10: 4.5e-6 and 6.7 %).
// Other precisions use the model of rescue_eps.
double elf_scale_eps(int precision) { return precision == KUMO_HIP_BF16X6 ? 3e-10 : 0.0; }

// Autotuner of the AO->MO tile. The points per workgroup (16 or 32) change
// only how points are shared among workgroups (every point's arithmetic and
// summation order are the same, so results are identical): the first launch
// of a (precision, derivative order, Hessian) on a molecule with at least
// kTuneBlocks screening blocks times both tiles on that batch (each launched
// twice, the second launch timed, so module loading and cold caches do not
// count) and the faster one is kept for the molecule's lifetime. Smaller
// batches, and batches being recorded into a graph, use 32 until a choice
// has been made.
constexpr int kTuneBlocks = 256;

int tuned_tile(Executor& ex, kumo_hip_molecule* mol, AoMoLaunch a, hipStream_t s) {
    if (a.precision == KUMO_HIP_FP64) return 16;
    if (ao_mo_wave_supported(a)) return 64;
    const int key = (a.precision * 4 + a.order) * 4 + a.hessian;
    {
        std::lock_guard<std::mutex> lock(mol->tune_mutex);
        const auto it = mol->tuned_tile.find(key);
        if (it != mol->tuned_tile.end()) return it->second;
    }
    if (ex.capture || a.nblk < kTuneBlocks) return 32;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KUMO_HIP(hipEventCreate(&e0));
    KUMO_HIP(hipEventCreate(&e1));
    const int cand[2] = {32, 16};
    float ms[2] = {0.f, 0.f};
    try {
        for (int i = 0; i < 2; ++i) {
            a.tile_points = cand[i];
            launch_ao_mo(a, s);
            KUMO_HIP(hipEventRecord(e0, s));
            launch_ao_mo(a, s);
            KUMO_HIP(hipEventRecord(e1, s));
            KUMO_HIP(hipEventSynchronize(e1));
            KUMO_HIP(hipEventElapsedTime(&ms[i], e0, e1));
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int best = ms[1] < ms[0] ? cand[1] : cand[0];
    std::lock_guard<std::mutex> lock(mol->tune_mutex);
    return mol->tuned_tile.emplace(key, best).first->second;
}

RescuePlan rescue_plan(Executor& ex, long long count, int nblk, int block_size, int tiles_per_block) {
    const std::size_t pts = static_cast<std::size_t>(nblk) * static_cast<std::size_t>(block_size);
    const std::size_t tiles = static_cast<std::size_t>(nblk) * static_cast<std::size_t>(tiles_per_block);
    const auto al = [](std::size_t x) { return (x + 255) / 256 * 256; };
    const std::size_t sizes[] = {static_cast<std::size_t>(count), 2 * pts, 4 * static_cast<std::size_t>(nblk),
                                 4 * tiles, 4 * tiles, 8};
    std::size_t total = 0;
    for (std::size_t x : sizes)
        total += al(x);
    ex.ws_rescue.resize(total);
    char* p = ex.ws_rescue.as<char>();
    void* ptr[6];
    for (int i = 0; i < 6; ++i) {
        ptr[i] = p;
        p += al(sizes[i]);
    }
    return {static_cast<uint8_t*>(ptr[0]), static_cast<uint16_t*>(ptr[1]), static_cast<int32_t*>(ptr[2]),
            static_cast<int32_t*>(ptr[3]), static_cast<int32_t*>(ptr[4]), static_cast<int32_t*>(ptr[5])};
}

} // namespace

// validated_point_count (src/grid_geometry.hpp:30-52).
long long validated_point_count(const kumo_hip_grid& g) {
    long long count = 1;
    for (int a = 0; a < 3; ++a) {
        if (g.counts[a] <= 0 || static_cast<unsigned long long>(g.counts[a]) >
                                    static_cast<unsigned long long>(std::numeric_limits<long long>::max() / count))
            invalid("invalid or oversized grid");
        count *= g.counts[a];
    }
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(g.origin[i])) invalid("invalid grid geometry: nonfinite origin or step");
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(g.steps[i])) invalid("invalid grid geometry: nonfinite origin or step");
    double d[3][3];
    for (int a = 0; a < 3; ++a) {
        const double* s = g.steps + 3 * a;
        const double len = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        if (!std::isfinite(len) || len <= 0.0) invalid("invalid grid geometry: zero step vector");
        for (int i = 0; i < 3; ++i)
            d[a][i] = s[i] / len;
    }
    const double det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                       d[1][0] * (d[0][1] * d[2][2] - d[0][2] * d[2][1]) +
                       d[2][0] * (d[0][1] * d[1][2] - d[0][2] * d[1][1]);
    if (std::abs(det) < 1e-12) invalid("invalid grid geometry: linearly dependent axes");
    return count;
}

// Constants of the derived fields computed with the host libm exactly as the
// reference computes them.
FieldsLaunch fields_launch(const double* channels, int from_spin, int closed, long long n, long long out_offset,
                          int ntasks, const FieldTask* tasks) {
    constexpr double pi = 3.141592653589793;
    FieldsLaunch f{channels, from_spin, closed, n, out_offset, ntasks, tasks, 0.3 * std::pow(6.0 * pi * pi, 2.0 / 3.0),
                   2.0 * std::cbrt(3.0 * pi * pi)};
    return f;
}

PointSource make_source(const kumo_hip_grid* grid, const double* device_points) {
    PointSource ps{};
    if (grid) {
        for (int i = 0; i < 3; ++i)
            ps.origin[i] = grid->origin[i];
        for (int i = 0; i < 9; ++i)
            ps.steps[i] = grid->steps[i];
        ps.ny = grid->counts[1];
        ps.nz = grid->counts[2];
        ps.points = nullptr;
    } else {
        ps.ny = ps.nz = 1;
        ps.points = device_points;
    }
    return ps;
}

Prepared prepare(const kumo_hip_molecule* mol, const PointsArg& pts, const kumo_hip_property* props,
                 std::size_t nprops, const kumo_hip_options& opt, bool points_on_device) {
    validate_options(opt, mol);
    Prepared r;
    r.grid = pts.grid != nullptr;
    if (r.grid) {
        r.n = validated_point_count(*pts.grid);
    } else {
        r.n = pts.n;
        if (r.n < 0) invalid("invalid point count");
        if (!points_on_device)
            for (long long q = 0; q < 3 * r.n; ++q)
                if (!std::isfinite(pts.points[q])) invalid("evaluation points must be finite");
    }
    r.plan = make_plan(props, nprops);
    r.backend = choose_backend(opt.backend, mol->eligible, r.plan);
    if (r.plan.orbital_energies) check_orbital_energy(mol, r.plan, opt.fukui_delta);
    for (std::size_t i = 0; i < nprops; ++i) {
        const auto& p = props[i];
        if (is_orbital_property(p.kind)) {
            const int nmo = (p.channel == 1 && mol->unrestricted) ? mol->nmo_b : mol->nmo_a;
            if (nmo == 0) runtime("FCHK has no MO coefficients");
            if (p.orbital_index > nmo) out_of_range("MO index outside available coefficient range");
        }
    }
    r.props.assign(props, props + nprops);
    std::size_t c = 0;
    for (std::size_t i = 0; i < nprops; ++i) {
        r.first_channel.push_back(c);
        c += static_cast<std::size_t>(channels(props[i].kind));
    }
    r.nout = c;
    for (int comp = 0; comp < 4; ++comp)
        if (uses_point_path(r.plan.c[static_cast<std::size_t>(comp)], r.backend, r.grid)) r.needed.push_back(comp);
    bool orb_ch[2] = {false, false};
    for (const auto& p : r.props) {
        if (is_orbital_property(p.kind) && !on_grid_path(p, r.backend, r.grid)) r.orbital_pass = true;
        if (on_grid_path(p, r.backend, r.grid) && p.kind == kDensity) ++r.grid_density_passes;
        if (on_grid_path(p, r.backend, r.grid) && p.kind == kOrbital) orb_ch[p.channel] = true;
    }
    r.grid_orbital_passes = (orb_ch[0] ? 1 : 0) + (orb_ch[1] ? 1 : 0);
    r.passes = (r.needed.empty() ? 0 : 1) + (r.plan.esp() ? 1 : 0) + (r.plan.orbital_energies ? 1 : 0) +
               (r.orbital_pass ? 1 : 0) + r.grid_density_passes + r.grid_orbital_passes;
    return r;
}

// The workspace sizes of the point pass of run_range, in the order it grows
// them: screening tables (tables()), channel moments, field tasks, and for a
// reduced precision the error scales, the rescue plan (rescue_plan()) and the
// rescue moments; for AO-P the staged panels and products. Each piece is
// rounded up to the arena's 256-byte alignment in `fixed`.
void point_pass_bytes(const kumo_hip_molecule* mol, const Prepared& pr, const kumo_hip_options& opt,
                      std::size_t& fixed, std::size_t& per_block) {
    const std::size_t nao = static_cast<std::size_t>(mol->n), ns = mol->basis.shells.size();
    const std::size_t mc = static_cast<std::size_t>(mol->max_chunks), bs = static_cast<std::size_t>(opt.block_size);
    const int ao_order = pr.grid ? pr.plan.max_order : std::max(1, pr.plan.max_order);
    fixed = 32 * 256 + static_cast<std::size_t>(pr.props.size()) * sizeof(FieldTask);
    per_block = 5 * nao + 17 * ns + 4 * (mc + 1) + 8 * mc * kChunkRows + 12; // tables
    if (pr.backend == KUMO_HIP_AO_MO) {
        const std::size_t groups = static_cast<std::size_t>((mol->slots + 127) / 128);
        per_block += groups * 2 * kMoments * bs * 8; // moments
        if (opt.precision != KUMO_HIP_FP64) {
            const std::size_t tpb = (bs + kRescueTile - 1) / kRescueTile;
            per_block += 4 * bs * 4;                                         // error scales
            per_block += 3 * bs + 4 + 8 * tpb;                               // rescue plan
            per_block += groups * 2 * kMoments * tpb * kRescueTile * 8;      // rescue moments
        }
    } else {
        const std::size_t pan = ao_order == 0 ? 1 : (ao_order == 1 ? 4 : 10);
        per_block += 2 * kMoments * bs * 8 + nao * pan * bs * 8 + nao * 4 * bs * 8;
    }
}

// The far field only where it is asked for (KUMO_HIP_ESP_FAR) and the
// tolerance is positive: the exact pair sums run at about 5e11 pair-points/s
// and outran the far field at every workload measured (100 000 surface
// points of a 20-atom molecule: 4.3 ms exact, 106 ms far field; an 80^3 grid
// of a 640-AO molecule: 110 ms, 591 ms), so KUMO_HIP_ESP_AUTO sums exactly.
bool use_far_field(const Executor& ex, const kumo_hip_options& opt) {
    return opt.potential_tolerance > 0.0 && opt.esp_far_field == KUMO_HIP_ESP_FAR && !ex.capture && !ex.no_alloc &&
           !ex.exact_esp;
}

// The potential at points [q0, q1) in target clusters under the far-field
// plan (evaluate_electrostatic_potential, src/evaluator.cpp:391-479): on a
// grid, tiles of up to 8^3 points about the centres of their boxes of grid
// indices, one plan whose radius covers the largest tile; at listed points,
// the axis-aligned cubes of edge 1.6 bohr holding them, about the cube
// centres, one plan covering half a cube diagonal. A value depends only on
// its point, its target and the plan, so a range holds the part of each
// target that falls in it. The targets run in batches sized by the
// workspace (per target: a near-pair list, a byte per pair and per source
// cluster, the local expansion).
void far_field_range(Executor& ex, kumo_hip_molecule* mol, const Prepared& pr, const PointSource& ps,
                     const kumo_hip_options& opt, double* out, long long q0, long long q1, const Advance& advance) {
    hipStream_t s = ex.stream;
    std::vector<FarTargetDev> targets;
    std::vector<long long> in_range; // points of each target in [q0, q1)
    std::vector<int64_t> idx;
    double radius = 0.0;
    if (pr.grid) {
        const long long ny = ps.ny, nz = ps.nz, nx = pr.n / (ny * nz);
        const long long counts[3] = {nx, ny, nz};
        long long tiles[3];
        double edges[3];
        for (int a = 0; a < 3; ++a) {
            edges[a] = static_cast<double>(std::min<long long>(kFarTileSide, counts[a]) - 1);
            tiles[a] = (counts[a] + kFarTileSide - 1) / kFarTileSide;
        }
        // grid_box_radius (src/potential_clusters.hpp:37-49)
        for (int corner = 0; corner < 8; ++corner) {
            double o[3], v[3];
            for (int a = 0; a < 3; ++a)
                o[a] = ((corner >> a) & 1 ? 0.5 : -0.5) * edges[a];
            for (int a = 0; a < 3; ++a)
                v[a] = (ps.steps[a] * o[0] + ps.steps[3 + a] * o[1]) + ps.steps[6 + a] * o[2];
            radius = std::max(radius, std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]));
        }
        const long long t0 = (q0 / (ny * nz)) / kFarTileSide, t1 = ((q1 - 1) / (ny * nz)) / kFarTileSide;
        for (long long ti = t0; ti <= t1; ++ti)
            for (long long tj = 0; tj < tiles[1]; ++tj)
                for (long long tk = 0; tk < tiles[2]; ++tk) {
                    const long long t[3] = {ti, tj, tk};
                    FarTargetDev tg{};
                    int count = 1;
                    double f[3];
                    for (int a = 0; a < 3; ++a) {
                        tg.lo[a] = static_cast<int32_t>(t[a] * kFarTileSide);
                        tg.hi[a] = static_cast<int32_t>(std::min<long long>(t[a] * kFarTileSide + kFarTileSide, counts[a]));
                        count *= tg.hi[a] - tg.lo[a];
                        f[a] = 0.5 * (tg.lo[a] + tg.hi[a] - 1);
                    }
                    const long long qmin = (static_cast<long long>(tg.lo[0]) * ny + tg.lo[1]) * nz + tg.lo[2];
                    const long long qmax = (static_cast<long long>(tg.hi[0] - 1) * ny + (tg.hi[1] - 1)) * nz + (tg.hi[2] - 1);
                    if (qmax < q0 || qmin >= q1) continue;
                    for (int a = 0; a < 3; ++a)
                        tg.c[a] = ps.origin[a] + ((ps.steps[a] * f[0] + ps.steps[3 + a] * f[1]) + ps.steps[6 + a] * f[2]);
                    tg.count = count;
                    long long held = count; // points of the tile in [q0, q1)
                    if (qmin < q0 || qmax >= q1) {
                        held = 0;
                        for (int i = tg.lo[0]; i < tg.hi[0]; ++i)
                            for (int j = tg.lo[1]; j < tg.hi[1]; ++j) {
                                const long long row = (static_cast<long long>(i) * ny + j) * nz;
                                held += std::max(0LL, std::min(row + tg.hi[2], q1) - std::max(row + tg.lo[2], q0));
                            }
                    }
                    targets.push_back(tg);
                    in_range.push_back(held);
                }
    } else {
        // cubes of listed points (src/evaluator.cpp:449-479, cluster_by
        // src/potential_clusters.hpp:51-76); the points come back from the
        // device, where the evaluation holds them
        const long long m = q1 - q0;
        std::vector<double> host(static_cast<std::size_t>(3 * m));
        KUMO_HIP(hipMemcpyAsync(host.data(), ps.points + 3 * q0, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        KUMO_HIP(hipStreamSynchronize(s));
        struct Item {
            std::array<int64_t, 3> cell;
            int64_t q;
        };
        std::vector<Item> items(static_cast<std::size_t>(m));
        for (long long j = 0; j < m; ++j) {
            Item& it = items[static_cast<std::size_t>(j)];
            for (int a = 0; a < 3; ++a) {
                const double index = std::floor(host[static_cast<std::size_t>(3 * j + a)] / kFarPointCube);
                if (!(std::abs(index) < 1e15)) invalid("point too far for the electrostatic potential");
                it.cell[static_cast<std::size_t>(a)] = static_cast<int64_t>(index);
            }
            it.q = q0 + j;
        }
        std::sort(items.begin(), items.end(),
                  [](const Item& a, const Item& b) { return a.cell != b.cell ? a.cell < b.cell : a.q < b.q; });
        idx.resize(items.size());
        for (std::size_t j = 0; j < items.size(); ++j) {
            idx[j] = items[j].q;
            if (j == 0 || items[j].cell != items[j - 1].cell) {
                FarTargetDev tg{};
                for (int a = 0; a < 3; ++a)
                    tg.c[a] = (static_cast<double>(items[j].cell[static_cast<std::size_t>(a)]) + 0.5) * kFarPointCube;
                tg.first = static_cast<int32_t>(j);
                targets.push_back(tg);
            }
            ++targets.back().count;
        }
        if (items.size() > static_cast<std::size_t>(std::numeric_limits<int32_t>::max()))
            invalid("too many points in one electrostatic potential range");
        for (const FarTargetDev& tg : targets)
            in_range.push_back(tg.count);
        radius = 0.5 * std::sqrt(3.0) * kFarPointCube;
    }
    const FarPlanDev& plan = mol->far_plan(radius, opt.potential_tolerance, s);
    const std::size_t nt = targets.size();
    const std::size_t npairs = static_cast<std::size_t>(plan.npairs);
    constexpr std::size_t kAlign = 256;
    const auto up = [](std::size_t b) { return (b + kAlign - 1) / kAlign * kAlign; };
    const std::size_t per_target = 4 * npairs + (npairs + static_cast<std::size_t>(plan.nclusters)) +
                                   far_hcount(kFarMultipole) * sizeof(double) + 4 + 64;
    constexpr std::size_t kBudget = std::size_t(1) << 30;
    const std::size_t batch = std::max<std::size_t>(1, std::min(nt, kBudget / per_target));
    const std::size_t b_targets = up(nt * sizeof(FarTargetDev)), b_idx = up(std::max<std::size_t>(1, idx.size()) * 8);
    const std::size_t b_local = up(batch * far_hcount(kFarMultipole) * sizeof(double)), b_nnear = up(batch * 4),
                      b_near = up(batch * npairs * 4),
                      b_work = up(batch * (npairs + static_cast<std::size_t>(plan.nclusters)));
    char* base = grow<char>(ex.ws_far, b_targets + b_idx + b_local + b_nnear + b_near + b_work);
    auto* d_targets = reinterpret_cast<FarTargetDev*>(base);
    auto* d_idx = reinterpret_cast<int64_t*>(base + b_targets);
    auto* d_local = reinterpret_cast<double*>(base + b_targets + b_idx);
    auto* d_nnear = reinterpret_cast<int32_t*>(base + b_targets + b_idx + b_local);
    auto* d_near = reinterpret_cast<int32_t*>(base + b_targets + b_idx + b_local + b_nnear);
    auto* d_work = reinterpret_cast<int8_t*>(base + b_targets + b_idx + b_local + b_nnear + b_near);
    upload_async(ex, d_targets, targets.data(), nt);
    if (!idx.empty()) upload_async(ex, d_idx, idx.data(), idx.size());
    const auto* pairs = mol->d_esp_pairs.as<EspPairDev>();
    const auto* coef = mol->d_esp_coef.as<double>();
    for (std::size_t b0 = 0; b0 < nt; b0 += batch) {
        const int nb = static_cast<int>(std::min(batch, nt - b0));
        launch_far_targets(plan, pairs, coef, ex.boys_table, d_targets + b0, nb, d_local, d_near, d_nnear, d_work, s);
        launch_far_points(pairs, mol->esp_groups, coef, ex.boys_table, mol->d_esp_atoms.as<EspAtomDev>(),
                          static_cast<int>(mol->atoms.size()), ps, q0, q1, d_targets + b0, nb, d_idx, d_local, d_near,
                          d_nnear, plan.npairs, out, s);
        sync(ex);
        long long done = 0;
        for (std::size_t t = b0; t < b0 + static_cast<std::size_t>(nb); ++t)
            done += in_range[t];
        advance(static_cast<std::size_t>(done));
    }
}

void run_range(Executor& ex, kumo_hip_molecule* mol, const Prepared& pr, const PointSource& ps,
               const kumo_hip_options& opt, double* const* dout, long long q0, long long q1, const Advance& advance) {
    hipStream_t s = ex.stream;
    const auto& props = pr.props;
    const std::size_t nprops = props.size();
    const long long n = pr.n;
    const bool grid = pr.grid;
    const int backend = pr.backend;
    const Plan& plan = pr.plan;
    const MolView mv = mol->view();
    const int nao = mol->n, nsh = static_cast<int>(mol->basis.shells.size());
    const int bs = opt.block_size;
    const double thr = opt.screening_threshold;
    if (q0 % bs != 0) invalid("evaluation ranges must start at a screening block boundary");
    q1 = std::min(q1, n);
    if (q1 <= q0) return;
    const long long range = q1 - q0;
    const auto& needed = pr.needed;

    // ---------------------------------------------------------------- point pass
    if (!needed.empty()) {
        const int ao_order = grid ? plan.max_order : std::max(1, plan.max_order);
        bool full_h = false, diag_h = false;
        for (int c = 0; c < 4; ++c) {
            full_h = full_h || plan.c[static_cast<std::size_t>(c)].full_hessian();
            diag_h = diag_h || plan.c[static_cast<std::size_t>(c)].hessian_diagonal();
        }
        const long long b_first = q0 / bs, b_end = (q1 + bs - 1) / bs;
        const long long nblocks = b_end - b_first;
        const long long target = 1ll << 21;
        int cap = static_cast<int>(std::min<long long>({nblocks, std::max(1ll, target / bs), 8192ll}));
        if (backend == KUMO_HIP_AO_P) {
            const long long per_block = static_cast<long long>(nao) * (ao_order == 0 ? 1 : (ao_order == 1 ? 4 : 10)) * bs;
            cap = static_cast<int>(std::max(1ll, std::min<long long>(cap, (512ll << 20) / (8 * per_block))));
        }
        if (ex.max_batch_blocks > 0) cap = static_cast<int>(std::min<long long>(cap, ex.max_batch_blocks));
        const double* sr2 = ex.no_alloc ? mol->screen_r2_resident(thr) : mol->screen_r2(thr, s);
        if (!sr2) invalid("the allocation-free evaluation needs the screening radii of its threshold resident "
                          "(an evaluation at that threshold through kumo_hip_evaluate_* makes them so)");
        // Folded rows of partially kept pure blocks (about 2 % of the rows of a
        // def2-TZVP-shaped basis at 1e-12) are formed once per batch into the
        // fold workspace: up to 16 per block, within 256 MiB; the others are
        // formed where each chunk is staged.
        int fold_rows = 0;
        if (backend == KUMO_HIP_AO_MO && !ex.no_alloc) {
            const long long per_row = 14ll * mol->slots + 12;
            fold_rows = static_cast<int>(std::min<long long>(16, (256ll << 20) / (per_row * cap)));
        }
        const auto ts = tables(ex, cap, nao, nsh, mol->max_chunks, fold_rows, mol->slots);
        const long long batch_max = static_cast<long long>(cap) * bs;
        int slot0 = 0, slot1 = 0;
        bool need_a = false, need_b = false;
        for (int c : needed) {
            if (c == kAlpha || c == kTotal || c == kDifference) need_a = true;
            if (c == kBeta || c == kTotal || c == kDifference) need_b = true;
        }
        const bool closed = !mol->unrestricted;
        if (backend == KUMO_HIP_AO_MO) {
            if (closed) {
                slot0 = 0;
                slot1 = mol->alpha_slots;
            } else {
                slot0 = need_a ? 0 : mol->alpha_slots;
                slot1 = need_b ? mol->slots : mol->alpha_slots;
            }
            if (slot1 == slot0) slot1 = slot0 + 16;
        }
        const int groups = backend == KUMO_HIP_AO_MO ? (slot1 - slot0 + 127) / 128 : 1;
        double* chm = grow<double>(ex.ws_moments, static_cast<std::size_t>(groups) * 2 * kMoments * batch_max);
        std::vector<FieldTask> tasks;
        for (std::size_t i = 0; i < nprops; ++i) {
            const auto& p = props[i];
            if (!from_point_path(p, backend, grid)) continue;
            FieldTask t{};
            t.kind = p.kind;
            t.component = p.component;
            for (int c = 0; c < channels(p.kind); ++c)
                t.out[c] = dout[pr.first_channel[i] + static_cast<std::size_t>(c)];
            tasks.push_back(t);
        }
        int elf_which = 0; // ELF requests on this pass: bit 0 total, 1 alpha, 2 beta
        int rho_which = 0; // density requests (the spin difference has no relative error bound)
        const auto bits = [](int component) {
            return component == kTotal ? 1 : (component == kAlpha ? 2 : (component == kBeta ? 4 : 0));
        };
        for (const auto& t : tasks) {
            if (t.kind == kElf) elf_which |= t.component == kDifference ? 1 : bits(t.component);
            if (t.kind == kDensity) rho_which |= bits(t.component);
        }
        FieldTask* dtasks = grow<FieldTask>(ex.ws_tasks, tasks.size());
        if (ex.no_alloc) launch_store_tasks(tasks.data(), static_cast<int>(tasks.size()), dtasks, s);
        else upload_async(ex, dtasks, tasks.data(), tasks.size());
        // Per batch: kept AO counts of its blocks, then its rescue tiles,
        // copied back asynchronously; read after one synchronisation at the
        // end (or after each batch when a progress callback must see
        // completed work).
        const long long nbatches = (b_end - b_first + cap - 1) / cap;
        const std::size_t slot = static_cast<std::size_t>(cap) + 2;
        int32_t* hst = ex.capture ? nullptr : ex.host_stats(static_cast<std::size_t>(nbatches) * slot);
        if (hst) std::fill(hst, hst + static_cast<std::size_t>(nbatches) * slot, 0);
        const bool stepwise = opt.progress != nullptr;
        double flops = 0.0;
        const int panels = ao_order <= 0 ? 1 : (ao_order == 1 ? 4 : (full_h ? 10 : (diag_h ? 7 : 5)));
        for (long long b0 = b_first; b0 < b_end; b0 += cap) {
            const int nb = static_cast<int>(std::min<long long>(cap, b_end - b0));
            const long long first = b0 * bs;
            const long long count = std::min<long long>(static_cast<long long>(nb) * bs, q1 - first);
            KeptTables kt = ts.t;
            kt.lists = backend == KUMO_HIP_AO_MO ? 0 : 1;
            launch_masks(mv, ps, n, bs, b0, nb, std::max(0, ao_order), thr, sr2, kt, s);
            if (backend == KUMO_HIP_AO_MO) {
                AoMoLaunch a{};
                a.mol = mv;
                a.ps = ps;
                a.mo = {mol->d_op64.as<double>(), mol->d_op32.as<float>(), mol->d_opbf.as<uint16_t>(),
                        mol->d_opnorm.as<float2>(), mol->operand_rows, mol->slots, mol->alpha_slots};
                if (kt.fold_per_block > 0) launch_fold(mv, a.mo, kt, nb, opt.precision, s);
                a.tables = ts.t;
                a.npoints = n;
                a.batch_first = first;
                a.batch_points = count;
                a.block_size = bs;
                a.first_block = b0;
                a.nblk = nb;
                a.order = std::max(0, ao_order);
                a.hessian = ao_order == 2 ? (full_h ? 2 : (diag_h ? 1 : 3)) : 0;
                a.precision = opt.precision;
                a.slot0 = slot0;
                a.slot1 = slot1;
                a.moments = chm;
                const bool rho_rescue = rho_which && rho_rescue_eps(opt.precision) > 0.0;
                const bool elf_scaled = elf_which && ao_order >= 1 && elf_scale_eps(opt.precision) > 0.0;
                // value scales [2][count], then (with gradient panels) gradient scales [2][count]
                a.rho_scale = rho_rescue || elf_scaled
                                  ? grow<float>(ex.ws_rscale, 4 * static_cast<std::size_t>(batch_max))
                                  : nullptr;
                a.tile_points = opt.tile_points ? opt.tile_points : tuned_tile(ex, mol, a, s);
                if (a.tile_points == 64 && !ao_mo_wave_supported(a))
                    invalid("tile_points 64 takes the split-bf16 precisions without full Hessians");
                record(ex, true);
                launch_ao_mo(a, s);
                record(ex, false);
                ++ex.last_launches;
                for (int g = 1; g < groups; ++g)
                    launch_add_into(chm, chm + static_cast<long long>(g) * 2 * kMoments * count, 2 * kMoments * count, s);
                if ((elf_which || rho_rescue) && opt.precision != KUMO_HIP_FP64) {
                    // FP64 rescue of the points whose ELF or density the
                    // contraction precision leaves uncertain (same masks,
                    // same chunks).
                    const int tpb = (bs + kRescueTile - 1) / kRescueTile;
                    const long long rslots = static_cast<long long>(nb) * tpb * kRescueTile;
                    const RescuePlan rp = rescue_plan(ex, count, nb, bs, tpb);
                    RescueCriteria rc{};
                    rc.elf_which = elf_which;
                    rc.elf_eps = rescue_eps(opt.precision);
                    rc.elf_tol = kElfRescueTol;
                    rc.rho_which = rho_rescue ? rho_which : 0;
                    rc.rho_scale = a.rho_scale;
                    rc.rho_eps = rho_rescue_eps(opt.precision);
                    rc.rho_tol = kRhoRescueTol;
                    rc.grad_scale = elf_scaled ? a.rho_scale + 2 * count : nullptr;
                    rc.elf_scale_eps = elf_scale_eps(opt.precision);
                    launch_rescue_plan(chm, count, closed ? 1 : 0, rc, bs, nb, kRescueTile, rp, s);
                    AoMoLaunch r = a;
                    r.precision = KUMO_HIP_FP64;
                    r.rho_scale = nullptr;
                    r.moments = grow<double>(ex.ws_rmom, static_cast<std::size_t>(groups) * 2 * kMoments * rslots);
                    r.batch_points = rslots;
                    r.r_tile_blk = rp.tile_blk;
                    r.r_tile_sub = rp.tile_sub;
                    r.r_total = rp.total;
                    r.r_count = rp.count;
                    r.r_offs = rp.offs;
                    record(ex, true, 1);
                    launch_ao_mo(r, s);
                    record(ex, false, 1);
                    launch_rescue_scatter(r.moments, rslots, groups, kRescueTile, bs, rp, chm, count, s);
                    if (hst)
                        KUMO_HIP(hipMemcpyAsync(hst + static_cast<std::size_t>((b0 - b_first) / cap) * slot + cap,
                                                rp.total, 8, hipMemcpyDeviceToHost, s));
                }
            } else {
                AoPLaunch a{};
                a.mol = mv;
                a.ps = ps;
                a.tables = ts.t;
                const bool tn = std::any_of(needed.begin(), needed.end(), [](int c) { return c != kDifference; });
                const bool sn = std::any_of(needed.begin(), needed.end(), [](int c) { return c != kTotal; });
                a.p[0] = tn ? mol->p_matrix(0, s) : nullptr;
                a.p[1] = sn ? mol->p_matrix(1, s) : nullptr;
                a.npoints = n;
                a.batch_first = first;
                a.batch_points = count;
                a.block_size = bs;
                a.first_block = b0;
                a.nblk = nb;
                a.order = ao_order;
                const int pan = ao_order == 0 ? 1 : (ao_order == 1 ? 4 : 10);
                a.phi = grow<double>(ex.ws_phi, static_cast<std::size_t>(nb) * nao * pan * bs);
                a.w = grow<double>(ex.ws_w, static_cast<std::size_t>(nb) * nao * 4 * bs);
                a.moments = chm;
                launch_fill(chm, 2 * kMoments * count, 0.0, s);
                record(ex, true);
                launch_ao_p(a, s);
                record(ex, false);
                ++ex.last_launches;
            }
            const FieldsLaunch f = fields_launch(chm, backend == KUMO_HIP_AO_MO ? 1 : 0,
                                                 backend == KUMO_HIP_AO_MO && closed ? 1 : 0, count, first,
                                                 static_cast<int>(tasks.size()), dtasks);
            record(ex, true, 2);
            launch_fields(f, s);
            record(ex, false, 2);
            if (!tasks.empty()) {
                // bytes the reduction must move: the channel moments read, the outputs written
                std::size_t outs = 0;
                for (const auto& t : tasks)
                    outs += static_cast<std::size_t>(channels(t.kind));
                const int read_ch = backend == KUMO_HIP_AO_MO && closed ? 1 : 2;
                ex.last_fields_points += count;
                ex.last_fields_bytes += count * static_cast<long long>(8 * (read_ch * kMoments + outs));
            }
            if (ex.capture) {
                advance(static_cast<std::size_t>(count));
                continue;
            }
            int32_t* hb = hst + static_cast<std::size_t>((b0 - b_first) / cap) * slot;
            KUMO_HIP(hipMemcpyAsync(hb, ts.t.nkept, static_cast<std::size_t>(nb) * 4, hipMemcpyDeviceToHost, s));
            if (stepwise) sync(ex);
            advance(static_cast<std::size_t>(count));
        }
        if (hst) {
            sync(ex);
            for (long long bb = 0; bb < nbatches; ++bb) {
                const int32_t* hb = hst + static_cast<std::size_t>(bb) * slot;
                const long long b0 = b_first + bb * cap;
                const int nb = static_cast<int>(std::min<long long>(cap, b_end - b0));
                ex.last_rescue_points += static_cast<long long>(hb[cap]) * kRescueTile;
                ex.last_rescue_flagged += hb[cap + 1];
                for (int b = 0; b < nb; ++b) {
                    const long long pts_b = std::min<long long>(bs, n - (b0 + b) * static_cast<long long>(bs));
                    const double m_b = hb[b];
                    if (backend == KUMO_HIP_AO_MO) {
                        const double mos = (need_a || closed ? mol->nalpha : 0) + (need_b && !closed ? mol->nbeta : 0);
                        flops += 2.0 * m_b * mos * panels * static_cast<double>(pts_b);
                    } else {
                        flops += 2.0 * m_b * m_b * (ao_order == 0 ? 1 : 4) * static_cast<double>(pts_b);
                    }
                }
            }
        }
        ex.last_flops += flops;
    }

    // ---------------------------------------------------------------- ESP
    if (plan.esp()) {
        mol->ensure_esp(s);
        std::size_t idx = 0;
        for (std::size_t i = 0; i < nprops; ++i)
            if (props[i].kind == kEsp) idx = pr.first_channel[i];
        if (use_far_field(ex, opt)) {
            far_field_range(ex, mol, pr, ps, opt, dout[idx], q0, q1, advance);
        } else {
            const long long chunk = 1 << 18;
            std::vector<EspGroup> groups = mol->esp_groups;
            const int nslices = esp_slices(groups, std::min(chunk, range));
            double* partial = grow<double>(ex.ws_esp, static_cast<std::size_t>(nslices) *
                                                          static_cast<std::size_t>(std::min(chunk, range)));
            for (long long p0 = q0; p0 < q1; p0 += chunk) {
                const long long cnt = std::min(chunk, q1 - p0);
                launch_esp(mol->d_esp_pairs.as<EspPairDev>(), groups, mol->d_esp_coef.as<double>(),
                           mol->d_esp_atoms.as<EspAtomDev>(), static_cast<int>(mol->atoms.size()), ex.boys_table, ps, p0,
                           cnt, partial, dout[idx] + p0, s);
                sync(ex);
                advance(static_cast<std::size_t>(cnt));
            }
        }
        for (std::size_t i = 0; i < nprops; ++i)
            if (props[i].kind == kEsp && pr.first_channel[i] != idx)
                KUMO_HIP(hipMemcpyAsync(dout[pr.first_channel[i]] + q0, dout[idx] + q0, range * sizeof(double),
                                        hipMemcpyDeviceToDevice, s));
    }

    // ---------------------------------------------------------- orbital energies
    // OrbitalEnergyFields::evaluate (src/orbital_energy_fields.hpp:195-217):
    // AO values without per-block screening, psi_i of every needed MO in FP64
    // on the MFMA kernel, the weighted sums per point, then the fields. The
    // masks use the smallest subnormal as threshold: an AO block is left out
    // only where every value of the block is zero (beyond the shell cutoff),
    // which changes no sum; the screening radius is the shell cutoff's.
    if (plan.orbital_energies) {
        mol->ensure_orbital_energy(s);
        const bool virt = plan.virtual_orbitals || plan.orbital_weighted_fukui;
        const int oslot1 = virt ? mol->oe_slots : mol->oe_occ_slots;
        const int groups = (oslot1 + 127) / 128;
        const auto weights = orbital_energy_weights(mol, plan, opt.fukui_delta);
        const long long b_first = q0 / bs, b_end = (q1 + bs - 1) / bs;
        const int cap = static_cast<int>(std::min<long long>({b_end - b_first, std::max(1ll, (1ll << 21) / bs), 8192ll}));
        // weights, then the sums [slot group][kOeSums][cap * bs]
        const std::size_t wsize = (weights.size() + 31) / 32 * 32;
        double* dw = grow<double>(ex.ws_oe, wsize + static_cast<std::size_t>(groups) * kOeSums * cap * bs);
        double* sums = dw + wsize;
        upload_async(ex, dw, weights.data(), weights.size());
        std::vector<FieldTask> tasks;
        for (std::size_t i = 0; i < nprops; ++i)
            if (is_orbital_energy(props[i].kind)) {
                FieldTask t{};
                t.kind = props[i].kind;
                t.component = props[i].component;
                t.out[0] = dout[pr.first_channel[i]];
                tasks.push_back(t);
            }
        FieldTask* dtasks = grow<FieldTask>(ex.ws_oe_tasks, tasks.size());
        upload_async(ex, dtasks, tasks.data(), tasks.size());
        const auto ts = tables(ex, cap, nao, nsh, mol->max_chunks);
        const double oe_thr = std::numeric_limits<double>::denorm_min();
        for (long long b0 = b_first; b0 < b_end; b0 += cap) {
            const int nb = static_cast<int>(std::min<long long>(cap, b_end - b0));
            const long long first = b0 * bs;
            const long long count = std::min<long long>(static_cast<long long>(nb) * bs, q1 - first);
            KeptTables kt = ts.t;
            kt.lists = 0;
            launch_masks(mv, ps, n, bs, b0, nb, 0, oe_thr, mol->screen_r2(oe_thr, s), kt, s);
            AoMoLaunch a{};
            a.mol = mv;
            a.ps = ps;
            a.mo = {mol->d_oe_op.as<double>(), nullptr, nullptr, nullptr, mol->operand_rows, mol->oe_slots,
                    mol->oe_slots};
            a.tables = ts.t;
            a.npoints = n;
            a.batch_first = first;
            a.batch_points = count;
            a.block_size = bs;
            a.first_block = b0;
            a.nblk = nb;
            a.order = 0;
            a.hessian = 0;
            a.precision = KUMO_HIP_FP64;
            a.slot0 = 0;
            a.slot1 = oslot1;
            a.tile_points = 16;
            a.moments = sums;
            a.weights = dw;
            record(ex, true);
            launch_ao_mo(a, s);
            record(ex, false);
            ++ex.last_launches;
            for (int g = 1; g < groups; ++g)
                launch_add_into(sums, sums + static_cast<long long>(g) * kOeSums * count, kOeSums * count, s);
            launch_orbital_energy_fields(sums, count, first, static_cast<int>(tasks.size()), dtasks, virt ? 1 : 0,
                                         plan.orbital_weighted_fukui ? 1 : 0, s);
            if (!ex.capture && opt.progress) sync(ex);
            advance(static_cast<std::size_t>(count));
        }
    }

    // ---------------------------------------------------------------- orbitals
    const auto run_orbitals = [&](const std::vector<std::size_t>& keys, int channel, int order) {
        std::vector<int> mos;
        for (auto i : keys)
            if (std::find(mos.begin(), mos.end(), props[i].orbital_index - 1) == mos.end())
                mos.push_back(props[i].orbital_index - 1);
        const bool beta = channel == 1 && mol->unrestricted;
        const double* coeff = beta ? mol->d_cb.as<double>() : mol->d_ca.as<double>();
        const int ldc = beta ? mol->nmo_b : mol->nmo_a;
        int32_t* dmo = grow<int32_t>(ex.ws_misc, mos.size());
        upload_async(ex, dmo, mos.data(), mos.size());
        const int NP = order == 0 ? 1 : (order == 1 ? 4 : 10);
        const long long chunk =
            std::max(1ll, std::min<long long>(range, (256ll << 20) / (8ll * NP * static_cast<long long>(mos.size()))));
        double* scratch = grow<double>(ex.ws_phi, static_cast<std::size_t>(mos.size()) * NP * chunk);
        for (long long p0 = q0; p0 < q1; p0 += chunk) {
            const long long cnt = std::min(chunk, q1 - p0);
            launch_orbitals(mv, ps, p0, cnt, coeff, ldc, dmo, static_cast<int>(mos.size()), order, scratch, cnt, s);
            for (auto i : keys) {
                const int k = static_cast<int>(std::find(mos.begin(), mos.end(), props[i].orbital_index - 1) - mos.begin());
                const int kind = props[i].kind;
                const int c0 = kind == kOrbital ? 0 : (kind == kOrbitalGradient ? 1 : 4);
                if ((order == 0) != (kind == kOrbital)) continue;
                for (int c = 0; c < channels(kind); ++c)
                    KUMO_HIP(hipMemcpyAsync(dout[pr.first_channel[i] + static_cast<std::size_t>(c)] + p0,
                                            scratch + (static_cast<long long>(k) * NP + c0 + c) * cnt, cnt * sizeof(double),
                                            hipMemcpyDeviceToDevice, s));
            }
            sync(ex);
        }
    };
    if (pr.orbital_pass) {
        for (int channel = 0; channel < 2; ++channel) {
            std::vector<std::size_t> values, derivs;
            int dorder = 0;
            for (std::size_t i = 0; i < nprops; ++i) {
                const auto& p = props[i];
                if (!is_orbital_property(p.kind) || on_grid_path(p, backend, grid) || p.channel != channel) continue;
                if (p.kind == kOrbital) {
                    values.push_back(i);
                } else {
                    derivs.push_back(i);
                    dorder = std::max(dorder, p.kind == kOrbitalGradient ? 1 : 2);
                }
            }
            if (!values.empty()) run_orbitals(values, channel, 0);
            if (!derivs.empty()) run_orbitals(derivs, channel, dorder);
        }
        advance(static_cast<std::size_t>(range));
    }

    // ---------------------------------------------------------------- grid path
    if (grid) {
        for (std::size_t i = 0; i < nprops; ++i) {
            const auto& p = props[i];
            if (!(on_grid_path(p, backend, grid) && p.kind == kDensity)) continue;
            // densities_blocked with one component (src/density.cpp:365-444).
            const long long b_first = q0 / bs, b_end = (q1 + bs - 1) / bs;
            const long long nblocks = b_end - b_first;
            const long long per_block = static_cast<long long>(nao) * bs;
            const int cap =
                static_cast<int>(std::max(1ll, std::min<long long>({nblocks, 8192ll, (256ll << 20) / (8 * per_block)})));
            const auto ts = tables(ex, cap, nao, nsh, mol->max_chunks);
            double* chm = grow<double>(ex.ws_moments, 2 * kMoments * static_cast<std::size_t>(cap) * bs);
            FieldTask t{};
            t.kind = kDensity;
            t.component = kTotal;
            t.out[0] = dout[pr.first_channel[i]];
            FieldTask* dtask = grow<FieldTask>(ex.ws_tasks, 1);
            upload_async(ex, dtask, &t, 1);
            for (long long b0 = b_first; b0 < b_end; b0 += cap) {
                const int nb = static_cast<int>(std::min<long long>(cap, b_end - b0));
                const long long first = b0 * bs;
                const long long count = std::min<long long>(static_cast<long long>(nb) * bs, q1 - first);
                launch_masks(mv, ps, n, bs, b0, nb, 0, thr, mol->screen_r2(thr, s), ts.t, s);
                AoPLaunch a{};
                a.mol = mv;
                a.ps = ps;
                a.tables = ts.t;
                a.p[0] = mol->p_matrix(p.component == kTotal ? 0 : (p.component == kDifference ? 1 : p.component == kAlpha ? 2 : 3), s);
                a.p[1] = nullptr;
                a.npoints = n;
                a.batch_first = first;
                a.batch_points = count;
                a.block_size = bs;
                a.first_block = b0;
                a.nblk = nb;
                a.order = 0;
                a.phi = grow<double>(ex.ws_phi, static_cast<std::size_t>(nb) * nao * bs);
                a.w = grow<double>(ex.ws_w, static_cast<std::size_t>(nb) * nao * 4 * bs);
                a.moments = chm;
                launch_fill(chm, 2 * kMoments * count, 0.0, s);
                launch_ao_p(a, s);
                const FieldsLaunch f = fields_launch(chm, 0, 0, count, first, 1, dtask);
                launch_fields(f, s);
                sync(ex);
                advance(static_cast<std::size_t>(count));
            }
        }
        for (int channel = 0; channel < 2; ++channel) {
            std::vector<std::size_t> keys;
            for (std::size_t i = 0; i < nprops; ++i)
                if (on_grid_path(props[i], backend, grid) && props[i].kind == kOrbital && props[i].channel == channel)
                    keys.push_back(i);
            if (keys.empty()) continue;
            run_orbitals(keys, channel, 0);
            advance(static_cast<std::size_t>(range));
        }
    }
    sync(ex);
}

void screening_masks(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const PointsArg& pts, int block_size,
                     int order, double threshold, uint8_t* kept) {
    Executor& ex = ctx->exec;
    if (block_size < 1 || block_size > 65536) invalid("block_size must be in [1, 65536]");
    if (order < 0 || order > 2) invalid("AO order must be 0, 1 or 2");
    const long long n = pts.grid ? validated_point_count(*pts.grid) : pts.n;
    const double* dpts = nullptr;
    if (!pts.grid) {
        dpts = grow<double>(ex.ws_points, static_cast<std::size_t>(3 * n));
        KUMO_HIP(hipMemcpyAsync(const_cast<double*>(dpts), pts.points, 3 * n * sizeof(double), hipMemcpyHostToDevice,
                                ex.stream));
    }
    const PointSource ps = make_source(pts.grid, dpts);
    const long long nblocks = (n + block_size - 1) / block_size;
    const int nao = mol->n, nsh = static_cast<int>(mol->basis.shells.size());
    const int cap = static_cast<int>(std::min<long long>(nblocks, 4096));
    const auto ts = tables(ex, std::max(1, cap), nao, nsh, mol->max_chunks);
    for (long long b0 = 0; b0 < nblocks; b0 += cap) {
        const int nb = static_cast<int>(std::min<long long>(cap, nblocks - b0));
        launch_masks(mol->view(), ps, n, block_size, b0, nb, order, threshold, const_cast<kumo_hip_molecule*>(mol)->screen_r2(threshold, ex.stream), ts.t, ex.stream);
        KUMO_HIP(hipMemcpyAsync(kept + b0 * nao, ts.t.kept, static_cast<std::size_t>(nb) * nao, hipMemcpyDeviceToHost,
                                ex.stream));
    }
    KUMO_HIP(hipStreamSynchronize(ex.stream));
}

void evaluate(kumo_hip_context* ctx, const kumo_hip_molecule* mol_c, const PointsArg& pts,
              const kumo_hip_property* props, std::size_t nprops, const kumo_hip_options& opt, double* const* outputs,
              int32_t* backend_used) {
    auto* mol = const_cast<kumo_hip_molecule*>(mol_c);
    Executor& ex = ctx->exec;
    hipStream_t s = ex.stream;
    const Prepared pr = prepare(mol, pts, props, nprops, opt, opt.device_outputs != 0);
    if (backend_used) *backend_used = pr.backend;
    ex.reset_stats();
    const long long n = pr.n;
    if (nprops == 0 || n == 0) return;
    std::vector<double*> dout(pr.nout);
    if (opt.device_outputs) {
        for (std::size_t c = 0; c < pr.nout; ++c)
            dout[c] = outputs[c];
    } else {
        ctx->ws_out.resize(pr.nout * static_cast<std::size_t>(n) * sizeof(double));
        for (std::size_t c = 0; c < pr.nout; ++c)
            dout[c] = ctx->ws_out.as<double>() + c * static_cast<std::size_t>(n);
    }
    const double* dpts = nullptr;
    if (!pr.grid) {
        if (opt.device_outputs) {
            dpts = pts.points;
        } else {
            dpts = grow<double>(ex.ws_points, static_cast<std::size_t>(3 * n));
            KUMO_HIP(hipMemcpyAsync(const_cast<double*>(dpts), pts.points, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        }
    }
    const PointSource ps = make_source(pts.grid, dpts);
    std::size_t done = 0;
    const std::size_t total = static_cast<std::size_t>(n) * static_cast<std::size_t>(pr.passes);
    const Advance advance = [&](std::size_t count) {
        done += count;
        if (opt.progress && opt.progress(done, total, opt.progress_user) != 0)
            throw Error(KUMO_HIP_ABORTED, "evaluation aborted by the progress callback");
    };
    run_range(ex, mol, pr, ps, opt, dout.data(), 0, n, advance);
    ex.collect_timing();
    if (!opt.device_outputs) {
        for (std::size_t c = 0; c < pr.nout; ++c)
            KUMO_HIP(hipMemcpyAsync(outputs[c], dout[c], n * sizeof(double), hipMemcpyDeviceToHost, s));
        KUMO_HIP(hipStreamSynchronize(s));
    }
}

void Executor::create(int device, const double* table) {
    KUMO_HIP(hipSetDevice(device));
    KUMO_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    boys_table = table;
}

void Executor::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& list : events) {
        for (auto& e : list) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        list.clear();
    }
    for (DeviceBuffer* b : {&ws_tables, &ws_moments, &ws_phi, &ws_w, &ws_points, &ws_tasks, &ws_misc,
                            &ws_rescue, &ws_rmom, &ws_rscale, &ws_oe, &ws_oe_tasks, &ws_esp})
        b->release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    if (hstats) (void)hipHostFree(hstats);
    hstats = nullptr;
    hstats_count = 0;
}

int32_t* Executor::host_stats(std::size_t count) {
    if (count > hstats_count) {
        // (called between evaluations: no copy into the old buffer is pending)
        if (hstats) (void)hipHostFree(hstats);
        hstats = nullptr;
        hstats_count = 0;
        KUMO_HIP(hipHostMalloc(reinterpret_cast<void**>(&hstats), count * sizeof(int32_t), hipHostMallocDefault));
        hstats_count = count;
    }
    return hstats;
}

void Executor::reset_stats() {
    events_used[0] = events_used[1] = events_used[2] = 0;
    last_kernel_ms = last_rescue_ms = last_fields_ms = 0.0;
    last_flops = 0.0;
    last_launches = last_rescue_points = last_rescue_flagged = last_fields_points = last_fields_bytes = 0;
}

void Executor::collect_timing() {
    double* dst[3] = {&last_kernel_ms, &last_rescue_ms, &last_fields_ms};
    for (int k = 0; k < 3; ++k) {
        double ms = 0.0;
        for (std::size_t e = 0; e < events_used[k]; ++e) {
            float t = 0.f;
            KUMO_HIP(hipEventElapsedTime(&t, events[k][e].first, events[k][e].second));
            ms += t;
        }
        *dst[k] = ms;
    }
}

} // namespace kumo
