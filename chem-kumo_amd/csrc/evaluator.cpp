// One evaluation on the device: the passes of Evaluator::Run
// (src/evaluator.cpp:556-802) in the reference order — point path, ESP,
// orbital pass, grid path — each over batches of whole screening blocks.
// Every screening block is evaluated by the same kernels whatever batch it
// falls in, so the results do not depend on the batch size or on other work
// on the device.
#include "plan.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace kumo {
namespace {

constexpr int kComp = 15;

struct Progress {
    const kumo_hip_options& opt;
    std::size_t done = 0, total = 0;
    void advance(std::size_t n) {
        done += n;
        if (opt.progress && opt.progress(done, total, opt.progress_user) != 0)
            throw Error(KUMO_HIP_ABORTED, "evaluation aborted by the progress callback");
    }
};

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
}

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

struct Ws {
    kumo_hip_context* ctx;
    template <typename T> T* get(DeviceBuffer& b, std::size_t count) {
        b.resize(count * sizeof(T));
        return b.as<T>();
    }
};

struct TableSet {
    KeptTables t;
    long long nblk_capacity = 0;
};

TableSet tables(kumo_hip_context* ctx, int nblk, int nao, int nsh) {
    const std::size_t b = static_cast<std::size_t>(nblk);
    const std::size_t bytes_kept = b * nao, bytes_list = 4 * b * nao, bytes_n = 4 * b, bytes_ks = 4 * b * nsh,
                      bytes_kb = 4 * b * nsh, bytes_kr = b * nsh, bytes_cb = 4 * b * (nsh + 1), bytes_nc = 4 * b;
    const auto al = [](std::size_t x) { return (x + 255) / 256 * 256; };
    const std::size_t total = al(bytes_kept) + al(bytes_list) + al(bytes_n) + al(bytes_ks) + al(bytes_kb) + al(bytes_kr) +
                              al(bytes_cb) + al(bytes_nc);
    ctx->ws_tables.resize(total);
    char* p = ctx->ws_tables.as<char>();
    TableSet s;
    s.t.kept = reinterpret_cast<uint8_t*>(p);
    p += al(bytes_kept);
    s.t.kept_list = reinterpret_cast<int32_t*>(p);
    p += al(bytes_list);
    s.t.nkept = reinterpret_cast<int32_t*>(p);
    p += al(bytes_n);
    s.t.kept_shell = reinterpret_cast<int32_t*>(p);
    p += al(bytes_ks);
    s.t.kept_bits = reinterpret_cast<uint32_t*>(p);
    p += al(bytes_kb);
    s.t.kept_row = reinterpret_cast<uint8_t*>(p);
    p += al(bytes_kr);
    s.t.chunk_begin = reinterpret_cast<int32_t*>(p);
    p += al(bytes_cb);
    s.t.nchunks = reinterpret_cast<int32_t*>(p);
    s.nblk_capacity = nblk;
    return s;
}

PointSource make_source(const PointsArg& pts, const double* device_points) {
    PointSource ps{};
    if (pts.grid) {
        for (int i = 0; i < 3; ++i)
            ps.origin[i] = pts.grid->origin[i];
        for (int i = 0; i < 9; ++i)
            ps.steps[i] = pts.grid->steps[i];
        ps.ny = pts.grid->counts[1];
        ps.nz = pts.grid->counts[2];
        ps.points = nullptr;
    } else {
        ps.ny = ps.nz = 1;
        ps.points = device_points;
    }
    return ps;
}

hipEvent_t next_event(kumo_hip_context* ctx, bool start) {
    if (ctx->events_used == ctx->events.size()) {
        hipEvent_t a, b;
        KUMO_HIP(hipEventCreate(&a));
        KUMO_HIP(hipEventCreate(&b));
        ctx->events.push_back({a, b});
    }
    auto& pr = ctx->events[ctx->events_used];
    if (!start) ++ctx->events_used;
    return start ? pr.first : pr.second;
}

} // namespace

// Constants of the derived fields computed with the host libm exactly as the
// reference computes them.
FieldsLaunch fields_launch(const double* moments, long long n, long long out_offset, int ntasks, const FieldTask* tasks) {
    constexpr double pi = 3.141592653589793;
    FieldsLaunch f{moments, n, out_offset, ntasks, tasks, 0.3 * std::pow(6.0 * pi * pi, 2.0 / 3.0),
                   2.0 * std::cbrt(3.0 * pi * pi)};
    return f;
}

void screening_masks(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const PointsArg& pts, int block_size,
                     int order, double threshold, uint8_t* kept) {
    if (block_size < 1 || block_size > 65536) invalid("block_size must be in [1, 65536]");
    if (order < 0 || order > 2) invalid("AO order must be 0, 1 or 2");
    const long long n = pts.grid ? validated_point_count(*pts.grid) : pts.n;
    const double* dpts = nullptr;
    if (!pts.grid) {
        ctx->ws_points.resize(static_cast<std::size_t>(std::max(1ll, 3 * n)) * sizeof(double));
        KUMO_HIP(hipMemcpyAsync(ctx->ws_points.get(), pts.points, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        dpts = ctx->ws_points.as<double>();
    }
    const PointSource ps = make_source(pts, dpts);
    const long long nblocks = (n + block_size - 1) / block_size;
    const int nao = mol->n, nsh = static_cast<int>(mol->basis.shells.size());
    const int cap = static_cast<int>(std::min<long long>(nblocks, 4096));
    const auto ts = tables(ctx, std::max(1, cap), nao, nsh);
    for (long long b0 = 0; b0 < nblocks; b0 += cap) {
        const int nb = static_cast<int>(std::min<long long>(cap, nblocks - b0));
        launch_masks(mol->view(), ps, n, block_size, b0, nb, order, threshold, ts.t, ctx->stream);
        KUMO_HIP(hipMemcpyAsync(kept + b0 * nao, ts.t.kept, static_cast<std::size_t>(nb) * nao, hipMemcpyDeviceToHost,
                                ctx->stream));
    }
    KUMO_HIP(hipStreamSynchronize(ctx->stream));
}

void evaluate(kumo_hip_context* ctx, const kumo_hip_molecule* mol_c, const PointsArg& pts,
              const kumo_hip_property* props, std::size_t nprops, const kumo_hip_options& opt, double* const* outputs,
              int32_t* backend_used) {
    auto* mol = const_cast<kumo_hip_molecule*>(mol_c);
    hipStream_t s = ctx->stream;
    validate_options(opt, mol);
    const bool grid = pts.grid != nullptr;
    long long n = 0;
    if (grid) {
        n = validated_point_count(*pts.grid);
    } else {
        n = pts.n;
        if (n < 0) invalid("invalid point count");
        if (!opt.device_outputs)
            for (long long q = 0; q < 3 * n; ++q)
                if (!std::isfinite(pts.points[q])) invalid("evaluation points must be finite");
    }
    const Plan plan = make_plan(props, nprops);
    const int backend = choose_backend(opt.backend, mol->eligible, plan);
    if (backend_used) *backend_used = backend;
    for (std::size_t i = 0; i < nprops; ++i) {
        const auto& p = props[i];
        if (is_orbital_property(p.kind)) {
            const int nmo = (p.channel == 1 && mol->unrestricted) ? mol->nmo_b : mol->nmo_a;
            if (nmo == 0) runtime("FCHK has no MO coefficients");
            if (p.orbital_index > nmo) out_of_range("MO index outside available coefficient range");
        }
    }
    ctx->events_used = 0;
    ctx->last_flops = ctx->last_padded_flops = 0.0;
    ctx->last_launches = 0;
    if (nprops == 0 || n == 0) return;

    // Outputs: device pointers per channel (staging buffers for host outputs).
    std::size_t nout = 0;
    for (std::size_t i = 0; i < nprops; ++i)
        nout += static_cast<std::size_t>(channels(props[i].kind));
    std::vector<double*> dout(nout);
    if (opt.device_outputs) {
        for (std::size_t c = 0; c < nout; ++c)
            dout[c] = outputs[c];
    } else {
        ctx->ws_out.resize(nout * static_cast<std::size_t>(n) * sizeof(double));
        for (std::size_t c = 0; c < nout; ++c)
            dout[c] = ctx->ws_out.as<double>() + c * static_cast<std::size_t>(n);
    }
    std::vector<std::size_t> first_channel(nprops);
    {
        std::size_t c = 0;
        for (std::size_t i = 0; i < nprops; ++i) {
            first_channel[i] = c;
            c += static_cast<std::size_t>(channels(props[i].kind));
        }
    }
    // Points on the device.
    const double* dpts = nullptr;
    if (!grid) {
        if (opt.device_outputs) {
            dpts = pts.points;
        } else {
            ctx->ws_points.resize(static_cast<std::size_t>(3 * n) * sizeof(double));
            KUMO_HIP(hipMemcpyAsync(ctx->ws_points.get(), pts.points, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
            dpts = ctx->ws_points.as<double>();
        }
    }
    const PointSource ps = make_source(pts, dpts);
    const MolView mv = mol->view();
    const int nao = mol->n, nsh = static_cast<int>(mol->basis.shells.size());
    const int bs = opt.block_size;
    const double thr = opt.screening_threshold;

    // Progress over all passes (src/evaluator.cpp:597-603).
    std::vector<int> needed;
    for (int c = 0; c < 4; ++c)
        if (uses_point_path(plan.c[static_cast<std::size_t>(c)], backend, grid)) needed.push_back(c);
    bool any_orbital_pass = false, esp = plan.esp();
    int grid_density_passes = 0, grid_orbital_passes = 0;
    bool orb_ch[2] = {false, false};
    for (std::size_t i = 0; i < nprops; ++i) {
        const auto& p = props[i];
        if (is_orbital_property(p.kind) && !on_grid_path(p, backend, grid)) any_orbital_pass = true;
        if (on_grid_path(p, backend, grid) && p.kind == kDensity) ++grid_density_passes;
        if (on_grid_path(p, backend, grid) && p.kind == kOrbital) orb_ch[p.channel] = true;
    }
    grid_orbital_passes = (orb_ch[0] ? 1 : 0) + (orb_ch[1] ? 1 : 0);
    Progress progress{opt};
    progress.total = static_cast<std::size_t>(n) *
                     static_cast<std::size_t>((needed.empty() ? 0 : 1) + (esp ? 1 : 0) + (any_orbital_pass ? 1 : 0) +
                                              grid_density_passes + grid_orbital_passes);

    // ---------------------------------------------------------------- point pass
    if (!needed.empty()) {
        const int ao_order = grid ? plan.max_order : std::max(1, plan.max_order);
        bool full_h = false;
        for (int c = 0; c < 4; ++c)
            full_h = full_h || plan.c[static_cast<std::size_t>(c)].full_hessian();
        const long long nblocks = (n + bs - 1) / bs;
        long long target = 1ll << 21;
        int cap = static_cast<int>(std::min<long long>({nblocks, std::max(1ll, target / bs), 8192ll}));
        if (backend == KUMO_HIP_AO_P) {
            const long long per_block = static_cast<long long>(nao) * (ao_order == 0 ? 1 : (ao_order == 1 ? 4 : 10)) * bs;
            cap = static_cast<int>(std::max(1ll, std::min<long long>(cap, (512ll << 20) / (8 * per_block))));
        }
        const auto ts = tables(ctx, cap, nao, nsh);
        const long long batch_max = static_cast<long long>(cap) * bs;
        // channel moments [groups][2][13][batch] and component moments [4][15][batch]
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
        double* chm = Ws{ctx}.get<double>(ctx->ws_moments, static_cast<std::size_t>(groups) * 2 * kMoments * batch_max);
        double* cmp = Ws{ctx}.get<double>(ctx->ws_comp, static_cast<std::size_t>(4) * kComp * batch_max);
        // Field tasks of the point path.
        std::vector<FieldTask> tasks;
        for (std::size_t i = 0; i < nprops; ++i) {
            const auto& p = props[i];
            if (!from_point_path(p, backend, grid)) continue;
            FieldTask t{};
            t.kind = p.kind;
            t.component = p.component;
            for (int c = 0; c < channels(p.kind); ++c)
                t.out[c] = dout[first_channel[i] + static_cast<std::size_t>(c)];
            tasks.push_back(t);
        }
        FieldTask* dtasks = Ws{ctx}.get<FieldTask>(ctx->ws_tasks, std::max<std::size_t>(1, tasks.size()));
        KUMO_HIP(hipMemcpyAsync(dtasks, tasks.data(), tasks.size() * sizeof(FieldTask), hipMemcpyHostToDevice, s));
        std::vector<int32_t> nkept(static_cast<std::size_t>(cap));
        double flops = 0.0;
        const int panels = ao_order <= 0 ? 1 : (ao_order == 1 ? 4 : (full_h ? 10 : 7));
        for (long long b0 = 0; b0 < nblocks; b0 += cap) {
            const int nb = static_cast<int>(std::min<long long>(cap, nblocks - b0));
            const long long first = b0 * bs;
            const long long count = std::min<long long>(static_cast<long long>(nb) * bs, n - first);
            launch_masks(mv, ps, n, bs, b0, nb, std::max(0, ao_order), thr, ts.t, s);
            if (backend == KUMO_HIP_AO_MO) {
                AoMoLaunch a{};
                a.mol = mv;
                a.ps = ps;
                a.mo = {mol->d_mo64.as<double>(), mol->d_mo32.as<float>(), mol->d_mobf.as<uint16_t>(), mol->slots,
                        mol->alpha_slots};
                a.tables = ts.t;
                a.npoints = n;
                a.batch_first = first;
                a.batch_points = count;
                a.block_size = bs;
                a.first_block = b0;
                a.nblk = nb;
                a.order = std::max(0, ao_order);
                a.hessian = ao_order == 2 ? (full_h ? 2 : 1) : 0;
                a.precision = opt.precision;
                a.slot0 = slot0;
                a.slot1 = slot1;
                a.moments = chm;
                KUMO_HIP(hipEventRecord(next_event(ctx, true), s));
                launch_ao_mo(a, s);
                KUMO_HIP(hipEventRecord(next_event(ctx, false), s));
                ++ctx->last_launches;
                if (groups > 1) {
                    // Sum the slot groups into group 0 (fixed order).
                    for (int g = 1; g < groups; ++g)
                        launch_add_into(chm, chm + static_cast<long long>(g) * 2 * kMoments * count, 2 * kMoments * count, s);
                }
                launch_combine(chm, cmp, count, 1, ao_order, closed ? 1 : 0, s);
            } else {
                AoPLaunch a{};
                a.mol = mv;
                a.ps = ps;
                a.tables = ts.t;
                const bool tn = std::any_of(needed.begin(), needed.end(), [](int c) { return c != kDifference; });
                const bool sn = std::any_of(needed.begin(), needed.end(), [](int c) { return c != kTotal; });
                a.p[0] = tn ? mol->p_matrix(0) : nullptr;
                a.p[1] = sn ? mol->p_matrix(1) : nullptr;
                a.npoints = n;
                a.batch_first = first;
                a.batch_points = count;
                a.block_size = bs;
                a.first_block = b0;
                a.nblk = nb;
                a.order = ao_order;
                const int pan = ao_order == 0 ? 1 : (ao_order == 1 ? 4 : 10);
                a.phi = Ws{ctx}.get<double>(ctx->ws_phi, static_cast<std::size_t>(nb) * nao * pan * bs);
                a.w = Ws{ctx}.get<double>(ctx->ws_w, static_cast<std::size_t>(nb) * nao * 4 * bs);
                a.moments = chm;
                launch_fill(chm, 2 * kMoments * count, 0.0, s);
                KUMO_HIP(hipEventRecord(next_event(ctx, true), s));
                launch_ao_p(a, s);
                KUMO_HIP(hipEventRecord(next_event(ctx, false), s));
                ++ctx->last_launches;
                launch_combine(chm, cmp, count, 0, ao_order, 0, s);
            }
            const FieldsLaunch f = fields_launch(cmp, count, first, static_cast<int>(tasks.size()), dtasks);
            launch_fields(f, s);
            // Algorithmic FLOPs from the kept counts.
            KUMO_HIP(hipMemcpyAsync(nkept.data(), ts.t.nkept, static_cast<std::size_t>(nb) * 4, hipMemcpyDeviceToHost, s));
            KUMO_HIP(hipStreamSynchronize(s));
            for (int b = 0; b < nb; ++b) {
                const long long pts_b = std::min<long long>(bs, n - (b0 + b) * static_cast<long long>(bs));
                const double m_b = nkept[static_cast<std::size_t>(b)];
                if (backend == KUMO_HIP_AO_MO) {
                    const double mos = (need_a || closed ? mol->nalpha : 0) + (need_b && !closed ? mol->nbeta : 0);
                    flops += 2.0 * m_b * mos * panels * static_cast<double>(pts_b);
                } else {
                    flops += 2.0 * m_b * m_b * (ao_order == 0 ? 1 : 4) * static_cast<double>(pts_b);
                }
            }
            progress.advance(static_cast<std::size_t>(count));
        }
        ctx->last_flops = flops;
    }

    // ---------------------------------------------------------------- ESP
    if (esp) {
        mol->ensure_esp();
        std::size_t idx = 0;
        for (std::size_t i = 0; i < nprops; ++i)
            if (props[i].kind == kEsp) idx = first_channel[i];
        const long long chunk = 1 << 18;
        for (long long q0 = 0; q0 < n; q0 += chunk) {
            const long long cnt = std::min(chunk, n - q0);
            launch_esp(mol->d_esp_pairs.as<EspPairDev>(), mol->esp_npairs, mol->d_esp_coef.as<double>(),
                       mol->d_esp_atoms.as<EspAtomDev>(), static_cast<int>(mol->atoms.size()), ctx->boys_table.as<double>(),
                       ps, q0, cnt, dout[idx] + q0, s);
            KUMO_HIP(hipStreamSynchronize(s));
            progress.advance(static_cast<std::size_t>(cnt));
        }
        // Duplicated ESP keys share the value.
        for (std::size_t i = 0; i < nprops; ++i)
            if (props[i].kind == kEsp && first_channel[i] != idx)
                KUMO_HIP(hipMemcpyAsync(dout[first_channel[i]], dout[idx], n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }

    // ---------------------------------------------------------------- orbitals
    const auto run_orbitals = [&](const std::vector<std::size_t>& keys, int channel, int order) {
        // distinct MO indices of this channel
        std::vector<int> mos;
        for (auto i : keys)
            if (std::find(mos.begin(), mos.end(), props[i].orbital_index - 1) == mos.end())
                mos.push_back(props[i].orbital_index - 1);
        const bool beta = channel == 1 && mol->unrestricted;
        const double* coeff = beta ? mol->d_cb.as<double>() : mol->d_ca.as<double>();
        const int ldc = beta ? mol->nmo_b : mol->nmo_a;
        int32_t* dmo = Ws{ctx}.get<int32_t>(ctx->ws_misc, mos.size());
        KUMO_HIP(hipMemcpyAsync(dmo, mos.data(), mos.size() * 4, hipMemcpyHostToDevice, s));
        const int NP = order == 0 ? 1 : (order == 1 ? 4 : 10);
        const long long chunk = std::max(1ll, std::min<long long>(n, (256ll << 20) / (8ll * NP * static_cast<long long>(mos.size()))));
        double* scratch = Ws{ctx}.get<double>(ctx->ws_phi, static_cast<std::size_t>(mos.size()) * NP * chunk);
        for (long long q0 = 0; q0 < n; q0 += chunk) {
            const long long cnt = std::min(chunk, n - q0);
            launch_orbitals(mv, ps, q0, cnt, coeff, ldc, dmo, static_cast<int>(mos.size()), order, scratch, cnt, s);
            for (auto i : keys) {
                const int k = static_cast<int>(std::find(mos.begin(), mos.end(), props[i].orbital_index - 1) - mos.begin());
                const int kind = props[i].kind;
                const int c0 = kind == kOrbital ? 0 : (kind == kOrbitalGradient ? 1 : 4);
                const int nc = channels(kind);
                if (order == 0 && kind != kOrbital) continue;
                if (order > 0 && kind == kOrbital) continue;
                for (int c = 0; c < nc; ++c)
                    KUMO_HIP(hipMemcpyAsync(dout[first_channel[i] + static_cast<std::size_t>(c)] + q0,
                                            scratch + (static_cast<long long>(k) * NP + c0 + c) * cnt, cnt * sizeof(double),
                                            hipMemcpyDeviceToDevice, s));
            }
            KUMO_HIP(hipStreamSynchronize(s));
        }
    };
    if (any_orbital_pass) {
        for (int channel = 0; channel < 2; ++channel) {
            std::vector<std::size_t> values, derivs;
            int dorder = 0;
            for (std::size_t i = 0; i < nprops; ++i) {
                const auto& p = props[i];
                if (!is_orbital_property(p.kind) || on_grid_path(p, backend, grid) || p.channel != channel) continue;
                if (p.kind == kOrbital) values.push_back(i);
                else {
                    derivs.push_back(i);
                    dorder = std::max(dorder, p.kind == kOrbitalGradient ? 1 : 2);
                }
            }
            if (!values.empty()) run_orbitals(values, channel, 0);
            if (!derivs.empty()) run_orbitals(derivs, channel, dorder);
        }
        progress.advance(static_cast<std::size_t>(n));
    }

    // ---------------------------------------------------------------- grid path
    if (grid) {
        for (std::size_t i = 0; i < nprops; ++i) {
            const auto& p = props[i];
            if (!(on_grid_path(p, backend, grid) && p.kind == kDensity)) continue;
            // densities_blocked with one component (src/density.cpp:365-444).
            const long long nblocks = (n + bs - 1) / bs;
            const long long per_block = static_cast<long long>(nao) * bs;
            const int cap = static_cast<int>(std::max(1ll, std::min<long long>({nblocks, 8192ll, (256ll << 20) / (8 * per_block)})));
            const auto ts = tables(ctx, cap, nao, nsh);
            double* chm = Ws{ctx}.get<double>(ctx->ws_moments, 2 * kMoments * static_cast<std::size_t>(cap) * bs);
            double* cmp = Ws{ctx}.get<double>(ctx->ws_comp, 4 * kComp * static_cast<std::size_t>(cap) * bs);
            FieldTask t{};
            t.kind = kDensity;
            t.component = kTotal;
            t.out[0] = dout[first_channel[i]];
            FieldTask* dtask = Ws{ctx}.get<FieldTask>(ctx->ws_tasks, 1);
            KUMO_HIP(hipMemcpyAsync(dtask, &t, sizeof(t), hipMemcpyHostToDevice, s));
            for (long long b0 = 0; b0 < nblocks; b0 += cap) {
                const int nb = static_cast<int>(std::min<long long>(cap, nblocks - b0));
                const long long first = b0 * bs;
                const long long count = std::min<long long>(static_cast<long long>(nb) * bs, n - first);
                launch_masks(mv, ps, n, bs, b0, nb, 0, thr, ts.t, s);
                AoPLaunch a{};
                a.mol = mv;
                a.ps = ps;
                a.tables = ts.t;
                a.p[0] = mol->p_matrix(p.component == kTotal ? 0 : (p.component == kDifference ? 1 : p.component == kAlpha ? 2 : 3));
                a.p[1] = nullptr;
                a.npoints = n;
                a.batch_first = first;
                a.batch_points = count;
                a.block_size = bs;
                a.first_block = b0;
                a.nblk = nb;
                a.order = 0;
                a.phi = Ws{ctx}.get<double>(ctx->ws_phi, static_cast<std::size_t>(nb) * nao * bs);
                a.w = Ws{ctx}.get<double>(ctx->ws_w, static_cast<std::size_t>(nb) * nao * 4 * bs);
                a.moments = chm;
                launch_fill(chm, 2 * kMoments * count, 0.0, s);
                launch_ao_p(a, s);
                launch_combine(chm, cmp, count, 0, 0, 1, s);
                const FieldsLaunch f = fields_launch(cmp, count, first, 1, dtask);
                launch_fields(f, s);
                KUMO_HIP(hipStreamSynchronize(s));
                progress.advance(static_cast<std::size_t>(count));
            }
        }
        for (int channel = 0; channel < 2; ++channel) {
            if (!orb_ch[channel]) continue;
            std::vector<std::size_t> keys;
            for (std::size_t i = 0; i < nprops; ++i)
                if (on_grid_path(props[i], backend, grid) && props[i].kind == kOrbital && props[i].channel == channel)
                    keys.push_back(i);
            run_orbitals(keys, channel, 0);
            progress.advance(static_cast<std::size_t>(n));
        }
    }

    // Duplicated property keys share their values (the reference evaluates a
    // repeated property once; ours are written per key by the passes above).
    KUMO_HIP(hipStreamSynchronize(s));
    // Kernel timing of the contraction launches.
    double ms = 0.0;
    for (std::size_t e = 0; e < ctx->events_used; ++e) {
        float t = 0.f;
        KUMO_HIP(hipEventElapsedTime(&t, ctx->events[e].first, ctx->events[e].second));
        ms += t;
    }
    ctx->last_kernel_ms = ms;
    ctx->last_padded_flops = ctx->last_flops;
    if (!opt.device_outputs) {
        for (std::size_t c = 0; c < nout; ++c)
            KUMO_HIP(hipMemcpyAsync(outputs[c], dout[c], n * sizeof(double), hipMemcpyDeviceToHost, s));
        KUMO_HIP(hipStreamSynchronize(s));
    }
}

} // namespace kumo
