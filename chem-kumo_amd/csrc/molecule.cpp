// Molecule loading and residency (src/molecule.cpp:20-185,
// src/backend_selection.hpp:43-78, src/mo_coefficients.hpp:21-34) and the
// host construction of the ESP primitive pairs
// (src/electrostatic_potential.cpp:117-229).
#include "fchk.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numbers>

namespace kumo {
namespace {

double frobenius(const std::vector<double>& m) {
    double s = 0.0;
    for (double v : m)
        s += v * v;
    return std::sqrt(s);
}

// P_alpha = C_a,occ C_a,occ^T and likewise for beta; checks total and spin
// against the FCHK matrices (validate_orbital_backend).
void assess(kumo_hip_molecule& m) {
    const int n = m.n;
    const auto& cb = m.unrestricted ? m.Cb : m.Ca;
    const int nmob = m.unrestricted ? m.nmo_b : m.nmo_a;
    if (m.Ca.empty() || m.nalpha > m.nmo_a || m.nbeta > nmob || (!m.unrestricted && m.nalpha != m.nbeta)) return;
    std::vector<double> at(static_cast<std::size_t>(m.nalpha) * n), bt(static_cast<std::size_t>(m.nbeta) * n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < m.nalpha; ++k)
            at[static_cast<std::size_t>(k) * n + i] = m.Ca[static_cast<std::size_t>(i) * m.nmo_a + k];
        for (int k = 0; k < m.nbeta; ++k)
            bt[static_cast<std::size_t>(k) * n + i] = cb[static_cast<std::size_t>(i) * nmob + k];
    }
    const std::size_t nn = static_cast<std::size_t>(n) * n;
    std::vector<double> pa(nn, 0.0), pb(nn, 0.0);
    for (int k = 0; k < m.nalpha; ++k) {
        const double* c = at.data() + static_cast<std::size_t>(k) * n;
        for (int i = 0; i < n; ++i) {
            const double ci = c[i];
            double* row = pa.data() + static_cast<std::size_t>(i) * n;
            for (int j = 0; j < n; ++j)
                row[j] += ci * c[j];
        }
    }
    for (int k = 0; k < m.nbeta; ++k) {
        const double* c = bt.data() + static_cast<std::size_t>(k) * n;
        for (int i = 0; i < n; ++i) {
            const double ci = c[i];
            double* row = pb.data() + static_cast<std::size_t>(i) * n;
            for (int j = 0; j < n; ++j)
                row[j] += ci * c[j];
        }
    }
    std::vector<double> rt(nn), rs(nn), et(nn), es(nn);
    for (std::size_t q = 0; q < nn; ++q) {
        rt[q] = pa[q] + pb[q];
        rs[q] = pa[q] - pb[q];
        et[q] = m.P[q] - rt[q];
        es[q] = m.S[q] - rs[q];
    }
    const double te = frobenius(et), se = frobenius(es);
    const double ts = std::max(frobenius(m.P), frobenius(rt));
    const double ss = std::max({frobenius(m.S), frobenius(rs), ts});
    m.total_rel = ts > 0 ? te / ts : te;
    m.spin_rel = ss > 0 ? se / ss : se;
    m.eligible = te <= 1e-10 + 1e-8 * ts && se <= 1e-10 + 1e-8 * ss;
}

// bf16 round to nearest even from float.
uint16_t bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
double bf16_value(uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}


// MO operand table rows of MO columns coef(ao, slot), slots 0..ns-1: per AO
// block the Cartesian rows g[k] (T^T C)[k] (pure transform and component
// factor folded in), then for pure blocks the plain pure rows C[r].
template <typename Coef>
std::vector<double> operand_table(const PreparedBasis& b, std::size_t ns, const Coef& coef) {
    std::vector<double> op(static_cast<std::size_t>(b.operand_rows) * ns, 0.0);
    for (std::size_t blk = 0; blk < b.blocks.size(); ++blk) {
        const HostBlock& hb = b.blocks[blk];
        const DevShell& d = b.shells[blk];
        const int nc = hb.ncart();
        for (int k = 0; k < nc; ++k) {
            const double gk = component_factor(hb.powers[0], hb.powers[static_cast<std::size_t>(k)]);
            double* row = op.data() + static_cast<std::size_t>(d.crow + k) * ns;
            for (std::size_t slot = 0; slot < ns; ++slot) {
                double v;
                if (hb.pure.empty()) {
                    v = coef(hb.offset + k, slot);
                } else {
                    v = 0.0;
                    for (int r = 0; r < hb.count; ++r)
                        v += hb.pure[static_cast<std::size_t>(r * nc + k)] * coef(hb.offset + r, slot);
                }
                row[slot] = gk * v;
            }
        }
        if (!hb.pure.empty())
            for (int r = 0; r < hb.count; ++r) {
                double* row = op.data() + static_cast<std::size_t>(d.prow + r) * ns;
                for (std::size_t slot = 0; slot < ns; ++slot)
                    row[slot] = coef(hb.offset + r, slot);
            }
    }
    return op;
}

void upload_device(kumo_hip_molecule& m, hipStream_t s) {
    const auto& b = m.basis;
    m.d_shells = upload(b.shells.data(), b.shells.size(), s);
    m.d_exps = upload(b.exps.data(), b.exps.size(), s);
    m.d_weights = upload(b.weights.data(), b.weights.size(), s);
    std::vector<double> tdense(static_cast<std::size_t>((kMaxL + 1) * kMaxPure * kMaxCart), 0.0);
    for (int l = 2; l <= kMaxL; ++l) {
        const auto t = gaussian_cartesian_to_pure(l);
        const int nc = (l + 1) * (l + 2) / 2;
        for (int r = 0; r < 2 * l + 1; ++r)
            for (int k = 0; k < nc; ++k)
                tdense[static_cast<std::size_t>((l * kMaxPure + r) * kMaxCart + k)] = t[static_cast<std::size_t>(r * nc + k)];
    }
    m.d_tdense = upload(tdense.data(), tdense.size(), s);
    m.d_p[0] = upload(m.P.data(), m.P.size(), s);
    m.d_p[1] = upload(m.S.data(), m.S.size(), s);
    if (!m.Ca.empty()) m.d_ca = upload(m.Ca.data(), m.Ca.size(), s);
    if (!m.Cb.empty()) m.d_cb = upload(m.Cb.data(), m.Cb.size(), s);
    // g[k] T[r][k]: the pure transform applied to unscaled Cartesian components.
    std::vector<double> tg(tdense.size(), 0.0);
    for (int l = 2; l <= kMaxL; ++l) {
        const auto pw = gaussian_cartesian_order(l);
        const int nc = (l + 1) * (l + 2) / 2;
        for (int r = 0; r < 2 * l + 1; ++r)
            for (int k = 0; k < nc; ++k) {
                const std::size_t q = static_cast<std::size_t>((l * kMaxPure + r) * kMaxCart + k);
                tg[q] = tdense[q] * component_factor(pw[0], pw[static_cast<std::size_t>(k)]);
            }
    }
    m.d_tg = upload(tg.data(), tg.size(), s);
    std::vector<double> gk(static_cast<std::size_t>((kMaxL + 1) * kMaxCart), 0.0);
    for (int l = 0; l <= kMaxL; ++l) {
        const auto pw = gaussian_cartesian_order(l);
        for (std::size_t k = 0; k < pw.size(); ++k)
            gk[static_cast<std::size_t>(l * kMaxCart) + k] = component_factor(pw[0], pw[k]);
    }
    m.d_gk = upload(gk.data(), gk.size(), s);
    if (b.shells.size() >= 65536) runtime("more than 65535 AO blocks are not supported by the HIP backend");
    m.d_lorder = upload(b.lorder.data(), b.lorder.size(), s);
    std::vector<ShellRec> rec(b.shells.size());
    for (std::size_t i = 0; i < b.shells.size(); ++i) {
        const DevShell& d = b.shells[i];
        ShellRec& r = rec[i];
        std::memset(&r, 0, sizeof(r));
        for (int a = 0; a < 3; ++a)
            r.o[a] = d.o[a];
        r.cutoff_r2 = d.cutoff_r2;
        r.l = d.l;
        r.nprim = d.nprim;
        r.prim = d.prim;
        for (int p = 0; p < d.nprim && p < kRecPrim; ++p) {
            const double a = b.exps[static_cast<std::size_t>(d.prim + p)];
            const double w = b.radw[static_cast<std::size_t>(d.prim + p)];
            r.e[p] = a;
            r.w[p] = w;
            // |w| max(1, 4 a^2) exp(-a r^2) < 1e-40 beyond this r^2
            const double lw = w != 0.0 ? std::log(std::fabs(w) * std::max(1.0, 4.0 * a * a)) : -1e300;
            r.skip[p] = a > 0.0 ? std::max(0.0, lw + 92.1) / a : 1e300;
        }
    }
    m.d_rec = upload(rec.data(), rec.size(), s);
    m.d_radw = upload(b.radw.data(), b.radw.size(), s);
    // Occupied-MO operand: alpha slots [0, A16), beta slots [A16, A16 + B16)
    // (closed shells contract the alpha channel only). An unrestricted
    // molecule keeps at least 16 (zero) slots per channel, so a channel with
    // no electrons is contracted as zeros and gives exactly zero moments.
    const bool closed = !m.unrestricted;
    const int a16 = closed ? (m.nalpha + 15) / 16 * 16 : std::max(16, (m.nalpha + 15) / 16 * 16);
    const int b16 = closed ? 0 : std::max(16, (m.nbeta + 15) / 16 * 16);
    m.alpha_slots = a16;
    m.slots = std::max(16, a16 + b16);
    m.operand_rows = b.operand_rows;
    m.max_chunks = max_chunks_bound(b.shells.data(), static_cast<int>(b.shells.size()));
    if (!m.Ca.empty() && m.nalpha <= m.nmo_a && (closed || m.nbeta <= m.nmo_b)) {
        const std::size_t ns = static_cast<std::size_t>(m.slots);
        const std::size_t nr = static_cast<std::size_t>(b.operand_rows);
        // Occupied coefficients of AO i in slot order.
        const auto coef = [&](int i, std::size_t slot) -> double {
            if (slot < static_cast<std::size_t>(m.nalpha))
                return m.Ca[static_cast<std::size_t>(i) * static_cast<std::size_t>(m.nmo_a) + slot];
            if (!closed && slot >= static_cast<std::size_t>(a16) && slot < static_cast<std::size_t>(a16 + m.nbeta))
                return m.Cb[static_cast<std::size_t>(i) * static_cast<std::size_t>(m.nmo_b) + (slot - static_cast<std::size_t>(a16))];
            return 0.0;
        };
        const std::vector<double> op = operand_table(b, ns, coef);
        std::vector<float> o32(op.size());
        // bf16 planes: each term is the bf16 rounding of what the previous
        // terms leave (residuals exact in FP64); three terms carry about 24
        // bits.
        std::vector<uint16_t> obf(kBf16Planes * op.size());
        for (std::size_t q = 0; q < op.size(); ++q) {
            const double v = op[q];
            o32[q] = static_cast<float>(v);
            double r = v;
            for (int t = 0; t < kBf16Planes; ++t) {
                const uint16_t h = bf16(static_cast<float>(r));
                obf[static_cast<std::size_t>(t) * op.size() + q] = h;
                r -= bf16_value(h);
            }
        }
        // Row norms over each channel's slots (the density error scale).
        std::vector<float> nrm(2 * nr);
        for (std::size_t r = 0; r < nr; ++r) {
            double sa = 0.0, sb = 0.0;
            for (std::size_t slot = 0; slot < ns; ++slot) {
                const double v = op[r * ns + slot];
                (slot < static_cast<std::size_t>(m.alpha_slots) ? sa : sb) += v * v;
            }
            nrm[2 * r] = static_cast<float>(std::sqrt(sa));
            nrm[2 * r + 1] = static_cast<float>(std::sqrt(sb));
        }
        m.d_opnorm = upload(nrm.data(), nrm.size(), s);
        m.d_op64 = upload(op.data(), op.size(), s);
        m.d_op32 = upload(o32.data(), o32.size(), s);
        m.d_opbf = upload(obf.data(), obf.size(), s);
    }
    KUMO_HIP(hipStreamSynchronize(s));
    m.device_bytes = 0;
    for (const DeviceBuffer* d : {&m.d_shells, &m.d_exps, &m.d_weights, &m.d_tdense, &m.d_radw, &m.d_tg, &m.d_rec, &m.d_p[0],
                                  &m.d_p[1], &m.d_ca, &m.d_cb, &m.d_op64, &m.d_op32, &m.d_opbf, &m.d_opnorm})
        m.device_bytes += static_cast<long long>(d->bytes());
}

std::vector<double> mo_matrix(const double* flat, int nmo, int n) {
    // FCHK: MO-major, AO fastest -> AO rows x MO columns.
    std::vector<double> c(static_cast<std::size_t>(n) * nmo);
    for (int mo = 0; mo < nmo; ++mo)
        for (int ao = 0; ao < n; ++ao)
            c[static_cast<std::size_t>(ao) * nmo + mo] = flat[static_cast<std::size_t>(mo) * n + ao];
    return c;
}

} // namespace

std::unique_ptr<kumo_hip_molecule> load_molecule(kumo_hip_context* ctx, const kumo_hip_molecule_desc& d) {
    auto m = std::make_unique<kumo_hip_molecule>();
    m->ctx = ctx;
    if (d.natoms <= 0 || !d.atomic_numbers || !d.coordinates) runtime("invalid FCHK geometry");
    for (int a = 0; a < d.natoms; ++a) {
        const int z = d.atomic_numbers[a];
        const double q = d.nuclear_charges ? d.nuclear_charges[a] : static_cast<double>(z);
        if (z <= 0 || z > 118 || !std::isfinite(d.coordinates[3 * a]) || !std::isfinite(d.coordinates[3 * a + 1]) ||
            !std::isfinite(d.coordinates[3 * a + 2]))
            runtime("invalid FCHK atom or coordinates");
        if (!std::isfinite(q) || q < 0.0) runtime("invalid FCHK nuclear charge");
        m->atoms.push_back({z, q, {d.coordinates[3 * a], d.coordinates[3 * a + 1], d.coordinates[3 * a + 2]}});
    }
    ShellRecords r;
    r.nshells = d.nshells;
    r.types = d.shell_types;
    r.nprim = d.primitives_per_shell;
    r.atom = d.shell_to_atom;
    r.xyz = d.shell_coordinates;
    r.exps = d.exponents;
    r.coef = d.coefficients;
    r.pcoef = d.p_coefficients;
    r.natoms = d.natoms;
    r.nbasis = d.nbasis;
    m->basis = prepare_basis(r);
    m->n = m->basis.nao;
    const std::size_t nn = static_cast<std::size_t>(m->n) * m->n;
    if (!d.p_total) runtime("missing fchk record: Total SCF Density");
    m->P.assign(d.p_total, d.p_total + nn);
    if (d.p_spin) m->S.assign(d.p_spin, d.p_spin + nn);
    else if (d.nalpha != d.nbeta) runtime("open-shell FCHK lacks Spin SCF Density");
    else m->S.assign(nn, 0.0);
    for (double v : m->P)
        if (!std::isfinite(v)) runtime("nonfinite AO density matrix");
    if (d.nalpha < 0 || d.nbeta < 0 || d.nalpha > m->n || d.nbeta > m->n) runtime("invalid electron counts in FCHK");
    m->nalpha = d.nalpha;
    m->nbeta = d.nbeta;
    if (d.c_alpha && d.nmo_alpha > 0) {
        if (d.nmo_alpha > m->n) runtime("invalid MO coefficient dimensions: Alpha MO coefficients");
        m->Ca = mo_matrix(d.c_alpha, d.nmo_alpha, m->n);
        m->nmo_a = d.nmo_alpha;
    }
    if (d.c_beta && d.nmo_beta > 0) {
        if (d.nmo_beta > m->n) runtime("invalid MO coefficient dimensions: Beta MO coefficients");
        m->Cb = mo_matrix(d.c_beta, d.nmo_beta, m->n);
        m->nmo_b = d.nmo_beta;
    }
    m->unrestricted = m->nmo_b != 0;
    if (m->nalpha != m->nbeta && !m->unrestricted) runtime("open-shell FCHK lacks Beta MO coefficients");
    if (d.e_alpha && m->nmo_a) m->Ea.assign(d.e_alpha, d.e_alpha + m->nmo_a);
    if (d.e_beta && m->nmo_b) m->Eb.assign(d.e_beta, d.e_beta + m->nmo_b);
    assess(*m);
    upload_device(*m, ctx->exec.stream);
    // the screening radii of the default threshold (1e-12) are resident from
    // the start, so evaluations at that threshold allocate nothing for them
    m->screen_r2(1e-12, ctx->exec.stream);
    return m;
}

std::unique_ptr<kumo_hip_molecule> load_fchk(kumo_hip_context* ctx, const std::string& path) {
    const Fchk f(path);
    kumo_hip_molecule_desc d{};
    const auto& z = f.integers("Atomic numbers");
    std::vector<int32_t> zi(z.begin(), z.end());
    const auto& xyz = f.reals("Current cartesian coordinates");
    if (xyz.size() != 3 * zi.size() || zi.empty()) runtime("invalid FCHK geometry");
    std::vector<double> charges;
    if (f.contains("Nuclear charges")) {
        charges = f.reals("Nuclear charges");
        if (charges.size() != zi.size()) runtime("invalid FCHK nuclear charge count");
    } else {
        for (const char* label : {"ECP-LMax", "ECP-LPSkip", "ECP-RNFroz", "ECP-NLP", "ECP-CLP1", "ECP-CLP2",
                                  "ECP-ZLP", "ECP-KFirst", "ECP-KLast"})
            if (f.contains(label))
                runtime("FCHK has an effective core potential (ECP) but no \"Nuclear charges\" record, so the "
                        "effective nuclear charges are unknown");
    }
    const auto& types = f.integers("Shell types");
    const auto& nprim = f.integers("Number of primitives per shell");
    const auto& centers = f.integers("Shell to atom map");
    if (types.size() != nprim.size() || types.size() != centers.size()) runtime("inconsistent shell arrays in fchk");
    std::vector<int32_t> ti(types.begin(), types.end()), ni(nprim.begin(), nprim.end()), ci(centers.begin(), centers.end());
    const auto& exps = f.reals("Primitive exponents");
    const auto& coef = f.reals("Contraction coefficients");
    const std::vector<double>* pc = f.contains("P(S=P) Contraction coefficients") ? &f.reals("P(S=P) Contraction coefficients") : nullptr;
    const auto& sxyz = f.reals("Coordinates of each shell");
    if (sxyz.size() != 3 * ti.size()) runtime("invalid shell-coordinate array in fchk");
    if (exps.size() != coef.size() || (pc && pc->size() != exps.size())) runtime("inconsistent primitive arrays in fchk");
    long long total_prims = 0;
    for (auto v : nprim) {
        if (v <= 0) runtime("invalid primitive count in fchk shell");
        total_prims += v;
    }
    if (static_cast<std::size_t>(total_prims) != exps.size()) runtime("unused primitive data after parsing fchk basis");
    const int n = static_cast<int>(f.integer("Number of basis functions"));
    if (n <= 0) runtime("invalid Number of basis functions");
    const auto unpack = [&](const std::vector<double>& packed) {
        const std::size_t nn = static_cast<std::size_t>(n);
        if (packed.size() != nn * (nn + 1) / 2) runtime("invalid packed AO density matrix size");
        std::vector<double> p(nn * nn);
        std::size_t q = 0;
        for (std::size_t i = 0; i < nn; ++i)
            for (std::size_t j = 0; j <= i; ++j, ++q)
                p[i * nn + j] = p[j * nn + i] = packed[q];
        return p;
    };
    const auto P = unpack(f.reals("Total SCF Density"));
    std::vector<double> S;
    if (f.contains("Spin SCF Density")) S = unpack(f.reals("Spin SCF Density"));
    if (!f.contains("Number of alpha electrons") || !f.contains("Number of beta electrons"))
        runtime("missing electron counts in FCHK");
    d.natoms = static_cast<int32_t>(zi.size());
    d.atomic_numbers = zi.data();
    d.nuclear_charges = charges.empty() ? nullptr : charges.data();
    d.coordinates = xyz.data();
    d.nshells = static_cast<int32_t>(ti.size());
    d.shell_types = ti.data();
    d.primitives_per_shell = ni.data();
    d.shell_to_atom = ci.data();
    d.shell_coordinates = sxyz.data();
    d.exponents = exps.data();
    d.coefficients = coef.data();
    d.p_coefficients = pc ? pc->data() : nullptr;
    d.nbasis = n;
    d.p_total = P.data();
    d.p_spin = S.empty() ? nullptr : S.data();
    d.nalpha = static_cast<int32_t>(f.integer("Number of alpha electrons"));
    d.nbeta = static_cast<int32_t>(f.integer("Number of beta electrons"));
    const auto mo = [&](const char* label, int32_t& nmo) -> const double* {
        if (!f.contains(label)) return nullptr;
        const auto& v = f.reals(label);
        if (v.size() % static_cast<std::size_t>(n) != 0 || v.size() / static_cast<std::size_t>(n) > static_cast<std::size_t>(n))
            runtime(std::string("invalid MO coefficient dimensions: ") + label);
        nmo = static_cast<int32_t>(v.size() / static_cast<std::size_t>(n));
        return v.data();
    };
    d.c_alpha = mo("Alpha MO coefficients", d.nmo_alpha);
    d.c_beta = mo("Beta MO coefficients", d.nmo_beta);
    if (f.contains("Alpha Orbital Energies")) d.e_alpha = f.reals("Alpha Orbital Energies").data();
    if (f.contains("Beta Orbital Energies")) d.e_beta = f.reals("Beta Orbital Energies").data();
    return load_molecule(ctx, d);
}

// Boys table F_m(i * 0.05), m < 39, i < 801 (src/boys_function.cpp:20-44).
std::vector<double> boys_table_host() {
    constexpr int orders = 39, points = 801, top = orders - 1;
    std::vector<double> t(static_cast<std::size_t>(orders * points));
    for (int i = 0; i < points; ++i) {
        const double T = i * 0.05;
        const double decay = std::exp(-T);
        double term = 1.0 / (2.0 * top + 1.0), sum = term;
        for (int k = 1; k < 10000 && term > 1e-18 * sum; ++k) {
            term *= 2.0 * T / (2.0 * top + 2.0 * k + 1.0);
            sum += term;
        }
        double* row = t.data() + static_cast<std::size_t>(i * orders);
        row[top] = decay * sum;
        for (int m = top; m > 0; --m)
            row[m - 1] = (2.0 * T * row[m] + decay) / (2.0 * m - 1.0);
    }
    return t;
}

// The device copy: the table, then for each ESP pair order L <= 2 kMaxL the
// seven orders L .. L + 6 of every point, padded to kBoysRowL doubles, so
// that a kernel of one L reads one short row per point from a table small
// enough to stay in the vector caches.
std::vector<double> boys_device_table() {
    std::vector<double> t = boys_table_host();
    constexpr int orders = 39, points = 801;
    t.resize(static_cast<std::size_t>(orders * points + (2 * kMaxL + 1) * points * kBoysRowL), 0.0);
    for (int L = 0; L <= 2 * kMaxL; ++L)
        for (int i = 0; i < points; ++i)
            for (int k = 0; k < 7; ++k)
                t[static_cast<std::size_t>(orders * points + (L * points + i) * kBoysRowL + k)] =
                    t[static_cast<std::size_t>(i * orders + L + k)];
    return t;
}

} // namespace kumo

const double* kumo_hip_molecule::p_matrix(int component, hipStream_t stream) {
    using namespace kumo;
    if (component <= 1) return d_p[component].as<double>();
    std::lock_guard<std::mutex> guard(lazy);
    if (!d_p[component].get()) {
        std::vector<double> m(P.size());
        for (std::size_t q = 0; q < P.size(); ++q)
            m[q] = 0.5 * (component == 2 ? (P[q] + S[q]) : (P[q] - S[q]));
        d_p[component] = upload(m.data(), m.size(), stream);
        KUMO_HIP(hipStreamSynchronize(stream));
    }
    return d_p[component].as<double>();
}

const double* kumo_hip_molecule::screen_r2(double threshold, hipStream_t stream) {
    using namespace kumo;
    std::lock_guard<std::mutex> guard(lazy);
    auto it = d_screen_r2.find(threshold);
    if (it == d_screen_r2.end()) {
        const auto r2 = screening_radii2(basis, threshold);
        it = d_screen_r2.emplace(threshold, upload(r2.data(), r2.size(), stream)).first;
        KUMO_HIP(hipStreamSynchronize(stream));
    }
    return it->second.as<double>();
}

const double* kumo_hip_molecule::screen_r2_resident(double threshold) {
    std::lock_guard<std::mutex> guard(lazy);
    const auto it = d_screen_r2.find(threshold);
    return it == d_screen_r2.end() ? nullptr : it->second.as<double>();
}

// Primitive pairs of the exact ESP: P contracted with the Hermite expansion
// coefficients of every pair of Cartesian primitives of blocks a <= b
// (src/electrostatic_potential.cpp:117-229), grouped by L = la + lb.
void kumo_hip_molecule::ensure_orbital_energy(hipStream_t stream) {
    using namespace kumo;
    std::lock_guard<std::mutex> guard(lazy);
    if (oe_ready) return;
    std::vector<OeSlot> slots;
    const auto pad = [&] {
        while (slots.size() % 16) slots.push_back({0, 0, -1, 0.0});
    };
    // Occupied MOs (src/orbital_energy_fields.hpp:160-187): restricted, the
    // first min(nalpha, nbeta) doubly and the rest up to max(nalpha, nbeta)
    // singly occupied; unrestricted, each channel's occupied MOs singly.
    const int nocc_r = std::max(nalpha, nbeta), ndouble = std::min(nalpha, nbeta);
    if (unrestricted) {
        for (int i = 0; i < nalpha; ++i) slots.push_back({1, 0, i, 1.0});
        for (int i = 0; i < nbeta; ++i) slots.push_back({1, 1, i, 1.0});
    } else {
        for (int i = 0; i < nocc_r; ++i) slots.push_back({1, 0, i, i < ndouble ? 2.0 : 1.0});
    }
    pad();
    oe_occ_slots = static_cast<int>(slots.size());
    if (unrestricted) {
        for (int i = nalpha; i < nmo_a; ++i) slots.push_back({2, 0, i, 0.0});
        for (int i = nbeta; i < nmo_b; ++i) slots.push_back({2, 1, i, 0.0});
    } else {
        for (int i = nocc_r; i < nmo_a; ++i) slots.push_back({2, 0, i, 0.0});
    }
    pad();
    if (slots.empty()) slots.assign(16, OeSlot{0, 0, -1, 0.0});
    const auto coef = [&](int ao, std::size_t slot) -> double {
        const OeSlot& q = slots[slot];
        if (q.kind == 0) return 0.0;
        const auto& C = q.channel == 1 ? Cb : Ca;
        const int nmo = q.channel == 1 ? nmo_b : nmo_a;
        return C[static_cast<std::size_t>(ao) * static_cast<std::size_t>(nmo) + static_cast<std::size_t>(q.mo)];
    };
    const std::vector<double> op = operand_table(basis, slots.size(), coef);
    d_oe_op = upload(op.data(), op.size(), stream);
    KUMO_HIP(hipStreamSynchronize(stream));
    oe_slots = static_cast<int>(slots.size());
    oe_slot = std::move(slots);
    device_bytes += static_cast<long long>(d_oe_op.bytes());
    oe_ready = true;
}

void kumo_hip_molecule::ensure_esp(hipStream_t stream) {
    using namespace kumo;
    std::lock_guard<std::mutex> guard(lazy);
    if (esp_ready) return;
    struct Triple {
        int t, u, v;
    };
    const auto triples = [](int L) {
        std::vector<Triple> out;
        for (int s = 0; s < L + 1; ++s)
            for (int t = s; t >= 0; --t)
                for (int u = s - t; u >= 0; --u)
                    out.push_back({t, u, s - t - u});
        return out;
    };
    // One Cartesian direction of a primitive pair, as McMurchie-Davidson coefficients E^{ij}_t:
    // (x-A)^i (x-B)^j = sum_t E^{ij}_t Lambda_t, Lambda_t = (d/dPx)^t exp(-p u^2), u = x - Px.
    // The product is first multiplied out as a polynomial in u, (u + pa)^i (u + pb)^j, and each
    // power u^k is then rewritten over the Lambda_t with u Lambda_t = Lambda_{t+1} / 2p + t Lambda_{t-1}.
    struct Herm {
        int lb, width;
        std::vector<double> e; // (i, j, t) at (i * (lb + 1) + j) * width + t
        Herm(int la, int lb_, double p, double pa, double pb) : lb(lb_), width(la + lb_ + 1) {
            const int L = la + lb;
            const double half = 0.5 / p;
            // u^k = sum_t power[k * width + t] Lambda_t
            std::vector<double> power(static_cast<std::size_t>(width * width), 0.0);
            power[0] = 1.0;
            for (int k = 1; k <= L; ++k)
                for (int t = 0; t <= k; ++t) {
                    double d = t > 0 ? half * power[static_cast<std::size_t>((k - 1) * width + t - 1)] : 0.0;
                    if (t + 1 <= k - 1) d += (t + 1) * power[static_cast<std::size_t>((k - 1) * width + t + 1)];
                    power[static_cast<std::size_t>(k * width + t)] = d;
                }
            e.assign(static_cast<std::size_t>((la + 1) * (lb + 1) * width), 0.0);
            std::vector<double> left(static_cast<std::size_t>(width), 0.0), poly;
            left[0] = 1.0; // (u + pa)^i, coefficient of u^k at k
            for (int i = 0; i <= la; ++i) {
                if (i > 0)
                    for (int k = i; k >= 0; --k)
                        left[static_cast<std::size_t>(k)] = (k > 0 ? left[static_cast<std::size_t>(k - 1)] : 0.0) +
                                                            pa * left[static_cast<std::size_t>(k)];
                poly = left; // (u + pa)^i (u + pb)^j
                for (int j = 0; j <= lb; ++j) {
                    if (j > 0)
                        for (int k = i + j; k >= 0; --k)
                            poly[static_cast<std::size_t>(k)] = (k > 0 ? poly[static_cast<std::size_t>(k - 1)] : 0.0) +
                                                                pb * poly[static_cast<std::size_t>(k)];
                    for (int t = 0; t <= i + j; ++t) {
                        double sum = 0.0;
                        for (int k = t; k <= i + j; k += 2)
                            sum += poly[static_cast<std::size_t>(k)] * power[static_cast<std::size_t>(k * width + t)];
                        e[static_cast<std::size_t>((i * (lb + 1) + j) * width + t)] = sum;
                    }
                }
            }
        }
        double get(int i, int j, int t) const {
            return (t < 0 || t > i + j) ? 0.0 : e[static_cast<std::size_t>((i * (lb + 1) + j) * width + t)];
        }
    };
    std::vector<std::vector<EspPairDev>> groups(2 * kMaxL + 1);
    std::vector<std::vector<double>> gcoef(2 * kMaxL + 1);
    std::vector<std::vector<int32_t>> gorder(2 * kMaxL + 1); // each pair's place in the reference's order
    int32_t made = 0;
    const auto& blocks = basis.blocks;
    const int nn = n;
    for (std::size_t ia = 0; ia < blocks.size(); ++ia) {
        const auto& A = blocks[ia];
        for (std::size_t ib = ia; ib < blocks.size(); ++ib) {
            const auto& B = blocks[ib];
            const int na = A.ncart(), nb = B.ncart();
            std::vector<double> left(static_cast<std::size_t>(na * B.count));
            for (int i = 0; i < na; ++i)
                for (int c = 0; c < B.count; ++c) {
                    double acc;
                    if (A.pure.empty()) {
                        acc = P[static_cast<std::size_t>(A.offset + i) * nn + static_cast<std::size_t>(B.offset + c)];
                    } else {
                        acc = 0.0;
                        for (int r = 0; r < A.count; ++r)
                            acc += A.pure[static_cast<std::size_t>(r * na + i)] *
                                   P[static_cast<std::size_t>(A.offset + r) * nn + static_cast<std::size_t>(B.offset + c)];
                    }
                    left[static_cast<std::size_t>(i * B.count + c)] = acc;
                }
            std::vector<double> cart(static_cast<std::size_t>(na * nb));
            bool any = false;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    double v;
                    if (B.pure.empty()) {
                        v = left[static_cast<std::size_t>(i * B.count + j)];
                    } else {
                        v = 0.0;
                        for (int c = 0; c < B.count; ++c)
                            v += left[static_cast<std::size_t>(i * B.count + c)] * B.pure[static_cast<std::size_t>(c * nb + j)];
                    }
                    if (ia != ib) v *= 2.0;
                    cart[static_cast<std::size_t>(i * nb + j)] = v;
                    any = any || v != 0.0;
                }
            if (!any) continue;
            const int la = A.l, lb = B.l, L = la + lb;
            const auto list = triples(L);
            const double ab[3] = {A.origin[0] - B.origin[0], A.origin[1] - B.origin[1], A.origin[2] - B.origin[2]};
            const double ab2 = (ab[0] * ab[0] + ab[1] * ab[1]) + ab[2] * ab[2];
            std::vector<double> pc(list.size());
            for (int p = 0; p < A.nprim(); ++p)
                for (int q = 0; q < B.nprim(); ++q) {
                    const double al = A.exps[static_cast<std::size_t>(p)], be = B.exps[static_cast<std::size_t>(q)];
                    const double ex = al + be;
                    double Q[3];
                    for (int a = 0; a < 3; ++a)
                        Q[a] = (al * A.origin[a] + be * B.origin[a]) / ex;
                    const double scale = std::exp(-al * be / ex * ab2) * 2.0 * std::numbers::pi / ex;
                    const Herm hx(la, lb, ex, Q[0] - A.origin[0], Q[0] - B.origin[0]);
                    const Herm hy(la, lb, ex, Q[1] - A.origin[1], Q[1] - B.origin[1]);
                    const Herm hz(la, lb, ex, Q[2] - A.origin[2], Q[2] - B.origin[2]);
                    std::fill(pc.begin(), pc.end(), 0.0);
                    for (int i = 0; i < na; ++i) {
                        const auto& pi = A.powers[static_cast<std::size_t>(i)];
                        for (int j = 0; j < nb; ++j) {
                            const auto& pj = B.powers[static_cast<std::size_t>(j)];
                            const double w = cart[static_cast<std::size_t>(i * nb + j)] *
                                             A.weights[static_cast<std::size_t>(i * A.nprim() + p)] *
                                             B.weights[static_cast<std::size_t>(j * B.nprim() + q)];
                            if (w == 0.0) continue;
                            for (std::size_t k = 0; k < list.size(); ++k) {
                                const auto& h = list[k];
                                if (h.t > pi.x + pj.x || h.u > pi.y + pj.y || h.v > pi.z + pj.z) continue;
                                pc[k] += w * hx.get(pi.x, pj.x, h.t) * hy.get(pi.y, pj.y, h.u) * hz.get(pi.z, pj.z, h.v);
                            }
                        }
                    }
                    if (std::all_of(pc.begin(), pc.end(), [](double c) { return c == 0.0; })) continue;
                    auto& g = groups[static_cast<std::size_t>(L)];
                    auto& gc = gcoef[static_cast<std::size_t>(L)];
                    g.push_back({ex, Q[0], Q[1], Q[2], L, static_cast<int32_t>(gc.size())});
                    gorder[static_cast<std::size_t>(L)].push_back(made++);
                    for (double c : pc)
                        gc.push_back(scale * c);
                }
        }
    }
    std::vector<EspPairDev> pairs;
    std::vector<double> coef;
    esp_groups.clear();
    esp_ref_order.assign(static_cast<std::size_t>(made), 0);
    for (std::size_t L = 0; L < groups.size(); ++L) {
        for (std::size_t k = 0; k < gorder[L].size(); ++k)
            esp_ref_order[static_cast<std::size_t>(gorder[L][k])] = static_cast<int32_t>(pairs.size() + k);
        const int base = static_cast<int>(coef.size());
        if (!groups[L].empty())
            esp_groups.push_back({static_cast<int32_t>(L), static_cast<int32_t>(pairs.size()),
                                  static_cast<int32_t>(pairs.size() + groups[L].size()), 1});
        for (auto p : groups[L]) {
            p.first += base;
            pairs.push_back(p);
        }
        coef.insert(coef.end(), gcoef[L].begin(), gcoef[L].end());
    }
    std::vector<EspAtomDev> at;
    for (const auto& a : atoms)
        at.push_back({a.xyz[0], a.xyz[1], a.xyz[2], a.charge});
    d_esp_pairs = upload(pairs.data(), pairs.size(), stream);
    d_esp_coef = upload(coef.data(), coef.size(), stream);
    d_esp_atoms = upload(at.data(), at.size(), stream);
    esp_npairs = static_cast<int>(pairs.size());
    KUMO_HIP(hipStreamSynchronize(stream));
    esp_pairs_host = std::move(pairs);
    esp_coef_host = std::move(coef);
    esp_ready = true;
}

// cluster_plan (src/potential_clusters.hpp:29-35): the radius enlarged by a
// relative 1e-10 plus 1e-12 bohr, then far_field_plan; uploaded once per
// (radius, tolerance).
const kumo::FarPlanDev& kumo_hip_molecule::far_plan(double radius, double tolerance, hipStream_t stream) {
    using namespace kumo;
    ensure_esp(stream);
    std::lock_guard<std::mutex> guard(lazy);
    const auto key = std::make_pair(radius, tolerance);
    auto it = far_plans.find(key);
    if (it != far_plans.end()) return it->second.dev;
    const FarPlanHost h =
        far_field_plan(esp_pairs_host, esp_coef_host, esp_ref_order, radius * (1.0 + 1e-10) + 1e-12, tolerance);
    FarPlanResident r;
    r.ref_order = upload(esp_ref_order.data(), esp_ref_order.size(), stream);
    r.level_d2 = upload(h.level_d2.data(), h.level_d2.size(), stream);
    r.clusters = upload(h.clusters.data(), h.clusters.size(), stream);
    r.cluster_pairs = upload(h.cluster_pairs.data(), h.cluster_pairs.size(), stream);
    r.cluster_of = upload(h.cluster_of.data(), h.cluster_of.size(), stream);
    r.moments = upload(h.moments.data(), h.moments.size(), stream);
    KUMO_HIP(hipStreamSynchronize(stream));
    FarPlanDev& d = r.dev;
    d.level_d2 = r.level_d2.as<float>();
    d.clusters = r.clusters.as<FarClusterDev>();
    d.cluster_pairs = r.cluster_pairs.as<int32_t>();
    d.cluster_of = r.cluster_of.as<int32_t>();
    d.moments = r.moments.as<double>();
    d.ref_order = r.ref_order.as<int32_t>();
    d.nclusters = static_cast<int>(h.clusters.size());
    d.npairs = static_cast<int>(esp_pairs_host.size());
    d.max_L = h.max_L;
    d.moment_count = h.moment_count;
    d.pair_tolerance = h.pair_tolerance;
    for (int j = 0; j < kFarLevels; ++j)
        d.levels[j] = h.levels[j];
    device_bytes += static_cast<long long>(r.level_d2.bytes() + r.clusters.bytes() + r.cluster_pairs.bytes() +
                                           r.cluster_of.bytes() + r.moments.bytes() + r.ref_order.bytes());
    return far_plans.emplace(key, std::move(r)).first->second.dev;
}
