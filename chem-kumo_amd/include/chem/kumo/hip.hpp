// C++20 mirror of chem-kumo's public evaluation API
// (include/chem/kumo/evaluator.hpp, include/chem/kumo/density.hpp), running on
// the MI355X through the C-ABI of include/kumo_hip.h. Names, argument meaning
// and exceptions follow the reference: a reference user keeps the same calls
// (Molecule::from_fchk, Evaluator(molecule, options).evaluate(grid, request),
// density::total.laplacian(), ...). The HIP build has no Eigen dependency:
// Vector3 / Matrix3 are plain arrays and points are passed as xyz triples.
//
// Header-only; link with libkumo_hip.so.
#pragma once

#include "../../../../include/kumo_hip.h"

#include <array>
#include <cmath>
#include <cstddef>
#include <filesystem>
#include <functional>
#include <map>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace chem::kumo {

enum class DensityComponent { Total, Alpha, Beta, Difference };    // include/chem/kumo/spin.hpp:8
enum class SpinChannel { Alpha, Beta };                             // include/chem/kumo/spin.hpp:12
enum class Backend { AUTO, AO_P, AO_MO };                           // include/chem/kumo/evaluator.hpp:403
enum class Precision { FP64, FP32, BF16X3, BF16X6 };                // HIP contraction arithmetic

// include/chem/kumo/evaluator.hpp:112-150 (same order, same values).
enum class PropertyKind {
    Density, Gradient, Hessian, Laplacian, Tau, Orbital, Elf, ReducedGradient, SignLambda2Density,
    ElectrostaticPotential, Lol, InteractionRegionIndicator, DensityOverlapRegionsIndicator,
    HamiltonianKineticEnergy, PotentialEnergyDensity, EnergyDensity, SpinPolarization, ShannonEntropyDensity,
    FisherInformationDensity, AverageLocalIonizationEnergy, LocalElectronAffinity, LocalElectronAttachmentEnergy,
    LocalElectronegativity, LocalHardness, OrbitalGradient, OrbitalHessian, OrbitalWeightedFukuiPlus,
    OrbitalWeightedFukuiMinus, OrbitalWeightedFukuiZero, OrbitalWeightedDualDescriptor, TauComponents,
    HamiltonianKineticEnergyComponents, PotentialKineticRatio, EnergyDensityPerElectron, KineticEnergyPerElectron,
    ElfTsirelsonStash, LolTsirelsonStash
};

inline constexpr double default_screening_threshold = 1e-12; // include/chem/kumo/density.hpp:28
inline constexpr std::size_t max_block_size = 65536;          // include/chem/kumo/density.hpp:32
inline constexpr double reduced_gradient_cap = 100.0;
inline constexpr double interaction_region_indicator_cap = 100.0;

using ProgressCallback = std::function<void(std::size_t done, std::size_t total)>;

using Vector3 = std::array<double, 3>;

struct Atom {
    int atomic_number = 0;
    double nuclear_charge = 0.0;
    Vector3 position{};
};

// Regular grid (include/chem/kumo/density.hpp:61-72); steps[a] is the step
// vector of axis a.
struct Grid {
    Vector3 origin{};
    std::array<Vector3, 3> steps{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    std::array<int, 3> counts{1, 1, 1};

    // Grid::around (src/grid.cpp:15-38).
    static Grid around(const std::vector<Atom>& atoms, double margin, double spacing) {
        if (atoms.empty() || !std::isfinite(margin) || margin < 0 || !std::isfinite(spacing) || spacing <= 0)
            throw std::invalid_argument("invalid grid parameters");
        Vector3 lo = atoms.front().position, hi = lo;
        for (const auto& a : atoms) {
            if (a.atomic_number <= 0) throw std::invalid_argument("invalid atom in grid");
            for (int k = 0; k < 3; ++k) {
                if (!std::isfinite(a.position[k])) throw std::invalid_argument("invalid atom in grid");
                lo[k] = std::min(lo[k], a.position[k]);
                hi[k] = std::max(hi[k], a.position[k]);
            }
        }
        Grid g;
        for (int k = 0; k < 3; ++k) {
            g.origin[k] = lo[k] - margin;
            const double count = std::ceil((hi[k] - lo[k] + 2 * margin) / spacing) + 1;
            if (!std::isfinite(count) || count > 2147483647.0) throw std::invalid_argument("grid too large");
            g.counts[k] = std::max(2, static_cast<int>(count));
            g.steps[k] = {0, 0, 0};
            g.steps[k][k] = spacing;
        }
        return g;
    }
    std::size_t size() const {
        return static_cast<std::size_t>(counts[0]) * static_cast<std::size_t>(counts[1]) *
               static_cast<std::size_t>(counts[2]);
    }
};

struct Property {
    PropertyKind kind;
    DensityComponent component = DensityComponent::Total;
    SpinChannel channel = SpinChannel::Alpha;
    std::size_t orbital_index = 0;
    friend bool operator==(const Property&, const Property&) = default;
    friend bool operator<(const Property& a, const Property& b) noexcept {
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.component != b.component) return a.component < b.component;
        if (a.channel != b.channel) return a.channel < b.channel;
        return a.orbital_index < b.orbital_index;
    }
};

struct DensityDescriptor {
    DensityComponent component;
    constexpr Property p(PropertyKind k) const { return {k, component}; }
    constexpr Property value() const { return p(PropertyKind::Density); }
    constexpr Property gradient() const { return p(PropertyKind::Gradient); }
    constexpr Property hessian() const { return p(PropertyKind::Hessian); }
    constexpr Property laplacian() const { return p(PropertyKind::Laplacian); }
    constexpr Property tau() const { return p(PropertyKind::Tau); }
    constexpr Property elf() const { return p(PropertyKind::Elf); }
    constexpr Property reduced_gradient() const { return p(PropertyKind::ReducedGradient); }
    constexpr Property sign_lambda2_density() const { return p(PropertyKind::SignLambda2Density); }
    constexpr Property electrostatic_potential() const { return p(PropertyKind::ElectrostaticPotential); }
    constexpr Property lol() const { return p(PropertyKind::Lol); }
    constexpr Property interaction_region_indicator() const { return p(PropertyKind::InteractionRegionIndicator); }
    constexpr Property density_overlap_regions_indicator() const {
        return p(PropertyKind::DensityOverlapRegionsIndicator);
    }
    constexpr Property hamiltonian_kinetic_energy() const { return p(PropertyKind::HamiltonianKineticEnergy); }
    constexpr Property potential_energy_density() const { return p(PropertyKind::PotentialEnergyDensity); }
    constexpr Property energy_density() const { return p(PropertyKind::EnergyDensity); }
    constexpr Property tau_components() const { return p(PropertyKind::TauComponents); }
    constexpr Property hamiltonian_kinetic_energy_components() const {
        return p(PropertyKind::HamiltonianKineticEnergyComponents);
    }
    constexpr Property potential_kinetic_ratio() const { return p(PropertyKind::PotentialKineticRatio); }
    constexpr Property energy_density_per_electron() const { return p(PropertyKind::EnergyDensityPerElectron); }
    constexpr Property elf_tsirelson_stash() const { return p(PropertyKind::ElfTsirelsonStash); }
    constexpr Property lol_tsirelson_stash() const { return p(PropertyKind::LolTsirelsonStash); }
    constexpr Property kinetic_energy_per_electron() const { return p(PropertyKind::KineticEnergyPerElectron); }
    constexpr Property spin_polarization() const { return p(PropertyKind::SpinPolarization); }
    constexpr Property shannon_entropy_density() const { return p(PropertyKind::ShannonEntropyDensity); }
    constexpr Property fisher_information_density() const { return p(PropertyKind::FisherInformationDensity); }
    // Orbital-energy fields (include/chem/kumo/evaluator.hpp:296-338; Total only).
    constexpr Property average_local_ionization_energy() const {
        return p(PropertyKind::AverageLocalIonizationEnergy);
    }
    constexpr Property local_electron_affinity() const { return p(PropertyKind::LocalElectronAffinity); }
    constexpr Property local_electron_attachment_energy() const {
        return p(PropertyKind::LocalElectronAttachmentEnergy);
    }
    constexpr Property local_electronegativity() const { return p(PropertyKind::LocalElectronegativity); }
    constexpr Property local_hardness() const { return p(PropertyKind::LocalHardness); }
    constexpr Property orbital_weighted_fukui_plus() const { return p(PropertyKind::OrbitalWeightedFukuiPlus); }
    constexpr Property orbital_weighted_fukui_minus() const { return p(PropertyKind::OrbitalWeightedFukuiMinus); }
    constexpr Property orbital_weighted_fukui_zero() const { return p(PropertyKind::OrbitalWeightedFukuiZero); }
    constexpr Property orbital_weighted_dual_descriptor() const {
        return p(PropertyKind::OrbitalWeightedDualDescriptor);
    }
};

namespace density {
inline constexpr DensityDescriptor total{DensityComponent::Total};
inline constexpr DensityDescriptor alpha{DensityComponent::Alpha};
inline constexpr DensityDescriptor beta{DensityComponent::Beta};
inline constexpr DensityDescriptor difference{DensityComponent::Difference};
} // namespace density

struct OrbitalDescriptor {
    SpinChannel channel;
    std::size_t index;
    Property value() const { return {PropertyKind::Orbital, DensityComponent::Total, channel, index}; }
    Property gradient() const { return {PropertyKind::OrbitalGradient, DensityComponent::Total, channel, index}; }
    Property hessian() const { return {PropertyKind::OrbitalHessian, DensityComponent::Total, channel, index}; }
};

namespace orbital {
inline OrbitalDescriptor alpha(std::size_t index) { return {SpinChannel::Alpha, index}; }
inline OrbitalDescriptor beta(std::size_t index) { return {SpinChannel::Beta, index}; }
} // namespace orbital

class Request {
public:
    Request& add(Property property) {
        properties_.push_back(property);
        return *this;
    }
    const std::vector<Property>& properties() const noexcept { return properties_; }

private:
    std::vector<Property> properties_;
};

struct EvaluationOptions {
    int threads = 0;
    std::size_t block_size = 128;
    Backend backend = Backend::AUTO;
    double screening_threshold = default_screening_threshold;
    double potential_tolerance = 1e-8;
    double fukui_delta = 0.1;
    ProgressCallback progress = {};
    Precision precision = Precision::FP64; // HIP only
    // HIP only: the ESP at potential_tolerance > 0 through the reference's far
    // field (true) or by the exact pair sums (false: within any tolerance, and
    // the faster of the two on the GPU).
    bool far_field = false;
};

struct PropertyField {
    std::size_t points = 0;
    std::vector<std::vector<double>> components;
    std::span<const double> component(std::size_t index = 0) const { return components.at(index); }
};

class Results {
public:
    const PropertyField& get(Property key) const { return fields_.at(key); }
    bool contains(Property key) const noexcept { return fields_.count(key) != 0; }
    Backend backend_used() const noexcept { return backend_used_; }

private:
    friend class Evaluator;
    std::map<Property, PropertyField> fields_;
    Backend backend_used_ = Backend::AO_P;
};

namespace detail {
// One process-wide context on device 0 (kumo_hip_init).
inline kumo_hip_context* context() {
    static std::unique_ptr<kumo_hip_context, int (*)(kumo_hip_context*)> ctx = [] {
        kumo_hip_context* c = nullptr;
        if (kumo_hip_init(0, &c) != KUMO_HIP_OK) throw std::runtime_error("no usable HIP device for the HIP backend");
        return std::unique_ptr<kumo_hip_context, int (*)(kumo_hip_context*)>(c, kumo_hip_release);
    }();
    return ctx.get();
}

[[noreturn]] inline void raise(int status) {
    char buf[1024] = {0};
    kumo_hip_last_error(context(), buf, sizeof(buf));
    switch (status) {
    case KUMO_HIP_INVALID_ARGUMENT: throw std::invalid_argument(buf);
    case KUMO_HIP_OUT_OF_RANGE: throw std::out_of_range(buf);
    default: throw std::runtime_error(buf);
    }
}
inline void check(int status) {
    if (status != KUMO_HIP_OK) raise(status);
}
inline kumo_hip_grid to_c(const Grid& grid) {
    kumo_hip_grid g{};
    for (int a = 0; a < 3; ++a) {
        g.origin[a] = grid.origin[a];
        g.counts[a] = grid.counts[a];
        for (int k = 0; k < 3; ++k)
            g.steps[3 * a + k] = grid.steps[a][k];
    }
    return g;
}
} // namespace detail

class Molecule {
public:
    static Molecule from_fchk(const std::string& path) {
        kumo_hip_molecule* m = nullptr;
        detail::check(kumo_hip_molecule_from_fchk(detail::context(), path.c_str(), &m));
        return Molecule(m);
    }
    const std::vector<Atom>& atoms() const noexcept { return atoms_; }
    std::size_t basis_size() const noexcept { return static_cast<std::size_t>(info_.nbasis); }
    bool unrestricted() const noexcept { return info_.unrestricted != 0; }
    bool ao_mo_eligible() const noexcept { return info_.ao_mo_eligible != 0; }
    std::size_t occupied_count(SpinChannel c = SpinChannel::Alpha) const {
        return static_cast<std::size_t>(c == SpinChannel::Alpha ? info_.nalpha : info_.nbeta);
    }
    const kumo_hip_molecule* handle() const noexcept { return mol_.get(); }

private:
    explicit Molecule(kumo_hip_molecule* m) : mol_(m, kumo_hip_molecule_release) {
        kumo_hip_molecule_info_get(m, &info_);
        std::vector<int32_t> z(static_cast<std::size_t>(info_.natoms));
        std::vector<double> q(z.size()), xyz(3 * z.size());
        kumo_hip_molecule_atoms(m, z.data(), q.data(), xyz.data());
        for (std::size_t a = 0; a < z.size(); ++a)
            atoms_.push_back({z[a], q[a], {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]}});
    }
    std::shared_ptr<kumo_hip_molecule> mol_;
    kumo_hip_molecule_info info_{};
    std::vector<Atom> atoms_;
};

inline std::size_t channels(PropertyKind k) {
    switch (k) {
    case PropertyKind::Gradient: case PropertyKind::OrbitalGradient: case PropertyKind::TauComponents:
    case PropertyKind::HamiltonianKineticEnergyComponents: return 3;
    case PropertyKind::Hessian: case PropertyKind::OrbitalHessian: return 6;
    default: return 1;
    }
}

// Evaluator (include/chem/kumo/evaluator.hpp:487-538). The Molecule must
// outlive the Evaluator, as in the reference.
class Evaluator {
public:
    explicit Evaluator(const Molecule& molecule, EvaluationOptions options = {})
        : molecule_(&molecule), options_(std::move(options)) {
        if (options_.block_size == 0 || options_.block_size > max_block_size)
            throw std::invalid_argument("evaluation block_size must be in [1, 65536]");
        if (options_.threads < 0) throw std::invalid_argument("invalid evaluation options");
        if (options_.threads > 1)
            throw std::invalid_argument("explicit multi-thread control is not implemented in the API adapter");
        if (!std::isfinite(options_.screening_threshold) || options_.screening_threshold < 0.0)
            throw std::invalid_argument("screening threshold must be finite and nonnegative");
        if (!std::isfinite(options_.potential_tolerance) || options_.potential_tolerance < 0.0)
            throw std::invalid_argument("potential tolerance must be finite and nonnegative");
        if (!std::isfinite(options_.fukui_delta) || options_.fukui_delta <= 0.0)
            throw std::invalid_argument("fukui_delta must be finite and positive");
        if (options_.backend == Backend::AO_MO && !molecule.ao_mo_eligible())
            throw std::invalid_argument("AO_MO backend requested but orbital representation is not eligible");
    }

    Results evaluate(const Grid& grid, const Request& request) const {
        const kumo_hip_grid g = detail::to_c(grid);
        if (grid.counts[0] <= 0 || grid.counts[1] <= 0 || grid.counts[2] <= 0)
            throw std::invalid_argument("invalid or oversized grid");
        return run(request, grid.size(), [&](const kumo_hip_property* p, std::size_t np, const kumo_hip_options* o,
                                             double* const* out, int32_t* used) {
            return kumo_hip_evaluate_grid(detail::context(), molecule_->handle(), &g, p, np, o, out, used);
        });
    }

    // Points as xyz triples (bohr), in order.
    Results evaluate(std::span<const Vector3> points, const Request& request) const {
        return run(request, points.size(), [&](const kumo_hip_property* p, std::size_t np, const kumo_hip_options* o,
                                               double* const* out, int32_t* used) {
            return kumo_hip_evaluate_points(detail::context(), molecule_->handle(),
                                            reinterpret_cast<const double*>(points.data()), points.size(), p, np, o,
                                            out, used);
        });
    }

    struct Field {
        Grid grid;
        std::vector<Atom> atoms;
        std::vector<double> values;
        // Field::write_cube (include/chem/kumo/density.hpp:116-121).
        void write_cube(const std::string& path, const std::string& comment = "Electron density") const {
            write(path, comment, 0);
        }
        // Field::write_orbital_cube (include/chem/kumo/density.hpp:122-128).
        void write_orbital_cube(const std::string& path, std::size_t orbital_index,
                                const std::string& comment = "Molecular orbital amplitude") const {
            if (orbital_index == 0 || orbital_index > 9999) throw std::invalid_argument("invalid orbital index for Cube");
            write(path, comment, orbital_index);
        }

    private:
        void write(const std::string& path, const std::string& comment, std::size_t orbital_index) const {
            if (values.size() != grid.size()) throw std::invalid_argument("field size mismatch");
            const kumo_hip_grid g = detail::to_c(grid);
            std::vector<int32_t> z;
            std::vector<double> q, xyz;
            for (const Atom& a : atoms) {
                z.push_back(a.atomic_number);
                q.push_back(a.nuclear_charge);
                xyz.insert(xyz.end(), a.position.begin(), a.position.end());
            }
            const auto n = static_cast<int32_t>(atoms.size());
            detail::check(orbital_index
                              ? kumo_hip_write_orbital_cube(detail::context(), path.c_str(), comment.c_str(), &g, n,
                                                            z.data(), q.data(), xyz.data(), values.data(), 0,
                                                            orbital_index)
                              : kumo_hip_write_cube(detail::context(), path.c_str(), comment.c_str(), &g, n, z.data(),
                                                    q.data(), xyz.data(), values.data(), 0));
        }
    };
    Field field(const Grid& grid, Property property) const {
        if (channels(property.kind) != 1)
            throw std::invalid_argument("Evaluator::field needs a one-component property (not a gradient, a "
                                        "Hessian or the Cartesian terms of G or K)");
        const auto r = evaluate(grid, Request().add(property));
        const auto v = r.get(property).component();
        return {grid, molecule_->atoms(), std::vector<double>(v.begin(), v.end())};
    }

private:
    template <typename Call> Results run(const Request& request, std::size_t n, Call&& call) const {
        Results res;
        std::vector<kumo_hip_property> props;
        std::vector<Property> keys;
        for (const auto& p : request.properties()) {
            if (res.fields_.count(p)) continue; // a property requested twice is evaluated once
            res.fields_[p] = PropertyField{n, std::vector<std::vector<double>>(channels(p.kind), std::vector<double>(n))};
            keys.push_back(p);
            props.push_back({static_cast<int32_t>(p.kind), static_cast<int32_t>(p.component),
                             static_cast<int32_t>(p.channel), static_cast<int32_t>(p.orbital_index)});
        }
        std::vector<double*> outs;
        for (const auto& k : keys)
            for (auto& c : res.fields_[k].components)
                outs.push_back(c.data());
        kumo_hip_options o;
        kumo_hip_default_options(&o);
        o.threads = options_.threads;
        o.block_size = static_cast<int32_t>(options_.block_size);
        o.backend = static_cast<int32_t>(options_.backend);
        o.precision = static_cast<int32_t>(options_.precision);
        o.screening_threshold = options_.screening_threshold;
        o.potential_tolerance = options_.potential_tolerance;
        o.fukui_delta = options_.fukui_delta;
        o.esp_far_field = options_.far_field ? KUMO_HIP_ESP_FAR : KUMO_HIP_ESP_AUTO;
        struct Hook {
            const ProgressCallback* cb;
            std::exception_ptr error;
        } hook{&options_.progress, nullptr};
        if (options_.progress) {
            o.progress = [](std::size_t d, std::size_t t, void* u) -> int {
                auto* h = static_cast<Hook*>(u);
                try {
                    (*h->cb)(d, t);
                    return 0;
                } catch (...) {
                    h->error = std::current_exception();
                    return 1;
                }
            };
            o.progress_user = &hook;
        }
        int32_t used = 0;
        const int rc = call(props.data(), props.size(), &o, outs.data(), &used);
        if (hook.error) std::rethrow_exception(hook.error); // the callback's own exception
        detail::check(rc);
        res.backend_used_ = static_cast<Backend>(used);
        return res;
    }

    const Molecule* molecule_;
    EvaluationOptions options_;
};

// ---- molecular surfaces (include/chem/kumo/surface.hpp) ----------------

inline constexpr double hartree_in_kcal_per_mol = 627.5094740631;  // include/chem/kumo/surface.hpp:18
inline constexpr double hartree_in_electronvolt = 27.211386245988; // include/chem/kumo/surface.hpp:20
inline constexpr double minimum_density_isovalue = 1e-8;           // include/chem/kumo/surface.hpp:90

// Closed, oriented triangle mesh (include/chem/kumo/surface.hpp:25-34):
// vertices in bohr, triangles counter-clockwise seen from outside.
struct TriangleMesh {
    std::vector<Vector3> vertices;
    std::vector<std::array<std::size_t, 3>> triangles;

    double area() const { return measure().first; }
    double volume() const { return measure().second; }

private:
    std::pair<double, double> measure() const {
        double area = 0.0, volume = 0.0;
        for (const auto& t : triangles) {
            const Vector3 &a = vertices.at(t[0]), &b = vertices.at(t[1]), &c = vertices.at(t[2]);
            const Vector3 u{b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v{c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const Vector3 n{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
            area += 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            // divergence theorem: (1/6) a . ((b - a) x (c - a)) = (1/6) a . (b x c)
            volume += (a[0] * n[0] + a[1] * n[1] + a[2] * n[2]) / 6.0;
        }
        return {area, volume};
    }
};

// Politzer-Murray statistics of a function on a surface
// (include/chem/kumo/surface.hpp:36-74).
struct SurfaceStatistics {
    double area = 0.0, positive_area = 0.0, negative_area = 0.0, nonpolar_area = 0.0;
    double minimum = 0.0, maximum = 0.0;
    std::size_t minimum_vertex = 0, maximum_vertex = 0;
    double average = 0.0, positive_average = 0.0, negative_average = 0.0;
    double positive_variance = 0.0, negative_variance = 0.0, total_variance = 0.0;
    double balance = 0.0, charge_separation = 0.0, polarity_index = 0.0;
    std::vector<std::size_t> local_minima, local_maxima;
};

struct SurfaceOptions { // include/chem/kumo/surface.hpp:108-116
    double isovalue = 1e-3;
    double spacing = 0.25;
    double potential_tolerance = 1e-8; // the HIP path sums the potential exactly (within any tolerance)
};

struct PotentialSurface { // include/chem/kumo/surface.hpp:118-123
    TriangleMesh mesh;
    std::vector<double> potential;
    SurfaceStatistics statistics;
};

struct IonizationEnergySurface { // include/chem/kumo/surface.hpp:156-166
    TriangleMesh mesh;
    std::vector<double> ionization_energy;
    SurfaceStatistics statistics;
};

struct MappedSurface { // include/chem/kumo/surface.hpp:176-182
    TriangleMesh mesh;
    std::vector<Property> properties;
    std::vector<std::vector<double>> values;
    std::vector<SurfaceStatistics> statistics;
    Backend backend_used = Backend::AO_P;
};

namespace detail {
inline SurfaceStatistics to_statistics(const kumo_hip_surface_stats& c, const std::vector<int64_t>& minima,
                                       const std::vector<int64_t>& maxima) {
    SurfaceStatistics s;
    s.area = c.area;
    s.positive_area = c.positive_area;
    s.negative_area = c.negative_area;
    s.nonpolar_area = c.nonpolar_area;
    s.minimum = c.minimum;
    s.maximum = c.maximum;
    s.minimum_vertex = static_cast<std::size_t>(c.minimum_vertex);
    s.maximum_vertex = static_cast<std::size_t>(c.maximum_vertex);
    s.average = c.average;
    s.positive_average = c.positive_average;
    s.negative_average = c.negative_average;
    s.positive_variance = c.positive_variance;
    s.negative_variance = c.negative_variance;
    s.total_variance = c.total_variance;
    s.balance = c.balance;
    s.charge_separation = c.charge_separation;
    s.polarity_index = c.polarity_index;
    s.local_minima.assign(minima.begin(), minima.end());
    s.local_maxima.assign(maxima.begin(), maxima.end());
    return s;
}

inline std::vector<int64_t> flat_triangles(const TriangleMesh& mesh) {
    std::vector<int64_t> t;
    t.reserve(3 * mesh.triangles.size());
    for (const auto& tri : mesh.triangles)
        for (std::size_t k : tri)
            t.push_back(static_cast<int64_t>(k));
    return t;
}

inline std::vector<double> flat_vertices(const TriangleMesh& mesh) {
    std::vector<double> v;
    v.reserve(3 * mesh.vertices.size());
    for (const auto& x : mesh.vertices)
        v.insert(v.end(), x.begin(), x.end());
    return v;
}

// A kumo_hip_surface read into host structures: the mesh, then each field's
// values and statistics.
class SurfaceHandle {
public:
    template <typename Make> explicit SurfaceHandle(Make&& make) {
        kumo_hip_surface* s = nullptr;
        check(make(&s));
        ptr_.reset(s);
        int64_t nv = 0, nt = 0;
        check(kumo_hip_surface_sizes(s, &nv, &nt, &nfields_, &backend_));
        std::vector<double> v(static_cast<std::size_t>(3 * nv));
        std::vector<int64_t> t(static_cast<std::size_t>(3 * nt));
        check(kumo_hip_surface_mesh(s, v.data(), t.data()));
        for (int64_t i = 0; i < nv; ++i)
            mesh_.vertices.push_back({v[3 * i], v[3 * i + 1], v[3 * i + 2]});
        for (int64_t i = 0; i < nt; ++i)
            mesh_.triangles.push_back({static_cast<std::size_t>(t[3 * i]), static_cast<std::size_t>(t[3 * i + 1]),
                                       static_cast<std::size_t>(t[3 * i + 2])});
    }
    const TriangleMesh& mesh() const noexcept { return mesh_; }
    int fields() const noexcept { return nfields_; }
    int backend() const noexcept { return backend_; }
    std::pair<std::vector<double>, SurfaceStatistics> field(int f) const {
        std::vector<double> values(mesh_.vertices.size());
        kumo_hip_surface_stats c{};
        check(kumo_hip_surface_field(ptr_.get(), f, values.data(), &c, nullptr, nullptr));
        std::vector<int64_t> mins(static_cast<std::size_t>(c.local_minima)), maxs(static_cast<std::size_t>(c.local_maxima));
        check(kumo_hip_surface_field(ptr_.get(), f, nullptr, &c, mins.data(), maxs.data()));
        return {std::move(values), to_statistics(c, mins, maxs)};
    }

private:
    std::unique_ptr<kumo_hip_surface, int (*)(kumo_hip_surface*)> ptr_{nullptr, kumo_hip_surface_release};
    TriangleMesh mesh_;
    int32_t nfields_ = 0, backend_ = -1;
};
} // namespace detail

// surface_statistics (include/chem/kumo/surface.hpp:76-84).
inline SurfaceStatistics surface_statistics(const TriangleMesh& mesh, std::span<const double> values,
                                            double polar_threshold = 10.0 / hartree_in_kcal_per_mol) {
    if (values.size() != mesh.vertices.size())
        throw std::invalid_argument("surface statistics need one value per vertex");
    const auto v = detail::flat_vertices(mesh);
    const auto t = detail::flat_triangles(mesh);
    kumo_hip_surface_stats c{};
    const auto call = [&](int64_t* mins, int64_t* maxs) {
        detail::check(kumo_hip_surface_statistics(detail::context(), v.data(),
                                                  static_cast<int64_t>(mesh.vertices.size()), t.data(),
                                                  static_cast<int64_t>(mesh.triangles.size()), values.data(),
                                                  polar_threshold, &c, mins, maxs, nullptr, nullptr));
    };
    call(nullptr, nullptr);
    std::vector<int64_t> mins(static_cast<std::size_t>(c.local_minima)), maxs(static_cast<std::size_t>(c.local_maxima));
    call(mins.data(), maxs.data());
    return detail::to_statistics(c, mins, maxs);
}

// density_isosurface (include/chem/kumo/surface.hpp:92-106).
inline TriangleMesh density_isosurface(const Molecule& molecule, double isovalue = 1e-3, double spacing = 0.25) {
    const detail::SurfaceHandle h([&](kumo_hip_surface** s) {
        return kumo_hip_density_isosurface(detail::context(), molecule.handle(), isovalue, spacing, s);
    });
    return h.mesh();
}

// electrostatic_potential_surface (include/chem/kumo/surface.hpp:125-131).
inline PotentialSurface electrostatic_potential_surface(const Molecule& molecule, const SurfaceOptions& options = {}) {
    const detail::SurfaceHandle h([&](kumo_hip_surface** s) {
        return kumo_hip_electrostatic_potential_surface(detail::context(), molecule.handle(), options.isovalue,
                                                        options.spacing, options.potential_tolerance, s);
    });
    auto [values, stats] = h.field(0);
    return {h.mesh(), std::move(values), std::move(stats)};
}

// average_local_ionization_energy (include/chem/kumo/surface.hpp:133-153).
inline std::vector<double> average_local_ionization_energy(const Molecule& molecule, std::span<const Vector3> points) {
    std::vector<double> xyz;
    xyz.reserve(3 * points.size());
    for (const auto& p : points)
        xyz.insert(xyz.end(), p.begin(), p.end());
    std::vector<double> out(points.size());
    detail::check(kumo_hip_average_local_ionization_energy(detail::context(), molecule.handle(), xyz.data(),
                                                           points.size(), out.data()));
    return out;
}

// local_ionization_energy_surface (include/chem/kumo/surface.hpp:168-173).
inline IonizationEnergySurface local_ionization_energy_surface(const Molecule& molecule, double isovalue = 1e-3,
                                                               double spacing = 0.25) {
    const detail::SurfaceHandle h([&](kumo_hip_surface** s) {
        return kumo_hip_local_ionization_energy_surface(detail::context(), molecule.handle(), isovalue, spacing, s);
    });
    auto [values, stats] = h.field(0);
    return {h.mesh(), std::move(values), std::move(stats)};
}

// map_on_density_isosurface (include/chem/kumo/surface.hpp:184-195): one
// entry per distinct property, in the order of first occurrence. The options'
// progress callback is not used on this path.
inline MappedSurface map_on_density_isosurface(const Molecule& molecule, const std::vector<Property>& properties,
                                               double isovalue = 1e-3, double spacing = 0.25,
                                               const EvaluationOptions& options = {}) {
    MappedSurface out;
    std::vector<kumo_hip_property> props;
    for (const auto& p : properties) {
        bool seen = false;
        for (const auto& q : out.properties)
            seen = seen || q == p;
        if (seen) continue;
        out.properties.push_back(p);
        props.push_back({static_cast<int32_t>(p.kind), static_cast<int32_t>(p.component),
                         static_cast<int32_t>(p.channel), static_cast<int32_t>(p.orbital_index)});
    }
    kumo_hip_options o;
    kumo_hip_default_options(&o);
    o.threads = options.threads;
    o.block_size = static_cast<int32_t>(options.block_size);
    o.backend = static_cast<int32_t>(options.backend);
    o.precision = static_cast<int32_t>(options.precision);
    o.screening_threshold = options.screening_threshold;
    o.potential_tolerance = options.potential_tolerance;
    o.fukui_delta = options.fukui_delta;
    const detail::SurfaceHandle h([&](kumo_hip_surface** s) {
        return kumo_hip_map_on_density_isosurface(detail::context(), molecule.handle(), props.data(), props.size(),
                                                  isovalue, spacing, &o, s);
    });
    out.mesh = h.mesh();
    for (int f = 0; f < h.fields(); ++f) {
        auto [values, stats] = h.field(f);
        out.values.push_back(std::move(values));
        out.statistics.push_back(std::move(stats));
    }
    out.backend_used = static_cast<Backend>(h.backend());
    return out;
}

// write_ply (include/chem/kumo/surface.hpp:197-210).
inline void write_ply(const std::filesystem::path& path, const TriangleMesh& mesh,
                      const std::vector<std::pair<std::string, std::vector<double>>>& fields,
                      const std::vector<std::string>& comments = {}) {
    const auto v = detail::flat_vertices(mesh);
    const auto t = detail::flat_triangles(mesh);
    std::vector<const char*> names, notes;
    std::vector<const double*> values;
    for (const auto& [name, vals] : fields) {
        if (vals.size() != mesh.vertices.size())
            throw std::invalid_argument("PLY property '" + name + "' needs one value per vertex");
        names.push_back(name.c_str());
        values.push_back(vals.data());
    }
    for (const auto& c : comments)
        notes.push_back(c.c_str());
    const std::string p = path.string();
    detail::check(kumo_hip_write_ply(detail::context(), p.c_str(), v.data(), static_cast<int64_t>(mesh.vertices.size()),
                                     t.data(), static_cast<int64_t>(mesh.triangles.size()),
                                     static_cast<int32_t>(fields.size()), names.data(), values.data(),
                                     static_cast<int32_t>(comments.size()), notes.data()));
}

} // namespace chem::kumo
