// Request planning: which density components need the point path, the AO
// derivative order, the contraction backend. Restates
// src/evaluation_plan.hpp:37-273 and src/automatic_backend.hpp:38-50 so that
// the HIP backend screens with exactly the reference's AO order.
#pragma once

#include "common.hpp"

#include "../../include/kumo_hip.h"

#include <algorithm>
#include <array>
#include <string>
#include <vector>

namespace kumo {

enum Kind {
    kDensity = 0, kGradient, kHessian, kLaplacian, kTau, kOrbital, kElf, kReducedGradient, kSignLambda2,
    kEsp, kLol, kIri, kDori, kKinetic, kPotentialEnergy, kEnergyDensity, kSpinPolarization, kShannon, kFisher,
    kAlie, kLea, kLeae, kElectronegativity, kHardness, kOrbitalGradient, kOrbitalHessian, kFukuiPlus, kFukuiMinus,
    kFukuiZero, kDualDescriptor, kTauComponents, kKineticComponents, kVirialRatio, kEnergyPerElectron,
    kKineticPerElectron, kElfTs, kLolTs, kKindCount
};
enum Component { kTotal = 0, kAlpha = 1, kBeta = 2, kDifference = 3 };

inline int channels(int kind) {
    switch (kind) {
    case kGradient: case kOrbitalGradient: case kTauComponents: case kKineticComponents: return 3;
    case kHessian: case kOrbitalHessian: return 6;
    default: return 1;
    }
}
inline bool is_orbital_property(int k) { return k == kOrbital || k == kOrbitalGradient || k == kOrbitalHessian; }
inline bool is_orbital_energy(int k) { return (k >= kAlie && k <= kHardness) || (k >= kFukuiPlus && k <= kDualDescriptor); }

struct ComponentPlan {
    bool density = false, gradient = false, hessian = false, laplacian = false, tau = false;
    bool elf = false, elf_input = false, reduced_gradient = false, sign_lambda2 = false;
    bool lol = false, lol_input = false, ge_input = false, gradient_input = false;
    bool hessian_input = false, energy_input = false, tau_input = false, value_input = false, esp = false;
    bool diagonal_input = false; // the Hessian diagonal element by element
    bool derivatives() const {
        return gradient || hessian || laplacian || tau || elf_input || reduced_gradient || sign_lambda2 || lol_input ||
               gradient_input || hessian_input || energy_input || tau_input || ge_input || value_input;
    }
    int ao_order() const {
        if (hessian || laplacian || sign_lambda2 || hessian_input || energy_input || ge_input) return 2;
        if (gradient || tau || elf_input || reduced_gradient || lol_input || gradient_input || tau_input || value_input)
            return 1;
        return density ? 0 : -1;
    }
    // Full Hessian moments (all six) are needed by Hessian, sign(lambda2) rho
    // and DORI; the others use its diagonal only.
    bool full_hessian() const { return hessian || sign_lambda2 || hessian_input; }
    // The other second-derivative users need only the Laplacian, except the
    // kinetic-energy components (tau_a - lap_a / 4).
    bool hessian_diagonal() const { return diagonal_input; }
};

struct Plan {
    std::array<ComponentPlan, 4> c{};
    bool orbitals = false;
    // Orbital-energy fields (src/evaluation_plan.hpp:93-95): the occupied
    // sums, plus the virtual-orbital sums, plus the orbital-weighted Fukui sums.
    bool orbital_energies = false, virtual_orbitals = false, orbital_weighted_fukui = false;
    int max_order = -1;
    bool esp() const { return c[kTotal].esp; }
};

// orbital_energy_name (src/evaluation_plan.hpp:112-128).
inline const char* orbital_energy_name(int kind) {
    switch (kind) {
    case kAlie: return "the average local ionization energy";
    case kLea: return "the local electron affinity";
    case kLeae: return "the local electron attachment energy";
    case kElectronegativity: return "the local electronegativity";
    case kHardness: return "the local hardness";
    case kFukuiPlus: return "the orbital-weighted Fukui function f+";
    case kFukuiMinus: return "the orbital-weighted Fukui function f-";
    case kFukuiZero: return "the orbital-weighted Fukui function f0";
    default: return "the orbital-weighted dual descriptor";
    }
}

inline void require_total(const kumo_hip_property& p, const char* what) {
    if (p.component != kTotal) invalid(std::string(what) + " is defined only for the total density");
}

inline void spin_inputs(Plan& plan, const kumo_hip_property& p, bool ComponentPlan::*input, const char* what) {
    if (p.component == kDifference) invalid(std::string(what) + " is not defined for the spin-difference density");
    if (p.component == kTotal) {
        plan.c[kAlpha].*input = true;
        plan.c[kBeta].*input = true;
    } else {
        plan.c[static_cast<std::size_t>(p.component)].*input = true;
    }
}

inline Plan make_plan(const kumo_hip_property* props, std::size_t n) {
    Plan plan;
    for (std::size_t i = 0; i < n; ++i) {
        const auto& p = props[i];
        if (p.kind < 0 || p.kind >= kKindCount) invalid("unknown property kind");
        if (p.component < 0 || p.component > 3) invalid("invalid density spin component");
        if (is_orbital_property(p.kind)) {
            if (p.channel != 0 && p.channel != 1) invalid("MO spin channel must be alpha or beta");
            if (p.orbital_index == 0) invalid("orbital indices are one-based");
            if (p.orbital_index < 0) out_of_range("MO index outside available coefficient range");
            plan.orbitals = true;
            plan.max_order = std::max(plan.max_order, 0);
            continue;
        }
        auto& c = plan.c[static_cast<std::size_t>(p.component)];
        switch (p.kind) {
        case kDensity: c.density = true; break;
        case kGradient: c.gradient = true; break;
        case kHessian: c.hessian = true; break;
        case kLaplacian: c.laplacian = true; break;
        case kTau: c.tau = true; break;
        case kElf: c.elf = true; spin_inputs(plan, p, &ComponentPlan::elf_input, "ELF"); break;
        case kReducedGradient: require_total(p, "the reduced density gradient"); c.reduced_gradient = true; break;
        case kSignLambda2: require_total(p, "sign(lambda2) rho"); c.sign_lambda2 = true; break;
        case kEsp: require_total(p, "the electrostatic potential"); c.esp = true; break;
        case kLol: c.lol = true; spin_inputs(plan, p, &ComponentPlan::lol_input, "LOL"); break;
        case kElfTs: spin_inputs(plan, p, &ComponentPlan::ge_input, "the Tsirelson-Stash ELF"); break;
        case kLolTs: spin_inputs(plan, p, &ComponentPlan::ge_input, "the Tsirelson-Stash LOL"); break;
        case kIri: require_total(p, "the interaction region indicator"); c.gradient_input = true; break;
        case kShannon: require_total(p, "the Shannon entropy density"); c.gradient_input = true; break;
        case kFisher: require_total(p, "the Fisher information density"); c.gradient_input = true; break;
        case kDori: require_total(p, "the density overlap regions indicator"); c.hessian_input = true; break;
        case kKinetic: c.energy_input = true; break;
        case kPotentialEnergy: require_total(p, "the potential energy density"); c.energy_input = true; break;
        case kEnergyDensity: require_total(p, "the energy density"); c.energy_input = true; break;
        case kKineticComponents: c.energy_input = true; c.diagonal_input = true; break;
        case kTauComponents: c.tau_input = true; break;
        case kVirialRatio: require_total(p, "|V| / G"); c.energy_input = true; break;
        case kEnergyPerElectron: require_total(p, "the energy density per electron"); c.energy_input = true; break;
        case kKineticPerElectron: require_total(p, "the kinetic energy per electron"); c.tau_input = true; break;
        case kSpinPolarization:
            if (p.component != kDifference) invalid("the spin polarization is defined only for the spin component");
            plan.c[kTotal].value_input = true;
            c.value_input = true;
            break;
        case kAlie:
            require_total(p, orbital_energy_name(p.kind));
            plan.orbital_energies = true;
            break;
        case kLea: case kLeae: case kElectronegativity: case kHardness:
            require_total(p, orbital_energy_name(p.kind));
            plan.orbital_energies = plan.virtual_orbitals = true;
            break;
        case kFukuiPlus: case kFukuiMinus: case kFukuiZero: case kDualDescriptor:
            require_total(p, orbital_energy_name(p.kind));
            plan.orbital_energies = plan.orbital_weighted_fukui = true;
            break;
        default:
            invalid("unknown property kind");
        }
    }
    for (const auto& c : plan.c)
        plan.max_order = std::max(plan.max_order, c.ao_order());
    return plan;
}

inline int choose_backend(int requested, bool eligible, const Plan& plan) {
    if (requested == KUMO_HIP_AO_P) return KUMO_HIP_AO_P;
    if (requested == KUMO_HIP_AO_MO) {
        if (!eligible) invalid("AO_MO requested for ineligible orbital representation");
        return KUMO_HIP_AO_MO;
    }
    if (requested != KUMO_HIP_AUTO) invalid("unknown evaluation backend");
    if (!eligible) return KUMO_HIP_AO_P;
    for (const auto& c : plan.c)
        if (c.ao_order() >= 0) return KUMO_HIP_AO_MO;
    return KUMO_HIP_AO_P;
}

inline bool uses_point_path(const ComponentPlan& c, int backend, bool grid) {
    return c.derivatives() || ((backend == KUMO_HIP_AO_MO || !grid) && c.density);
}
inline bool on_grid_path(const kumo_hip_property& p, int backend, bool grid) {
    return grid && (p.kind == kOrbital || (p.kind == kDensity && backend == KUMO_HIP_AO_P));
}
inline bool from_point_path(const kumo_hip_property& p, int backend, bool grid) {
    return !on_grid_path(p, backend, grid) && !is_orbital_property(p.kind) && p.kind != kEsp &&
           !is_orbital_energy(p.kind);
}

} // namespace kumo
