// Host-side basis preparation: normalized AO blocks, Cartesian->pure
// transforms and shell cutoff radii, restating the BasisEvaluator constructor
// of the reference (src/basis_math.cpp:27-66, src/basis_evaluator.cpp:98-202).
#pragma once

#include "device.hpp"

#include <cstdint>
#include <vector>

namespace kumo {

struct CartPower {
    int x, y, z;
};

// Gaussian FCHK order of Cartesian components (src/basis_evaluator.cpp:135-152).
std::vector<CartPower> gaussian_cartesian_order(int l);
// (2l+1) x ncart row-major, rows m = 0, +1, -1, ... (src/basis_evaluator.cpp:155-175).
std::vector<double> gaussian_cartesian_to_pure(int l);

struct HostBlock {
    double origin[3];
    int l = 0, atom = 0;
    std::vector<double> exps;
    std::vector<CartPower> powers;
    std::vector<double> weights; // ncart x nprim row-major
    std::vector<double> pure;    // count x ncart row-major, empty if Cartesian
    int offset = 0, count = 0;
    double cutoff_r2 = 0.0;
    int ncart() const { return static_cast<int>(powers.size()); }
    int nprim() const { return static_cast<int>(exps.size()); }
};

struct ShellRecords {
    int nshells = 0;
    const int32_t* types = nullptr;
    const int32_t* nprim = nullptr;
    const int32_t* atom = nullptr; // one-based
    const double* xyz = nullptr;
    const double* exps = nullptr;
    const double* coef = nullptr;
    const double* pcoef = nullptr;
    int natoms = 0;
    int nbasis = 0;
};

// Angular factor g of Cartesian component p relative to the block's first
// component: primitive_norm(a, p) = g * primitive_norm(a, first) for every a.
double component_factor(const CartPower& first, const CartPower& p);

struct PreparedBasis {
    std::vector<HostBlock> blocks;
    int nao = 0;
    int max_l = 0;
    int operand_rows = 0; // Cartesian rows of all blocks, then pure rows of pure blocks
    // Device tables
    std::vector<DevShell> shells;
    std::vector<double> exps, weights, radw;
    std::vector<PureTerm> pure_terms;
    PureIndex pure_index{};
    std::vector<int32_t> lorder; // blocks by descending l, ascending index within an l
};

// Squared radius per block beyond which the rigorous bound of
// src/basis_evaluator.cpp:98-127 keeps every AO of the block and its first and
// second derivatives at or below `threshold` (the cutoff radius when the
// threshold is below the cutoff's 1e-14).
std::vector<double> screening_radii2(const PreparedBasis& basis, double threshold);

// Validates the shell records as src/basis.cpp:21-120 does and prepares the
// blocks; throws kumo::Error (invalid argument / runtime) on bad records.
PreparedBasis prepare_basis(const ShellRecords& r);

} // namespace kumo
