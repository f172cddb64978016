// Data layouts shared by the host code and the HIP kernels.
//
// A molecule lives in HBM as structure-of-arrays tables built once at load:
//  - shells[]: one entry per normalized AO block (an SP shell gives an s and a
//    p block), in FCHK AO order;
//  - exps[], weights[]: primitive exponents and contraction weights
//    (coefficient x primitive normalization), weights row-major per block
//    (Cartesian component x primitive);
//  - pure[]: sparse Cartesian->pure transform entries per angular momentum,
//    ordered by Cartesian component so that each pure AO accumulates its terms
//    in increasing component order;
//  - radw[]: per primitive the contraction weight of the block's first
//    Cartesian component; weights[k][p] = g[k] radw[p] with g[k] depending on
//    l and the component only, so a block's components are g[k] times
//    x^i y^j z^k R(r) with R(r) = sum_p radw[p] exp(-a_p r^2);
//  - the MO operand table of the AO->MO contraction: per block, Cartesian rows
//    g[k] (T^T C)[k] (the pure transform and the component factor folded into
//    the occupied-MO coefficients), then for pure blocks the plain pure rows
//    C[r]; one row per [MO slot] range (alpha then beta, each padded to 16),
//    in FP64, FP32 and three bf16 split planes;
//  - P total and P spin, row-major FP64.
#pragma once

#include <cstdint>

namespace kumo {

constexpr int kMaxL = 5;                 // h shells; higher l is rejected at load
constexpr int kMaxCart = (kMaxL + 1) * (kMaxL + 2) / 2;
constexpr int kMaxPure = 2 * kMaxL + 1;
constexpr int kMaxPrim = 64;

struct DevShell {
    double o[3];
    double cutoff_r2;
    int32_t l;
    int32_t nprim;
    int32_t prim;   // first primitive in exps[]
    int32_t wofs;   // first weight in weights[] (ncart x nprim, row-major)
    int32_t ncart;
    int32_t count;  // AOs of the block
    int32_t ao;     // first AO
    int32_t pure;   // 1 if the block is pure (l >= 2)
    int32_t atom;
    int32_t crow;   // first Cartesian row of the block in the MO operand table
    int32_t prow;   // first pure row of the block in the MO operand table (-1 if Cartesian)
    int32_t pad;
};

// A block's data for the contraction kernels in one contiguous record, so
// that a lane reads everything it needs for a row with independent loads:
// centre, cutoff, l and up to kRecPrim primitives (exponent, factored
// weight radw, skip radius); blocks with more primitives read exps[] /
// radw[] instead. Beyond its skip radius, r^2 > skip[p], primitive p adds
// less than 1e-40 to each radial sum (|radw| max(1, 4 a^2) exp(-a r^2) <
// 1e-40), and the wave-uniform contraction kernel leaves its exponential
// out.
constexpr int kRecPrim = 8;
struct ShellRec {
    double o[3];
    double cutoff_r2;
    int32_t l, nprim, prim, pad;
    double e[kRecPrim];
    double w[kRecPrim];
    double skip[kRecPrim];
};

// Sparse transform entry: pure row `row` += coef * Cartesian component `cart`.
struct PureTerm {
    double coef;
    int32_t cart;
    int32_t row;
};

// Offsets of the pure terms of angular momentum l in the PureTerm table.
struct PureIndex {
    int32_t first[kMaxL + 2];
};

// Point source of one evaluation: a grid or an explicit list.
struct PointSource {
    double origin[3];
    double steps[9]; // column-major: column a = step vector of axis a
    int64_t ny, nz;
    const double* points; // xyz per point, or null for a grid
};

// Per-point moments of one density channel, as DensityDerivatives
// (include/chem/kumo/density.hpp:151-160): value, gradient x y z, tau terms
// x y z, Hessian xx xy xz yy yz zz.
constexpr int kMoments = 13;
enum Moment { M_VALUE = 0, M_GX, M_GY, M_GZ, M_TX, M_TY, M_TZ, M_HXX, M_HXY, M_HXZ, M_HYY, M_HYZ, M_HZZ };

} // namespace kumo
