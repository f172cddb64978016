// Far-field approximation of the electrostatic potential under an error
// budget (far_field_plan, src/electrostatic_potential.cpp:363-699; target
// clusters, src/potential_clusters.hpp:29-105 and src/evaluator.cpp:391-479).
//
// The plan (host, once per molecule, cluster radius and tolerance):
//  - per primitive pair the squared distances from a target-cluster centre
//    beyond which its local expansion of order kFarLocal about that centre is
//    within each of kFarLevels error levels (half the tolerance times
//    10^(-j/2)), from a bound on the Taylor remainder of its multipole terms
//    and on the erfc tail of its Gaussian charge;
//  - source clusters: the pairs grouped by the 2-bohr cube holding their
//    centre and by exponent interval [4^k, 4^(k+1)), each with its multipole
//    moments to order kFarMultipole about the centre of its bounding box, and
//    the distance beyond which its translation to a target cluster is within
//    a share of a quarter of the tolerance.
// On the GPU, per target cluster: the source clusters far enough are
// translated into a local expansion, the other pairs are expanded where the
// pair budget allows (smallest error levels first; at the level that fits
// only in part, pairs in index order) and the rest are summed exactly at
// every point; each point gets nuclear - (local polynomial + exact pairs).
#pragma once

#include "kernels.hpp"

#include <array>
#include <cstdint>
#include <vector>

namespace kumo {

constexpr int kFarLocal = 6;       // order of a pair's local expansion (src/electrostatic_potential.hpp:64)
constexpr int kFarMultipole = 10;  // order of a source cluster's moments and of the local expansion (:66)
constexpr int kFarLevels = 33;     // error levels (:76)
constexpr double kFarCell = 2.0;   // edge of a source cluster's cube, bohr (:77)
constexpr double kFarPointCube = 1.6; // edge of a target cube of listed points (include/chem/kumo/evaluator.hpp:170)
constexpr int kFarTileSide = 8;    // grid points per edge of a target tile (src/evaluator.cpp:399)

__host__ __device__ constexpr int far_hcount(int s) { return s < 0 ? 0 : (s + 1) * (s + 2) * (s + 3) / 6; }

// Source cluster on the device: centre, the distance from a target centre
// beyond which it is translated, its pairs' range in the plan's pair list
// and their largest order L (its moments of shift order m vanish beyond
// total order L + m).
struct FarClusterDev {
    double c[3];
    double min_distance;
    int32_t first, count;
    int32_t max_L, pad;
};

struct FarPlanHost {
    double radius = 0.0, tolerance = 0.0, pair_tolerance = 0.0;
    double levels[kFarLevels] = {};
    int max_L = 0;
    int moment_count = 0;                 // far_hcount(max_L + kFarMultipole)
    std::vector<float> level_d2;          // [pairs][kFarLevels], +inf where no distance reaches the level
    std::vector<FarClusterDev> clusters;
    std::vector<int32_t> cluster_pairs;   // pairs of each cluster, in pair order
    std::vector<int32_t> cluster_of;      // [pairs] the cluster of each pair
    std::vector<double> moments;          // [clusters][kFarMultipole + 1][moment_count]
};

// The plan for the molecule's primitive pairs (the exact ESP's), target
// clusters of `radius` bohr and error `tolerance` > 0; order[k] is the pair
// that comes k-th in the reference's order (source clusters are numbered and
// filled in that order).
FarPlanHost far_field_plan(const std::vector<EspPairDev>& pairs, const std::vector<double>& coef,
                           const std::vector<int32_t>& order, double radius, double tolerance);

// Per-molecule device copy of a plan.
struct FarPlanDev {
    const float* level_d2;
    const FarClusterDev* clusters;
    const int32_t* cluster_pairs;
    const int32_t* cluster_of;
    const double* moments;
    const int32_t* ref_order; // the pairs in the reference's order (the budget's order within a level)
    int nclusters, npairs, max_L, moment_count;
    double pair_tolerance;
    double levels[kFarLevels];
};

// Target clusters of one launch: centre and points. A grid tile holds the
// grid points with indices lo <= (i, j, k) < hi (count = their number) that
// fall in the launch's point range; a cube of listed points holds the point
// indices idx[first + k], k < count.
struct FarTargetDev {
    double c[3];
    int32_t lo[3], hi[3];
    int32_t first, count;
};

// For target clusters [0, ntargets): the local expansion (local[t][far_hcount(kFarMultipole)])
// and the near pairs (near[t][npairs], nnear[t]); work holds npairs +
// nclusters bytes per target.
void launch_far_targets(const FarPlanDev& plan, const EspPairDev* pairs, const double* coef, const double* boys_table,
                        const FarTargetDev* targets, int ntargets, double* local, int32_t* near, int32_t* nnear,
                        int8_t* work, hipStream_t stream);
// The potential at the points of targets [0, ntargets) with indices in [q0,
// q1) (grid tiles; listed points: idx): out[index] = nuclear - (local
// polynomial) - (the sums of its near pairs, one pair group (L) after
// another; the groups are the exact ESP's pair ranges by L).
void launch_far_points(const EspPairDev* pairs, const std::vector<EspGroup>& groups, const double* coef,
                       const double* boys_table, const EspAtomDev* atoms, int natoms, const PointSource& ps,
                       long long q0, long long q1, const FarTargetDev* targets, int ntargets, const int64_t* idx,
                       const double* local, const int32_t* near, const int32_t* nnear, int npairs, double* out,
                       hipStream_t stream);

} // namespace kumo
