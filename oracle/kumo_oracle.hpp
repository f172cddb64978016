// ORACLE — test infrastructure only. A CPU FP64 restatement of the reference
// chem-kumo 2.1.0 FCHK->grid evaluator, used as the correctness checker for
// the HIP path and as the timed CPU baseline. Only tests/, the smoke check of
// __graft_entry__.py and the CPU-baseline leg of bench.py may load it; the
// product library never links it.
//
// Every function cites the reference file:line it restates. The arithmetic
// order of the AO values, the shell cutoff, the screening masks and the grid
// coordinates follows the reference (no FMA contraction: built with
// -ffp-contract=off). Eigen's GEMM summation order cannot be reproduced
// without Eigen; sums over AOs and MOs here run sequentially.
//
// Pinning status of each function (how a plausible mistake would be caught)
// is listed in tests/test_oracle_*.py; functions without an external pin say
// "parity unpinned" there.
#pragma once

#include <cstddef>
#include <cstdint>

extern "C" {

/// Property keys shared with the product C-ABI: `kind` uses the numbering of
/// the reference enum PropertyKind (include/chem/kumo/evaluator.hpp:112-150),
/// `component` DensityComponent (include/chem/kumo/spin.hpp:8), `channel`
/// SpinChannel (spin.hpp:12), `orbital_index` is one-based.
struct orc_property {
    int kind;
    int component;
    int channel;
    int orbital_index;
};

struct orc_options {
    int block_size;             // 1..65536 (reference default 128)
    int backend;                // 0 AUTO, 1 AO_P, 2 AO_MO
    double screening_threshold; // reference default 1e-12
    double fukui_delta;         // hartree, > 0 (reference default 0.1)
    double potential_tolerance; // hartree/e, >= 0 (reference default 1e-8); 0 is exact
};

/// Basis from the raw FCHK shell records (src/basis.cpp:21-120).
void* orc_basis_create(int nshell, const int* shell_types, const int* nprim, const int* shell_atom,
                       const double* shell_xyz, const double* exponents, const double* coefficients,
                       const double* p_coefficients, char* error, int error_len);
void orc_basis_destroy(void* basis);
int orc_basis_nao(const void* basis);
int orc_basis_nblocks(const void* basis);
/// Per normalized block: l, AO offset, AO count, cutoff r^2.
void orc_basis_blocks(const void* basis, int* l, int* offset, int* count, double* cutoff_r2);
/// AO values (order 0: evaluate_batch) or derivative panels (order 1, 2:
/// evaluate_derivatives_batch). Output column-major nao x (panels*n), as the
/// reference (include/chem/kumo/basis_evaluator.hpp:104-117).
int orc_ao_eval(const void* basis, int n, const double* points, int order, double* out);
/// Analytic overlap matrix (src/basis_evaluator.cpp:210-278), row-major nao x nao.
void orc_overlap(const void* basis, double* out);
/// Cartesian-to-pure matrix (src/basis_evaluator.cpp:155-175), row-major.
int orc_pure_transform(int l, double* out);

/// Molecule: P total/spin (row-major n x n), MO coefficients as read from an
/// FCHK record (MO-major, AO fastest: src/mo_coefficients.hpp:21-34).
void* orc_molecule_create(void* basis, int natoms, const int* atomic_numbers,
                          const double* charges, const double* xyz, const double* p_total,
                          const double* p_spin, int nmo_alpha, const double* c_alpha,
                          int nmo_beta, const double* c_beta, int nalpha, int nbeta,
                          char* error, int error_len);
void orc_molecule_destroy(void* molecule);
/// SCF orbital energies of each channel (FCHK "Alpha/Beta Orbital Energies").
void orc_molecule_set_energies(void* molecule, int n_alpha, const double* e_alpha, int n_beta,
                               const double* e_beta);
int orc_molecule_eligible(const void* molecule);

/// Points: a grid (origin[3], steps[9] column-major as Eigen Matrix3d,
/// counts[3]) when `points` is null, otherwise n points (x,y,z per point).
/// Outputs: outputs[p] has channels(kind) x npoints doubles. Returns 0 on
/// success; nonzero with a message in `error` otherwise. `backend_used` gets
/// the resolved backend (1 AO_P, 2 AO_MO).
int orc_evaluate(const void* molecule, const double* origin, const double* steps,
                 const int* counts, std::int64_t npoints, const double* points,
                 const orc_property* props, int nprops, const orc_options* options,
                 double** outputs, int* backend_used, char* error, int error_len);

/// Kept-AO masks of the point path: one byte per (block, AO), blocks of
/// `block_size` consecutive points, AO order `ao_order` (0, 1, 2).
int orc_masks(const void* molecule, const double* origin, const double* steps, const int* counts,
              std::int64_t npoints, const double* points, int block_size, int ao_order,
              double threshold, std::uint8_t* kept);

/// Boys function F_0..F_m(T) (src/boys_function.cpp:47-76).
int orc_boys(double T, int m, double* out);
/// Exact electrostatic potential at n points (src/electrostatic_potential.cpp:117-361).
int orc_esp(const void* molecule, std::int64_t n, const double* points, double* out);
std::int64_t orc_esp_pair_count(const void* molecule);
/// The potential of a molecule (pairs contracted with P once) for the
/// far-field functions below; null with a message on error.
void* orc_esp_create(const void* molecule, char* error, int error_len);
void orc_esp_destroy(void* esp);
/// far_field_plan (src/electrostatic_potential.cpp:497-501) for target
/// clusters of `radius` bohr; null with a message on error.
void* orc_far_plan(const void* esp, double radius, double tolerance, char* error, int error_len);
void orc_far_plan_destroy(void* plan);
/// sizes[0..5]: pairs, finite level entries, source clusters, moment_count,
/// largest pair order, and the pair tolerance and radius in values[0..1].
void orc_far_plan_info(const void* plan, std::int64_t* sizes, double* values);
/// Source cluster c: centre[3], radius, min_distance, its pair count, and
/// its moments of full order (moment_count values, (t, u, v) ordered by
/// t + u + v, then t and u descending) when `moments` is not null.
int orc_far_plan_cluster(const void* plan, std::int64_t c, double* centre, double* radius, double* min_distance,
                         std::int64_t* npairs, double* moments);
/// Level distances of pair i (33 squared distances as floats).
int orc_far_plan_levels(const void* plan, std::int64_t i, float* d2);
/// evaluate_cluster (src/electrostatic_potential.cpp:701-728) of n points
/// about `centre`; stats[0..2]: source clusters translated, pairs expanded,
/// pairs evaluated exactly. Returns 0, or 1 with a message.
int orc_far_cluster(const void* esp, const void* plan, std::int64_t n, const double* points, const double* centre,
                    double* out, std::int64_t* stats, char* error, int error_len);
/// Grid point coordinates (src/grid_geometry.hpp:59-68), 3 x n.
void orc_grid_points(const double* origin, const double* steps, const int* counts,
                     std::int64_t start, std::int64_t n, double* out);
/// Gaussian Cube file of a field (src/cube.cpp:70-111, 253-262); orbital
/// nonzero: an orbital Cube with the one-based orbital_index. Returns 0, 1
/// (invalid argument: nonfinite sample, orbital index) or 3 (file error).
int orc_write_field_cube(const char* path, const char* comment, const double* origin, const double* steps,
                         const int* counts, int natoms, const int* z, const double* charges, const double* xyz,
                         const double* values, std::int64_t n, int orbital, std::int64_t orbital_index, char* error,
                         int error_len);
/// Scalar function for isosurfaces: values (n) and gradients (3 per point) at n points.
typedef void (*orc_scalar_fn)(const double* points, std::int64_t n, double* values, double* gradients, void* user);
/// extract_isosurface (src/surface.cpp:150-331) of grid samples; the mesh is
/// kept in an opaque handle (orc_mesh_*). Returns null with a message on error.
void* orc_isosurface(const double* origin, const double* steps, const int* counts, const double* samples,
                     double isovalue, orc_scalar_fn fn, void* user, int* status, char* error, int error_len);
/// density_isosurface (src/surface.cpp:341-369, 531-533) from the oracle's densities.
void* orc_density_isosurface(const void* molecule, double isovalue, double spacing, int* status, char* error,
                             int error_len);
void orc_mesh_sizes(const void* mesh, std::int64_t* nvertices, std::int64_t* ntriangles);
void orc_mesh_data(const void* mesh, double* vertices, std::int64_t* triangles);
void orc_mesh_free(void* mesh);
/// surface_statistics (src/surface.cpp:418-529): stats[17] in the order area,
/// positive_area, negative_area, nonpolar_area, minimum, maximum,
/// minimum_vertex, maximum_vertex, average, positive_average,
/// negative_average, positive_variance, negative_variance, total_variance,
/// balance, charge_separation, polarity_index; counts[2] the numbers of local
/// minima and maxima (their vertices written when the arrays are not null).
int orc_surface_statistics(const double* vertices, std::int64_t nvertices, const std::int64_t* triangles,
                           std::int64_t ntriangles, const double* values, double polar_threshold, double* stats,
                           std::int64_t* counts, std::int64_t* minima, std::int64_t* maxima, char* error,
                           int error_len);
int orc_threads(void);
// Threads of the following calls (OpenMP); nonpositive keeps the current count.
void orc_set_threads(int threads);
}
