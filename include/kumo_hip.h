/*
 * kumo_hip — C-ABI of the MI355X (gfx950) FCHK->grid field evaluator.
 *
 * The drop-in boundary for chem-kumo's evaluation path. Every entry point is
 * extern "C", takes plain pointers and sizes, and returns a status code
 * (KUMO_HIP_OK on success). Status codes map one to one onto the exception
 * types the reference throws (include/chem/kumo/evaluator.hpp:497-506):
 * INVALID_ARGUMENT -> std::invalid_argument, OUT_OF_RANGE -> std::out_of_range,
 * RUNTIME_ERROR / DEVICE_ERROR -> std::runtime_error. The message of the last
 * failure on a context is returned by kumo_hip_last_error().
 *
 * Property keys, density components, spin channels and backends use the
 * numbering of the reference enums (include/chem/kumo/evaluator.hpp:112-150,
 * 403; include/chem/kumo/spin.hpp:8-12).
 */
#ifndef KUMO_HIP_H
#define KUMO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum kumo_hip_status {
    KUMO_HIP_OK = 0,
    KUMO_HIP_INVALID_ARGUMENT = 1,
    KUMO_HIP_OUT_OF_RANGE = 2,
    KUMO_HIP_RUNTIME_ERROR = 3,
    KUMO_HIP_DEVICE_ERROR = 4,
    KUMO_HIP_ABORTED = 5
};

/* Backend {AUTO, AO_P, AO_MO} (include/chem/kumo/evaluator.hpp:403): the
 * density contraction; every one of them runs on the GPU here. */
enum kumo_hip_backend { KUMO_HIP_AUTO = 0, KUMO_HIP_AO_P = 1, KUMO_HIP_AO_MO = 2 };

/* Arithmetic of the dense contractions (AO->MO and AO.P). AO values, the
 * screening masks, the per-point reductions, the derived fields, the Boys
 * function and the ESP are FP64 in every mode. */
enum kumo_hip_precision {
    KUMO_HIP_FP64 = 0,
    KUMO_HIP_FP32 = 1,
    KUMO_HIP_BF16X3 = 2, /* hi/lo bf16 split, 3 MFMA products, FP32 accumulation */
    KUMO_HIP_BF16X6 = 3  /* hi/mid/lo bf16 split, 6 MFMA products, FP32 accumulation */
};

typedef struct kumo_hip_context kumo_hip_context;
typedef struct kumo_hip_molecule kumo_hip_molecule;

/* Property key (include/chem/kumo/evaluator.hpp:166-180). orbital_index is
 * one-based and used by the Orbital kinds only. */
typedef struct kumo_hip_property {
    int32_t kind;
    int32_t component;
    int32_t channel;
    int32_t orbital_index;
} kumo_hip_property;

/* Regular grid r(i,j,k) = origin + steps * (i,j,k), k fastest
 * (include/chem/kumo/density.hpp:61-72). steps is column-major: column a
 * (steps[3a..3a+2]) is the step vector of axis a. */
typedef struct kumo_hip_grid {
    double origin[3];
    double steps[9];
    int32_t counts[3];
} kumo_hip_grid;

/* Progress report (include/chem/kumo/density.hpp:34-39); a nonzero return
 * aborts the evaluation with KUMO_HIP_ABORTED. */
typedef int (*kumo_hip_progress_fn)(size_t done, size_t total, void* user);

/* EvaluationOptions (include/chem/kumo/evaluator.hpp:406-441) plus the
 * precision of the contractions. Zero-initialise, then call
 * kumo_hip_default_options(). */
typedef struct kumo_hip_options {
    int32_t threads;            /* 0 or 1, as the reference */
    int32_t block_size;         /* 1..65536, points per screening block */
    int32_t backend;            /* enum kumo_hip_backend */
    int32_t precision;          /* enum kumo_hip_precision */
    double screening_threshold; /* >= 0; 0 keeps every AO */
    double potential_tolerance; /* >= 0: error bound of the ESP's far-field approximation (0: exact) */
    double fukui_delta;         /* > 0 */
    int32_t device_outputs;     /* 1: outputs are device pointers (HBM) */
    int32_t tile_points;        /* AO->MO workgroup tile in points: 16 or 32 (the tile kernel:
                                   identical results for both) or 64 (the wave-uniform kernel of
                                   the split-bf16 precisions without full Hessians: the same
                                   products, its own Phi arithmetic); 0: the wave-uniform kernel
                                   where it applies, else chosen by the autotuner (timed once per
                                   molecule, precision and derivative order). */
    kumo_hip_progress_fn progress;
    void* progress_user;
    int32_t esp_far_field;      /* ESP at potential_tolerance > 0: KUMO_HIP_ESP_FAR takes the
                                   reference's far field (8^3 grid tiles, 1.6-bohr point cubes;
                                   within the tolerance of the exact potential);
                                   KUMO_HIP_ESP_AUTO and KUMO_HIP_ESP_EXACT sum every pair
                                   exactly, which is within any tolerance and on this GPU the
                                   faster of the two. Batches and allocation-free evaluations
                                   always sum exactly. */
} kumo_hip_options;

enum kumo_hip_esp_method { KUMO_HIP_ESP_AUTO = 0, KUMO_HIP_ESP_FAR = 1, KUMO_HIP_ESP_EXACT = 2 };

/* Molecule from raw FCHK records (src/basis.cpp:21-120, src/molecule.cpp:153-180).
 * Matrices are row-major n x n (full, symmetric); MO coefficients as in the
 * FCHK record: MO-major with the AO index fastest (src/mo_coefficients.hpp:21-34).
 * c_beta is NULL for a restricted calculation; p_spin may be NULL (zero);
 * energies may be NULL. */
typedef struct kumo_hip_molecule_desc {
    int32_t natoms;
    const int32_t* atomic_numbers;
    const double* nuclear_charges; /* NULL: the atomic numbers */
    const double* coordinates;     /* 3 * natoms, bohr */
    int32_t nshells;
    const int32_t* shell_types;    /* FCHK "Shell types": -1 SP, <= -2 pure */
    const int32_t* primitives_per_shell;
    const int32_t* shell_to_atom;  /* one-based */
    const double* shell_coordinates; /* 3 * nshells */
    const double* exponents;
    const double* coefficients;
    const double* p_coefficients;  /* SP shells; may be NULL without SP shells */
    int32_t nbasis;
    const double* p_total;
    const double* p_spin;
    int32_t nalpha, nbeta;
    int32_t nmo_alpha, nmo_beta;
    const double* c_alpha;
    const double* c_beta;
    const double* e_alpha;
    const double* e_beta;
} kumo_hip_molecule_desc;

typedef struct kumo_hip_molecule_info {
    int32_t natoms, nbasis, nblocks, max_l;
    int32_t nalpha, nbeta, nmo_alpha, nmo_beta;
    int32_t unrestricted;
    int32_t ao_mo_eligible;             /* ElectronicStructureInfo::ao_mo_eligible() */
    double total_relative_error, spin_relative_error;
    int64_t device_bytes;               /* HBM held by the molecule */
} kumo_hip_molecule_info;

/* ---- context (replaces nothing in the reference: it owns the device) ---- */
int kumo_hip_init(int device, kumo_hip_context** ctx);
int kumo_hip_release(kumo_hip_context* ctx);
/* Copy of the last error message of the context (NUL-terminated). */
int kumo_hip_last_error(const kumo_hip_context* ctx, char* buffer, size_t length);
void kumo_hip_default_options(kumo_hip_options* options);

/* ---- molecule residency ----
 * Molecule::from_fchk (include/chem/kumo/density.hpp:215-221): reads a plain
 * or gzip FCHK, prepares the basis (BasisEvaluator ctor,
 * src/basis_evaluator.cpp:177-202), checks the MOs against P
 * (src/backend_selection.hpp:43-78) and keeps basis, P and C in HBM. */
int kumo_hip_molecule_from_fchk(kumo_hip_context* ctx, const char* path, kumo_hip_molecule** mol);
/* The same from records in memory. */
int kumo_hip_molecule_create(kumo_hip_context* ctx, const kumo_hip_molecule_desc* desc,
                             kumo_hip_molecule** mol);
int kumo_hip_molecule_info_get(const kumo_hip_molecule* mol, kumo_hip_molecule_info* info);
int kumo_hip_molecule_release(kumo_hip_molecule* mol);
/* Prepared AO blocks (src/basis_evaluator.cpp:177-202): per block the angular
 * momentum, first AO, AO count and squared cutoff radius. Arrays of nblocks. */
int kumo_hip_molecule_blocks(const kumo_hip_molecule* mol, int32_t* l, int32_t* offset, int32_t* count,
                             double* cutoff_r2);
/* Atoms: atomic numbers, charges, coordinates (3 per atom). */
int kumo_hip_molecule_atoms(const kumo_hip_molecule* mol, int32_t* z, double* charges, double* xyz);

/* ---- evaluation (Evaluator::evaluate, include/chem/kumo/evaluator.hpp:487-538) ----
 * outputs holds one pointer per component of every property, in request
 * order (sum over properties of channels(kind)); each receives npoints
 * doubles in Field order (grid) or point order. backend_used may be NULL. */
int kumo_hip_evaluate_grid(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const kumo_hip_grid* grid,
                           const kumo_hip_property* properties, size_t nproperties,
                           const kumo_hip_options* options, double* const* outputs, int32_t* backend_used);
/* points: xyz of each point (3 * npoints doubles, host memory; device memory
 * when options->device_outputs is set). */
int kumo_hip_evaluate_points(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points,
                             size_t npoints, const kumo_hip_property* properties, size_t nproperties,
                             const kumo_hip_options* options, double* const* outputs, int32_t* backend_used);

/* ---- allocation-free evaluation ----
 * kumo_hip_evaluate_points without device memory allocation, for callers that
 * cannot allocate: code under HIP graph capture, tight serving loops. No
 * counterpart in the reference (it has no device memory). points (3 * npoints
 * doubles) and the outputs are device pointers; the work is issued on
 * `stream` (a hipStream_t; NULL: the context's stream) and the call returns
 * without synchronising it. Density-family and derived-field properties only
 * (KUMO_HIP_INVALID_ARGUMENT for MO amplitudes, the electrostatic potential
 * and orbital-energy fields, and for more than 16 properties); the screening
 * threshold must be the default 1e-12 or one an allocating evaluation of the
 * molecule has used. Workspaces come from one 64 MiB device array per GPU,
 * allocated with the library: points beyond one batch of it are evaluated in
 * several batches, and a molecule whose single screening block does not fit
 * is rejected with KUMO_HIP_INVALID_ARGUMENT. Calls that overlap on one GPU
 * share that array and race; results equal kumo_hip_evaluate_points'. */
int kumo_hip_evaluate_points_noalloc(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points,
                                     size_t npoints, const kumo_hip_property* properties, size_t nproperties,
                                     const kumo_hip_options* options, double* const* outputs, void* stream,
                                     int32_t* backend_used);

/* ---- kernel-level seams used by the parity tests ----
 * BasisEvaluator::evaluate_batch / evaluate_derivatives_batch
 * (include/chem/kumo/basis_evaluator.hpp:99-117): order 0, 1 or 2; out is
 * column-major nbasis x (panels * npoints), panels = 1, 4 or 10. */
int kumo_hip_ao_eval(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points, size_t npoints,
                     int32_t order, double* out);
/* significant_aos per screening block (src/ao_screening.hpp:35-50) as the
 * point path applies it (src/density.cpp:454-467): one byte per
 * (block, AO); grid may be NULL when points is given. */
int kumo_hip_screening_masks(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const kumo_hip_grid* grid,
                             const double* points, size_t npoints, int32_t block_size, int32_t ao_order,
                             double threshold, uint8_t* kept);
/* Boys function F_0..F_m(T) on the device (src/boys_function.cpp:47-76),
 * out is n x (m + 1). */
int kumo_hip_boys(kumo_hip_context* ctx, const double* T, size_t n, int32_t m, double* out);
/* The Boys function of the ESP pair loop (src/boys_function.cpp:47-76 as the
 * pair sum of src/electrostatic_potential.cpp:286-330 calls it): the same
 * table + Taylor / closed-form scheme with constant divisions turned into
 * products and the square root above T = 40 from the hardware estimate,
 * orders 0..top (0 <= top <= 10, the pair orders of shells up to h), out is
 * n x (top + 1). For testing that variant on its own. */
int kumo_hip_boys_esp(kumo_hip_context* ctx, const double* T, size_t n, int32_t top, double* out);

/* ---- batched evaluation ----
 * A fixed list of (molecule, grid) items sharing one property list and one
 * set of options, evaluated across `streams` HIP streams (0: 4). outputs
 * holds, for item i, its channel pointers at outputs[i * nchannels + c]
 * (nchannels = sum over properties of their component counts); they are
 * device pointers and must stay valid while the batch lives. With use_graph,
 * the first kumo_hip_batch_run records every launch of every item into one
 * HIP graph and each later run replays it; without it, each run issues the
 * launches again. Each item's results equal kumo_hip_evaluate_grid's on that
 * item (results do not depend on how items share the device). The options'
 * progress callback is not used. No counterpart in the reference (its
 * Evaluator evaluates one grid per call); it serves callers that evaluate
 * many molecules per step. */
typedef struct kumo_hip_batch kumo_hip_batch;
int kumo_hip_batch_create(kumo_hip_context* ctx, size_t nitems, const kumo_hip_molecule* const* molecules,
                          const kumo_hip_grid* grids, const kumo_hip_property* properties, size_t nproperties,
                          const kumo_hip_options* options, int32_t streams, double* const* outputs,
                          kumo_hip_batch** batch);
int kumo_hip_batch_run(kumo_hip_batch* batch, int32_t use_graph);
/* Device time (ms) of the last run, measured with HIP events around it. */
int kumo_hip_batch_last_ms(const kumo_hip_batch* batch, double* ms);
/* kumo_hip_batch_run with use_graph = 2 issues every launch again without a
 * graph but, as in the graph, without host synchronisation (the streams'
 * launches overlap as on replay), bracketing each contraction, FP64 rescue
 * and field-reduction launch with HIP events on its stream; use_graph = 3
 * does the same with the streams chained one after another, so that no two
 * launches overlap and each bracketed duration is the kernel's own.
 * Afterwards this returns, summed over the run: the contraction kernels'
 * device time (ms, each launch from its start to its end on its stream),
 * their launches and algorithmic FLOPs (2 x kept AOs x MOs x panels x points
 * per screening block, read back in the batch's first run), the rescue
 * launches' time, point slots (16 per tile) and flagged points, and the field
 * reductions' time and bytes moved. Measurement only. */
int kumo_hip_batch_kernel_stats(const kumo_hip_batch* batch, double* kernel_ms, int64_t* launches, double* flops,
                                double* rescue_ms, int64_t* rescue_points, double* fields_ms, int64_t* fields_bytes,
                                int64_t* rescue_flagged);
int kumo_hip_batch_release(kumo_hip_batch* batch);

/* ---- serving loop ----
 * A request server over resident molecules: `workers` threads, each with its
 * own HIP stream and workspaces (0: 4), evaluate admitted requests in
 * block-aligned ranges of about chunk_points points (0: 131072), always
 * taking the next range of the admitted request with the fewest points left,
 * so that point and line requests are not queued behind cubes and ranges of
 * different requests run on the device at the same time. A request is
 * admitted when its device memory (outputs and listed points) fits in
 * hbm_budget bytes (0: 8 GiB) beside the requests in flight, in arrival
 * order; one request alone is always admitted. Results are copied to the
 * host `outputs` (one pointer per component, as kumo_hip_evaluate_*), which
 * must stay valid until kumo_hip_server_wait returns, and equal those of a
 * lone kumo_hip_evaluate_* call. No counterpart in the reference (a library
 * call per grid there); it is for serving many clients from one GPU. */
typedef struct kumo_hip_server kumo_hip_server;
typedef struct kumo_hip_server_options {
    int32_t workers;
    int32_t reserved;
    int64_t hbm_budget;
    int64_t chunk_points;
} kumo_hip_server_options;
typedef struct kumo_hip_request_info {
    int32_t status;       /* kumo_hip_status of the request */
    int32_t backend_used;
    double queue_ms;      /* submission -> first range started */
    double run_ms;        /* first range started -> results on the host */
    double latency_ms;    /* submission -> results on the host */
    char error[256];
} kumo_hip_request_info;
int kumo_hip_server_create(kumo_hip_context* ctx, const kumo_hip_server_options* options, kumo_hip_server** server);
/* grid or points (host memory, 3 * npoints doubles; copied at submission). */
int kumo_hip_server_submit(kumo_hip_server* server, const kumo_hip_molecule* mol, const kumo_hip_grid* grid,
                           const double* points, size_t npoints, const kumo_hip_property* properties,
                           size_t nproperties, const kumo_hip_options* options, double* const* outputs,
                           uint64_t* request);
/* Blocks until the request is done, fills info and forgets the request. */
int kumo_hip_server_wait(kumo_hip_server* server, uint64_t request, kumo_hip_request_info* info);
/* done = 1 once the request's results are on the host. */
int kumo_hip_server_query(kumo_hip_server* server, uint64_t request, int32_t* done);
/* Finishes the requests in flight, then stops the workers. */
int kumo_hip_server_release(kumo_hip_server* server);

/* ---- Cube output ----
 * Field::write_cube (include/chem/kumo/density.hpp:116-121; src/cube.cpp:70-111,
 * 253-255): a Gaussian Cube file in bohr, "chem-kumo 2.1.0" as the first line,
 * `comment` (NULL: "Electron density") as the second, the grid and the atoms
 * (nuclear_charges NULL: the atomic numbers), then the npoints samples in
 * Field order (k fastest) as printf("%13.5e") fields, six per line and a line
 * break after each run along the third axis. The bytes are those of the
 * reference's writer; the samples are formatted on the GPU. values holds the
 * samples in host memory, or in device memory (HBM) when values_on_device is
 * nonzero (e.g. an evaluation's device outputs). A nonfinite sample gives
 * INVALID_ARGUMENT before the file is opened; a file that cannot be opened or
 * written gives RUNTIME_ERROR. */
int kumo_hip_write_cube(kumo_hip_context* ctx, const char* path, const char* comment, const kumo_hip_grid* grid,
                        int32_t natoms, const int32_t* atomic_numbers, const double* nuclear_charges,
                        const double* coordinates, const double* values, int32_t values_on_device);
/* Field::write_orbital_cube (include/chem/kumo/density.hpp:122-128; src/cube.cpp:257-262):
 * as kumo_hip_write_cube, with a negative atom count and the line "    1" plus
 * the one-based orbital_index (width 5) after the atoms; comment NULL:
 * "Molecular orbital amplitude". orbital_index 0 or above 9999 gives
 * INVALID_ARGUMENT. */
int kumo_hip_write_orbital_cube(kumo_hip_context* ctx, const char* path, const char* comment,
                                const kumo_hip_grid* grid, int32_t natoms, const int32_t* atomic_numbers,
                                const double* nuclear_charges, const double* coordinates, const double* values,
                                int32_t values_on_device, size_t orbital_index);
/* The GPU formatter of the Cube writer on its own: n host values written as
 * printf("%*.*e", precision + 8, precision, v), precision + 8 bytes each (no
 * terminators), 0 <= precision <= 12. exact_only nonzero decides every
 * rounding on the exact big-integer path instead of the double-double one
 * (for testing). No counterpart in the reference. */
int kumo_hip_format_scientific(kumo_hip_context* ctx, const double* values, size_t n, int32_t precision,
                               int32_t exact_only, char* text);

/* ---- molecular surfaces (include/chem/kumo/surface.hpp) ----
 * A kumo_hip_surface holds a closed triangle mesh of a total-density
 * isosurface (vertices in bohr, triangles counter-clockwise seen from
 * outside) and zero or more per-vertex fields, each with its statistics. The
 * density on the construction grid, its value and gradient at the vertices
 * during their placement, and the mapped fields are evaluated on the GPU. */
typedef struct kumo_hip_surface kumo_hip_surface;

/* SurfaceStatistics (include/chem/kumo/surface.hpp:36-74); the local extrema
 * lists are returned by kumo_hip_surface_field / kumo_hip_surface_statistics. */
typedef struct kumo_hip_surface_stats {
    double area, positive_area, negative_area, nonpolar_area;
    double minimum, maximum;
    int64_t minimum_vertex, maximum_vertex;
    double average, positive_average, negative_average;
    double positive_variance, negative_variance, total_variance;
    double balance, charge_separation, polarity_index;
    int64_t local_minima, local_maxima; /* counts */
} kumo_hip_surface_stats;

/* density_isosurface (include/chem/kumo/surface.hpp:92-106, src/surface.cpp:
 * 150-331, 341-369): no fields. */
int kumo_hip_density_isosurface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double isovalue,
                                double spacing, kumo_hip_surface** surface);
/* electrostatic_potential_surface (include/chem/kumo/surface.hpp:108-131,
 * src/surface.cpp:612-662): field 0 the potential (exact at every tolerance),
 * statistics with the polar threshold 10 kcal/mol. */
int kumo_hip_electrostatic_potential_surface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double isovalue,
                                             double spacing, double potential_tolerance,
                                             kumo_hip_surface** surface);
/* local_ionization_energy_surface (include/chem/kumo/surface.hpp:155-173,
 * src/surface.cpp:594-610): field 0 the average local ionization energy. */
int kumo_hip_local_ionization_energy_surface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double isovalue,
                                             double spacing, kumo_hip_surface** surface);
/* map_on_density_isosurface (include/chem/kumo/surface.hpp:175-195,
 * src/surface.cpp:664-699): one field per distinct one-component property,
 * statistics with polar threshold 0; options as for an evaluation. */
int kumo_hip_map_on_density_isosurface(kumo_hip_context* ctx, const kumo_hip_molecule* mol,
                                       const kumo_hip_property* properties, size_t nproperties, double isovalue,
                                       double spacing, const kumo_hip_options* options,
                                       kumo_hip_surface** surface);
/* average_local_ionization_energy at n points (include/chem/kumo/surface.hpp:
 * 133-153, src/surface.cpp:563-592). */
int kumo_hip_average_local_ionization_energy(kumo_hip_context* ctx, const kumo_hip_molecule* mol,
                                             const double* points, size_t npoints, double* out);
/* Sizes of a surface; backend_used of map_on_density_isosurface (else -1). */
int kumo_hip_surface_sizes(const kumo_hip_surface* surface, int64_t* nvertices, int64_t* ntriangles,
                           int32_t* nfields, int32_t* backend_used);
/* The mesh: 3 * nvertices coordinates, 3 * ntriangles vertex indices. */
int kumo_hip_surface_mesh(const kumo_hip_surface* surface, double* vertices, int64_t* triangles);
/* Field `field`: nvertices values, its statistics, and the local minima /
 * maxima (stats.local_minima / local_maxima entries each); NULL skips. */
int kumo_hip_surface_field(const kumo_hip_surface* surface, int32_t field, double* values,
                           kumo_hip_surface_stats* stats, int64_t* local_minima, int64_t* local_maxima);
int kumo_hip_surface_release(kumo_hip_surface* surface);
/* surface_statistics (include/chem/kumo/surface.hpp:76-84) of vertex values
 * on a caller's mesh; the extrema arrays may be NULL (call once for the
 * counts). area and volume: TriangleMesh::area / volume (may be NULL). */
int kumo_hip_surface_statistics(kumo_hip_context* ctx, const double* vertices, int64_t nvertices,
                                const int64_t* triangles, int64_t ntriangles, const double* values,
                                double polar_threshold, kumo_hip_surface_stats* stats, int64_t* local_minima,
                                int64_t* local_maxima, double* area, double* volume);
/* write_ply (include/chem/kumo/surface.hpp:197-210, src/surface.cpp:701-720):
 * ASCII PLY with per-vertex fields named names[f]. */
int kumo_hip_write_ply(kumo_hip_context* ctx, const char* path, const double* vertices, int64_t nvertices,
                       const int64_t* triangles, int64_t ntriangles, int32_t nfields, const char* const* names,
                       const double* const* values, int32_t ncomments, const char* const* comments);

/* ---- timing hooks for the benchmark ----
 * Device time (ms) of the last evaluation's dominant contraction kernel,
 * summed over its launches and measured with HIP events on the stream the
 * kernel ran on, and the number of launches and of algorithmic FLOPs
 * (2 * kept AOs * MOs * panels * points, summed over blocks). */
int kumo_hip_last_kernel_stats(const kumo_hip_context* ctx, double* kernel_ms, int64_t* launches,
                               double* algorithmic_flops, double* padded_flops);

/* Device time (ms) of the last evaluation's FP64 rescue launches, the
 * number of point slots they covered (points in 16-point tiles) and the
 * number of flagged points among them. With a
 * reduced contraction precision, the points whose density (relative error
 * bound above 5e-7) or ELF (error bound above 3e-6) the precision leaves
 * uncertain are recomputed in FP64; no counterpart in the reference (its CPU
 * path is FP64 throughout). */
int kumo_hip_last_rescue_stats(const kumo_hip_context* ctx, double* rescue_ms, int64_t* rescue_points,
                               int64_t* rescue_flagged);

/* Device time (ms) of the last evaluation's field reductions (the density
 * components and derived fields formed from the channel moments, store_block
 * of src/evaluator.cpp:127-353), the points they covered and the bytes they
 * must move (channel moments read, outputs written): measurement only, no
 * counterpart in the reference. */
int kumo_hip_last_field_stats(const kumo_hip_context* ctx, double* fields_ms, int64_t* points, int64_t* bytes);

/* The last Cube write of the context: wall time of the call (ms), device
 * time of the text kernels and of the text copies to host memory (ms, HIP
 * events on the context's stream), samples and file bytes written:
 * measurement only, no counterpart in the reference. */
int kumo_hip_last_cube_stats(const kumo_hip_context* ctx, double* total_ms, double* kernel_ms, double* copy_ms,
                             int64_t* values, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
