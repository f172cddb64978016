// Host-callable launchers of the HIP kernels.
#pragma once

#include "ao_eval.hpp"
#include "device.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace kumo {

// Rows of one contraction chunk: kept AO blocks are packed into chunks of at
// most kChunkRows operand rows.
constexpr int kChunkRows = 32;
constexpr int kBf16Planes = 3; // bf16 terms of the MO operand table

// How a kept AO block enters the AO->MO contraction.
// Dropped AOs of a partially kept block get the operand row -1 (zero
// coefficients) in the chunk's row list.
enum EntryMode : int32_t {
    kEntryCart = 0,       // every AO kept: the block's Cartesian rows (transform folded into C)
    kEntryCartMasked = 1, // Cartesian block, some AOs dropped
    kEntryPure = 2,       // pure block, some AOs dropped: its Cartesian rows with the kept pure
                          // AOs' coefficients folded when the chunk is staged
};
// Operand row code of a folded row: kFoldedRow - (entry << 5 | Cartesian
// component) (codes below -1; -1 is a zero row), formed where the chunk is
// staged; or, for the rows the fold workspace holds, kFoldedWs - (workspace
// row), formed once per batch by launch_fold (codes <= kFoldedWs).
constexpr int32_t kFoldedRow = -2;
constexpr int32_t kFoldedWs = -(1 << 29);

// Per screening block outputs of the mask kernel.
struct KeptTables {
    uint8_t* kept;          // [nblk][nao] significant_aos flags
    int32_t* kept_list;     // [nblk][nao] kept AO indices, ascending
    int32_t* nkept;         // [nblk]
    int32_t* kept_shell;    // [nblk][nshells] AO blocks with at least one kept AO, ascending
    uint32_t* kept_bits;    // [nblk][nshells] bit r: AO row r of that block is kept
    int32_t* nkshell;       // [nblk]
    // AO->MO contraction plan: entries ordered by descending l (so that the
    // lanes generating one chunk mostly share the block's code path), packed
    // next-fit into chunks.
    int32_t* ent;           // [nblk][nshells] block index | mode << 24
    uint32_t* ent_bits;     // [nblk][nshells] kept bits of the entry's block
    uint8_t* ent_row;       // [nblk][nshells] first row of the entry in its chunk
    int32_t* chunk_begin;   // [nblk][max_chunks + 1] first entry of each chunk
    // [nblk][max_chunks][kChunkRows][2]: operand table row (-1: zero
    // coefficients; below -1: a folded row, kFoldedRow) and the Phi row it
    // multiplies, block << 15 | mode << 13 | x, y, z powers of the Cartesian
    // component (3 bits each), or -1 past the chunk's rows.
    int32_t* chunk_rows;
    int32_t* nchunks;       // [nblk]
    int max_chunks;
    // Fold workspace: up to fold_per_block folded rows per block (0: none;
    // the rows beyond it are formed at staging), row b * fold_per_block + j
    // of block b: its FP64 coefficients over every MO slot, the upper bounds
    // of its norms, its entry-relative code, and the rows per block.
    double* fold_rows;      // [fold_cap][fold_slots]
    void* fold_lo;          // the same rows as staged in FP32 ([fold_cap][fold_slots]) or as
                            // kBf16Planes bf16 planes ([plane][fold_cap][fold_slots])
    float2* fold_norm;      // [fold_cap]
    int32_t* fold_desc;     // [fold_cap] kFoldedRow - (entry << 5 | component)
    int32_t* fold_n;        // [nblk]
    int fold_per_block;
    int fold_slots;
    int fold_cap;           // rows allocated: blocks of the tables x fold_per_block
    // 1: also write kept, kept_list, kept_shell, kept_bits and nkshell (the
    // AO-P path and the masks entry point); 0: the AO->MO plan and nkept only.
    int lists;
};

// Upper bound on the chunks of any screening block of a basis: next-fit
// packing of blocks sorted by l closes a chunk of l-blocks only when it holds
// at least floor(kChunkRows / ncart(l)) of them.
int max_chunks_bound(const DevShell* shells, int nshells);

// significant_aos (src/ao_screening.hpp:35-50) for screening blocks
// [first_block, first_block + nblk) of `block_size` points; order selects
// the AO panels screened (0: values, 1: + gradient, 2: + Hessian).
void launch_masks(const MolView& m, const PointSource& ps, long long npoints, int block_size, long long first_block,
                  int nblk, int order, double threshold, const double* screen_r2, const KeptTables& t,
                  hipStream_t stream);

// AO panels at listed points, column-major nao x (panels * n).
void launch_ao_eval(const MolView& m, const PointSource& ps, long long n, int order, double* out, hipStream_t stream);

// MO operand table of the AO->MO contraction ([plane][row][slot]; rows as
// DevShell::crow / prow; slots alpha then beta, each padded to 16).
struct MoOperand {
    const double* op64;     // FP64, one plane
    const float* op32;      // FP32, one plane
    const uint16_t* opbf;   // kBf16Planes bf16 planes, largest term first
    const float2* norm;     // per row: Euclidean norm over the alpha slots, over the beta slots
    int rows;               // rows per plane
    int slots;
    int alpha_slots;        // slots of the alpha channel (multiple of 16)
};

// Fused AO-MO contraction (density_derivatives_mo, src/density.cpp:234-351):
// for every point of the batch, the moments of each channel, written to
// moments[channel][kMoments][npoints_batch] (FP64). slot range [slot0,
// slot1) selects the channels contracted.
// Points of a batch recomputed in FP64 (ELF rescue), in tiles of points of
// one screening block.
struct RescuePlan {
    uint8_t* flag;       // [batch points]
    uint16_t* offs;      // [batch points]: flagged offsets per block, compacted
    int32_t* count;      // [blocks]
    int32_t* tile_blk;   // [max tiles]
    int32_t* tile_sub;   // [max tiles]
    int32_t* total;      // [2]: tiles in use, flagged points
};
// What a point is rescued for (output bits: 0 total, 1 alpha, 2 beta).
//  ELF: the channel moments carry a relative error elf_eps; rescue where the
//   ELF could then be off by more than elf_tol.
//  Density: the reduced-precision contraction of channel c is off by about
//   rho_eps * sum_i |Psi_i| sum_k |C_ik Phi_k| <= rho_eps * sqrt(rho_c) s_c
//   (Cauchy-Schwarz over the MOs), with s_c = sum_k |Phi_k| ||C_k,c|| from
//   the contraction kernel; rescue where 2 rho_eps s_c / sqrt(rho_c) (or its
//   rho-weighted mean over both channels for the total) exceeds rho_tol.
struct RescueCriteria {
    int elf_which;
    double elf_eps, elf_tol;
    int rho_which;
    const float* rho_scale; // [2][n] s_alpha, s_beta
    double rho_eps, rho_tol;
    // When not null, the ELF test bounds each channel's errors from its own
    // scales: rho_scale ([2][n]) and grad_scale ([2][n], the gradient
    // scales), with rounding unit elf_scale_eps, instead of elf_eps (tau +
    // |grad rho|^2 / 8 rho).
    const float* grad_scale;
    double elf_scale_eps;
};
// Flags the points to recompute and groups them into tiles of `tile` points
// of one screening block.
void launch_rescue_plan(const double* channel_moments, long long n, int closed, const RescueCriteria& c,
                        int block_size, int nblk, int tile, const RescuePlan& r, hipStream_t stream);
void launch_rescue_scatter(const double* rescue_moments, long long rslots, int groups, int tile, int block_size,
                           const RescuePlan& r, double* channel_moments, long long n, hipStream_t stream);

// Per-point orbital sums of the orbital-energy fields (OrbitalEnergySums,
// src/orbital_energy_fields.hpp:16-60), each sum_i w_s(i) |psi_i|^2 over MO
// slots i with the weights:
enum OeSum : int {
    kOeOccDensity = 0, // n_i (occupied)
    kOeOccEnergy,      // n_i |e_i| (occupied)
    kOeVirDensity,     // 1 (virtual)
    kOeVirEnergy,      // e_i (virtual)
    kOeAttachment,     // w e_i (virtual, e_i < 0; w = 2 restricted, 1 unrestricted)
    kOeFukuiPlus,      // g_i (virtual, restricted closed shell)
    kOeFukuiMinus,     // g_i (occupied, restricted closed shell)
    kOeSums
};

struct AoMoLaunch {
    MolView mol;
    PointSource ps;
    MoOperand mo;
    KeptTables tables;
    long long npoints;       // all points of the evaluation
    long long batch_first;   // first point of the batch (a block boundary)
    long long batch_points;  // points in the batch
    int block_size;
    long long first_block;   // screening block of batch_first
    int nblk;
    int order;               // AO arithmetic: 0 value kernel, 1/2 derivative kernel
    int hessian;             // 0 none, 1 diagonal, 2 full, 3 trace only (in the xx moment)
    int precision;           // kumo_hip_precision
    int slot0, slot1;
    int tile_points;         // points per workgroup: 16, 32 or 64 (the wave-uniform kernel); 0: 32; FP64 always 16
    double* moments;         // [2][kMoments][batch_points]
    // Reduced precisions, when not null: [2][batch_points] sum_k |Phi_k| ||C_k,c||
    // over the kept operand rows for channel c (slot group 0 only), then, when
    // the panels include the gradient, [2][batch_points]
    // sum_k (|d/dx Phi_k| + |d/dy Phi_k| + |d/dz Phi_k|) ||C_k,c||.
    float* rho_scale;
    // Weighted mode (FP64, order 0), when not null: [kOeSums][mo.slots]
    // weight of each MO slot in each orbital sum; the moments array then
    // receives [slot group][kOeSums][batch_points].
    const double* weights;
    // Rescue mode (FP64 only): workgroup t evaluates the flagged points of
    // tile t of the plan and writes moments at slot t * 16 + j.
    const int32_t* r_tile_blk;
    const int32_t* r_tile_sub;
    const int32_t* r_total;
    const int32_t* r_count;
    const uint16_t* r_offs;
};
void launch_ao_mo(const AoMoLaunch& a, hipStream_t stream);
constexpr int kRescueTile = 16; // points per rescue tile (the FP64 kernel's tile)
// The wave-uniform contraction (tile_points 64: 64 points per workgroup, AO
// blocks generated one per wave at every point of the tile) for the
// split-bf16 precisions, every panel set but the full Hessian, outside
// rescue and weighted modes; launch_ao_mo dispatches to it.
bool ao_mo_wave_supported(const AoMoLaunch& a);
void launch_ao_mo_wave(const AoMoLaunch& a, hipStream_t stream);
// The folded rows the masks of nblk blocks placed in the fold workspace
// (t.fold_per_block > 0), from the FP64 operand table mo: in FP64 and as
// staged in `precision` (FP32 or bf16 planes; nothing more for FP64).
void launch_fold(const MolView& m, const MoOperand& mo, const KeptTables& t, int nblk, int precision,
                 hipStream_t stream);

// AO-P contraction (density_derivatives_shared / densities_blocked,
// src/density.cpp:149-229, 365-444): W = P_kept [Phi Phi_x Phi_y Phi_z] and
// the per-point moments of each matrix (0: total / alpha, 1: spin / beta).
struct AoPLaunch {
    MolView mol;
    PointSource ps;
    KeptTables tables;
    const double* p[2];      // row-major nao x nao, null if unused
    long long npoints, batch_first, batch_points;
    int block_size;
    long long first_block;
    int nblk;
    int order;               // 0 (densities_blocked) or 1, 2
    double* phi;             // workspace: [nblk][nao][panels][block_size]
    double* w;               // workspace: [nblk][nao][4][block_size]
    double* moments;         // [2][kMoments][batch_points]
};
void launch_ao_p(const AoPLaunch& a, hipStream_t stream);

// Per-point derived fields (store_block, src/evaluator.cpp:248-353).
struct FieldTask {
    int kind, component;
    double* out[6];
};
// The density components (total, alpha, beta, difference; combine,
// src/density.cpp:50-60) are formed per point from the two channels' moments
// and the derived fields computed from them in the same pass.
struct FieldsLaunch {
    const double* channels;  // [2][kMoments][n]: alpha, beta (from_spin) or total, spin difference
    int from_spin;           // 1: the channels are alpha and beta
    int closed;              // 1: closed shell (alpha and beta channels equal; the second is not read)
    long long n;
    long long out_offset;    // index of the batch's first point in the outputs
    int ntasks;
    const FieldTask* tasks;  // device array
    double fermi;            // 0.3 (6 pi^2)^(2/3), as src/evaluator.cpp:128-129 computes it
    double rdg_scale;        // 2 (3 pi^2)^(1/3), src/evaluator.cpp:148
};
FieldsLaunch fields_launch(const double* channels, int from_spin, int closed, long long n, long long out_offset,
                          int ntasks, const FieldTask* tasks);
void launch_fields(const FieldsLaunch& f, hipStream_t stream);

// The orbital-energy fields (store_orbital_energy_block, src/evaluator.cpp:
// 481-520) of n points from their sums [kOeSums][n]: ALIE, LEA, LEAE, local
// electronegativity and hardness (0 where a denominator vanishes), the
// orbital-weighted Fukui functions f+, f-, f0 and the dual descriptor.
// virtuals / fukui: the virtual-orbital / Fukui sums were formed.
void launch_orbital_energy_fields(const double* sums, long long n, long long out_offset, int ntasks,
                                  const FieldTask* tasks, int virtuals, int fukui, hipStream_t stream);

// dst[i] += src[i]; dst[i] = value.
void launch_add_into(double* dst, const double* src, long long n, hipStream_t stream);
void launch_fill(double* dst, long long n, double value, hipStream_t stream);

// Cube sample text (Field::write_cube, src/cube.cpp:104-109): values
// [first, first + count) of a field whose third grid axis has nz points,
// written as their printf("%13.5e") fields with the line breaks of the file
// to text, which receives the bytes from offset cube_text_offset(first, nz)
// of the sample section on.
long long cube_text_offset(long long i, long long nz);
void launch_cube_text(const double* values, long long first, long long count, long long nz, long long text_first,
                      unsigned char* text, hipStream_t stream);
// printf("%*.*e", P + 8, P, v) in P + 8 bytes per value (0 <= P <= 12);
// exact_only: every rounding decided by exact big-integer comparison.
void launch_sci_slots(const double* values, long long n, int precision, int exact_only, char* out,
                      hipStream_t stream);
// *bad = number of infinite or NaN values.
void launch_count_nonfinite(const double* values, long long n, int* bad, hipStream_t stream);

// MO amplitudes (and derivatives) of selected MOs at points:
// out[mo][component][n] for components value (order 0) or value, gradient,
// Hessian (xx, xy, xz, yy, yz, zz).
void launch_orbitals(const MolView& m, const PointSource& ps, long long first, long long n, const double* coeff,
                     int ldc, const int* mo_index, int nmo, int order, double* out, long long ld_out,
                     hipStream_t stream);

// Electrostatic potential at points (src/electrostatic_potential.cpp:231-361).
// Boys table on the device: F_m(i * 0.05) for m < kBoysOrders, i < kBoysPoints
// (row i at i * kBoysOrders), then per pair order L the orders L .. L + 6 of
// each point in rows of kBoysRowL doubles.
constexpr int kBoysOrders = 39, kBoysPoints = 801, kBoysRowL = 8;
struct EspPairDev {
    double p, qx, qy, qz;
    int32_t L;
    int32_t first;
};
struct EspAtomDev {
    double x, y, z, charge;
};
// The pairs [first, last) of one L = la + lb, evaluated in `slices` contiguous
// slices (each writes a partial sum per point).
struct EspGroup {
    int32_t L, first, last, slices;
};
// Sets the slices of every group for n points; returns their total (the rows
// of the partial-sum workspace, n doubles each).
int esp_slices(std::vector<EspGroup>& groups, long long n);
// out[j] = V at point first + j, j < n; partial holds esp_slices(groups, n) x n doubles.
void launch_esp(const EspPairDev* pairs, const std::vector<EspGroup>& groups, const double* coef,
                const EspAtomDev* atoms, int natoms, const double* boys_table, const PointSource& ps, long long first,
                long long n, double* partial, double* out, hipStream_t stream);
void launch_boys(const double* table, const double* T, long long n, int m, double* out, hipStream_t stream);
// Copies n <= kMaxStoredTasks field tasks to dst through kernel arguments.
constexpr int kMaxStoredTasks = 16;
void launch_store_tasks(const FieldTask* tasks, int n, FieldTask* dst, hipStream_t stream);
// The ESP pair loop's Boys function (orders 0..top, top <= 10) at n arguments.
void launch_boys_esp(const double* table, const double* T, long long n, int top, double* out, hipStream_t stream);

} // namespace kumo
