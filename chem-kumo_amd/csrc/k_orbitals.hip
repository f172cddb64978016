// MO amplitudes and their derivatives (src/orbitals.cpp:66-102 on grids,
// src/evaluator.cpp:688-730 at points): psi = C_sel^T phi over all AOs, no
// screening. One thread per point evaluates every AO block in FP64 and
// accumulates up to kOrbChunk selected MOs in registers.
#include "kernels.hpp"

#include "common.hpp"

namespace kumo {
namespace {

constexpr int kOrbChunk = 4;

template <int ORDER>
__global__ void __launch_bounds__(128) orbital_kernel(MolView m, PointSource ps, long long first, long long n,
                                                      const double* coeff, int ldc, const int* mo_index, int nmo,
                                                      double* out, long long ld_out) {
    constexpr unsigned PSET = ORDER == 0 ? kPanelsValue : (ORDER == 1 ? kPanelsGrad : kPanelsAll);
    constexpr int NP = popcount10(PSET);
    const long long j = blockIdx.x * 128ll + threadIdx.x;
    if (j >= n) return;
    double px, py, pz;
    point_coords(ps, first + j, px, py, pz);
    int idx[kOrbChunk];
#pragma unroll
    for (int k = 0; k < kOrbChunk; ++k)
        idx[k] = k < nmo ? mo_index[k] : mo_index[0];
    double acc[kOrbChunk][NP];
#pragma unroll
    for (int k = 0; k < kOrbChunk; ++k)
#pragma unroll
        for (int i = 0; i < NP; ++i)
            acc[k][i] = 0.0;
    for (int s = 0; s < m.nshells; ++s) {
        const DevShell sh = m.shells[s];
        const double x = px - sh.o[0], y = py - sh.o[1], z = pz - sh.o[2];
        const double r2 = sq_norm(x, y, z);
        if (!(r2 <= sh.cutoff_r2)) continue;
        eval_block<ORDER, PSET>(sh, m, x, y, z, r2, [&](int row, const double (&v)[NP]) __attribute__((always_inline)) {
            const double* c = coeff + static_cast<long long>(sh.ao + row) * ldc;
#pragma unroll
            for (int k = 0; k < kOrbChunk; ++k) {
                const double ck = c[idx[k]];
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    acc[k][i] += ck * v[i];
            }
        });
    }
#pragma unroll
    for (int k = 0; k < kOrbChunk; ++k)
        if (k < nmo)
#pragma unroll
            for (int i = 0; i < NP; ++i)
                out[(static_cast<long long>(k) * NP + i) * ld_out + j] = acc[k][i];
}

} // namespace

void launch_orbitals(const MolView& m, const PointSource& ps, long long first, long long n, const double* coeff,
                     int ldc, const int* mo_index, int nmo, int order, double* out, long long ld_out,
                     hipStream_t stream) {
    if (n <= 0 || nmo <= 0) return;
    const int NP = order == 0 ? 1 : (order == 1 ? 4 : 10);
    const dim3 grid(static_cast<unsigned>((n + 127) / 128));
    for (int k0 = 0; k0 < nmo; k0 += kOrbChunk) {
        const int cnt = nmo - k0 < kOrbChunk ? nmo - k0 : kOrbChunk;
        double* o = out + static_cast<long long>(k0) * NP * ld_out;
        switch (order) {
        case 0: orbital_kernel<0><<<grid, 128, 0, stream>>>(m, ps, first, n, coeff, ldc, mo_index + k0, cnt, o, ld_out); break;
        case 1: orbital_kernel<1><<<grid, 128, 0, stream>>>(m, ps, first, n, coeff, ldc, mo_index + k0, cnt, o, ld_out); break;
        default: orbital_kernel<2><<<grid, 128, 0, stream>>>(m, ps, first, n, coeff, ldc, mo_index + k0, cnt, o, ld_out); break;
        }
        KUMO_HIP(hipGetLastError());
    }
}

} // namespace kumo
