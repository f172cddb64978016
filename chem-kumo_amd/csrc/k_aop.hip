// AO-P contraction (src/density.cpp:149-229 density_derivatives_shared,
// :365-444 densities_blocked): per screening block, the kept AO panels Phi
// are staged in HBM, W = P_kept * [Phi Phi_x Phi_y Phi_z] is formed on the
// FP64 matrix cores (v_mfma_f64_16x16x4) from LDS tiles of P, gathered by
// the block's kept list, and of Phi, and the per-point moments are the dot
// products of Phi and W over the kept rows.
#include "kernels.hpp"

#include "common.hpp"

namespace kumo {
namespace {

// Phi[blk][kept row i][panel][point] for the block's kept AOs (FP64).
template <int ORDER>
__global__ void __launch_bounds__(128) aop_phi_kernel(AoPLaunch a, int panels) {
    constexpr unsigned PSET = ORDER == 0 ? kPanelsValue : (ORDER == 1 ? kPanelsGrad : kPanelsAll);
    constexpr int NP = popcount10(PSET);
    const int blk_local = blockIdx.y;
    const long long blk = a.first_block + blk_local;
    const long long bstart = blk * a.block_size;
    const int bcount = static_cast<int>(min(static_cast<long long>(a.block_size), a.npoints - bstart));
    const int j = blockIdx.x * 128 + threadIdx.x;
    if (j >= bcount) return;
    const MolView& m = a.mol;
    const int nao = m.nao;
    const int B = a.block_size;
    const int nsh = m.nshells;
    const int32_t* ks = a.tables.kept_shell + static_cast<long long>(blk_local) * nsh;
    const uint32_t* kb = a.tables.kept_bits + static_cast<long long>(blk_local) * nsh;
    const int nkshell = a.tables.nkshell[blk_local];
    double* phi = a.phi + static_cast<long long>(blk_local) * nao * panels * B;
    double px, py, pz;
    point_coords(a.ps, bstart + j, px, py, pz);
    int row = 0; // kept row counter (ascending AO order)
    for (int q = 0; q < nkshell; ++q) {
        const DevShell sh = m.shells[ks[q]];
        const uint32_t bits = kb[q];
        const double x = px - sh.o[0], y = py - sh.o[1], z = pz - sh.o[2];
        const double r2 = sq_norm(x, y, z);
        // kept-row index of each AO row of the block
        int rowidx[kMaxCart];
        int rr = row;
#pragma unroll
        for (int r = 0; r < kMaxCart; ++r) {
            rowidx[r] = ((bits >> r) & 1u) ? rr : -1;
            rr += (bits >> r) & 1u;
        }
        if (r2 <= sh.cutoff_r2) {
            eval_block<ORDER, PSET>(sh, m, x, y, z, r2, [&](int r, const double (&v)[NP]) __attribute__((always_inline)) {
                if (rowidx[r] >= 0)
#pragma unroll
                    for (int i = 0; i < NP; ++i)
                        phi[(static_cast<long long>(rowidx[r]) * panels + i) * B + j] = v[i];
            });
        } else {
#pragma unroll
            for (int r = 0; r < kMaxCart; ++r)
                if (rowidx[r] >= 0)
                    for (int i = 0; i < NP; ++i)
                        phi[(static_cast<long long>(rowidx[r]) * panels + i) * B + j] = 0.0;
        }
        row = rr;
    }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// W[blk][i][c][j] = sum_k P(kept_i, kept_k) Phi[k][c][j], c < wpanels, on
// FP64 MFMA. A workgroup computes a 64 x 64 tile of W (rows i, columns
// (c, j) = (col / bcount, col % bcount)); wave w its rows 16 w .. 16 w + 15
// against the four 16-column subtiles. K runs over the kept rows in chunks
// of 16 staged in LDS: P gathered by the kept list (zero past m), Phi rows
// read along the points. Operand layouts of v_mfma_f64_16x16x4: lane l
// holds A[l % 16][l / 16] and B[l / 16][l % 16], and accumulator element r
// is D[l / 16 + 4 r][l % 16].
constexpr int WT = 64, WK = 16;
__global__ void __launch_bounds__(256) aop_w_kernel(AoPLaunch a, int panels, int wpanels) {
    __shared__ double ps_[WT][WK + 1];
    __shared__ double fs_[WK][WT];
    const int blk_local = blockIdx.z;
    const long long blk = a.first_block + blk_local;
    const long long bstart = blk * a.block_size;
    const int bcount = static_cast<int>(min(static_cast<long long>(a.block_size), a.npoints - bstart));
    const int m = a.tables.nkept[blk_local];
    const int i0 = blockIdx.y * WT;
    const int cols = wpanels * bcount;
    const int col0 = blockIdx.x * WT;
    if (i0 >= m || col0 >= cols) return; // uniform over the workgroup
    const int nao = a.mol.nao, B = a.block_size;
    const int32_t* kl = a.tables.kept_list + static_cast<long long>(blk_local) * nao;
    const double* phi = a.phi + static_cast<long long>(blk_local) * nao * panels * B;
    double* w = a.w + static_cast<long long>(blk_local) * nao * 4 * B;
    const double* P = a.p[0];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    // the staging items of this thread: 4 of the P tile, 4 of the Phi tile
    int prow[4], pcol[4], fk[4], fcol[4];
    long long fofs[4];
    bool fok[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + 256 * u;
        prow[u] = e / WK;
        pcol[u] = e % WK;
        fk[u] = e / WT;
        fcol[u] = e % WT;
        const int col = col0 + fcol[u];
        fok[u] = col < cols;
        const int c = fok[u] ? col / bcount : 0, j = fok[u] ? col % bcount : 0;
        fofs[u] = static_cast<long long>(c) * B + j;
    }
    const int pi[4] = {i0 + prow[0] < m ? kl[i0 + prow[0]] : -1, i0 + prow[1] < m ? kl[i0 + prow[1]] : -1,
                       i0 + prow[2] < m ? kl[i0 + prow[2]] : -1, i0 + prow[3] < m ? kl[i0 + prow[3]] : -1};
    f64x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
        acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < m; k0 += WK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + pcol[u];
            ps_[prow[u]][pcol[u]] =
                (pi[u] >= 0 && k < m) ? P[static_cast<long long>(pi[u]) * nao + kl[k]] : 0.0;
            const int kf = k0 + fk[u];
            fs_[fk[u]][fcol[u]] =
                (fok[u] && kf < m) ? phi[static_cast<long long>(kf) * panels * B + fofs[u]] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < WK; ks += 4) {
            const double av = ps_[16 * wave + r16][ks + g];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, fs_[ks + g][16 * t + r16], acc[t], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int col = col0 + 16 * t + r16;
        if (col >= cols) continue;
        const int c = col / bcount, j = col % bcount;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = i0 + 16 * wave + g + 4 * r;
            if (i < m) w[(static_cast<long long>(i) * 4 + c) * B + j] = acc[t][r];
        }
    }
}

// Moments of one matrix (contract_derivatives, src/density.cpp:26-44).
__global__ void __launch_bounds__(128) aop_moments_kernel(AoPLaunch a, int panels, int order, int channel) {
    const int blk_local = blockIdx.y;
    const long long blk = a.first_block + blk_local;
    const long long bstart = blk * a.block_size;
    const int bcount = static_cast<int>(min(static_cast<long long>(a.block_size), a.npoints - bstart));
    const int j = blockIdx.x * 128 + threadIdx.x;
    if (j >= bcount) return;
    const int nao = a.mol.nao, B = a.block_size;
    const int m = a.tables.nkept[blk_local];
    const double* phi = a.phi + static_cast<long long>(blk_local) * nao * panels * B;
    const double* w = a.w + static_cast<long long>(blk_local) * nao * 4 * B;
    double q[kMoments];
#pragma unroll
    for (int k = 0; k < kMoments; ++k)
        q[k] = 0.0;
    constexpr int hc[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int i = 0; i < m; ++i) {
        const double* ph = phi + static_cast<long long>(i) * panels * B + j;
        const double* wi = w + static_cast<long long>(i) * 4 * B + j;
        const double y = wi[0];
        q[0] += ph[0] * y;
        if (order >= 1) {
            for (int ax = 0; ax < 3; ++ax) {
                const double ga = ph[(1 + ax) * B];
                q[1 + ax] += ga * y;
                q[4 + ax] += ga * wi[(1 + ax) * B];
            }
            if (order == 2) {
                int h = 0;
                for (int ax = 0; ax < 3; ++ax)
                    for (int c2 = ax; c2 < 3; ++c2, ++h)
                        q[7 + h] += ph[(4 + hc[ax][c2]) * B] * y + ph[(1 + ax) * B] * wi[(1 + c2) * B];
            }
        }
    }
    const long long point = bstart + j - a.batch_first;
    double* dst = a.moments + static_cast<long long>(channel) * kMoments * a.batch_points;
    for (int k = 0; k < kMoments; ++k) {
        const double f = k == 0 ? 1.0 : (k <= 3 ? 2.0 : (k <= 6 ? 0.5 : 2.0));
        dst[k * a.batch_points + point] = f * q[k];
    }
}

} // namespace

void launch_ao_p(const AoPLaunch& a, hipStream_t stream) {
    if (a.nblk <= 0) return;
    const int panels = a.order == 0 ? 1 : (a.order == 1 ? 4 : 10);
    const int wpanels = a.order == 0 ? 1 : 4;
    const dim3 gphi(static_cast<unsigned>((a.block_size + 127) / 128), static_cast<unsigned>(a.nblk));
    switch (a.order) {
    case 0: aop_phi_kernel<0><<<gphi, 128, 0, stream>>>(a, panels); break;
    case 1: aop_phi_kernel<1><<<gphi, 128, 0, stream>>>(a, panels); break;
    default: aop_phi_kernel<2><<<gphi, 128, 0, stream>>>(a, panels); break;
    }
    KUMO_HIP(hipGetLastError());
    const dim3 gw(static_cast<unsigned>((wpanels * a.block_size + WT - 1) / WT),
                  static_cast<unsigned>((a.mol.nao + WT - 1) / WT), static_cast<unsigned>(a.nblk));
    for (int ch = 0; ch < 2; ++ch) {
        if (!a.p[ch]) continue;
        // a.p[ch] is the matrix of channel ch (the caller forms combinations).
        AoPLaunch b = a;
        b.p[0] = a.p[ch];
        aop_w_kernel<<<gw, 256, 0, stream>>>(b, panels, wpanels);
        KUMO_HIP(hipGetLastError());
        aop_moments_kernel<<<gphi, 128, 0, stream>>>(a, panels, a.order, ch);
        KUMO_HIP(hipGetLastError());
    }
}

} // namespace kumo
