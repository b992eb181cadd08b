// Column stage with register-resident z lines.
//
// Each z column of a band lives in G adjacent lanes (M = N3/G values per
// lane, the layouts of fft_reg.hpp): no LDS exchange and no block barrier, so
// the kernels are pure streams of the sphere coefficients in and the column
// buffer out (and back).  A block of 256 threads takes 256/G columns of one
// band; blocks are dealt band-major to the XCDs (xcd_band_tile).
//
// zmap[c * N3 + z]: row of psi (0-based) of the sphere point at column c,
// height z, or -1.  With PR the sphere box lies inside |z| <= N3/4 (checked by
// the caller): the register slots whose z is outside it on every lane are
// known zeros of the inverse transform and unused outputs of the forward one.
#pragma once

#include "fft_reg.hpp"

namespace cpwgpu {

constexpr int COLR_T = 256;

// psi(G) -> colbuf[band][z][col]: gather of the sphere points of each column,
// inverse z transform in registers, store.
template <typename R, int N3, int G, bool PR>
__global__ __launch_bounds__(COLR_T) void k_col_bwd_reg(const double2 *__restrict__ psi, long ldp,
                                                      cplx<R> *__restrict__ colbuf,
                                                      const int *__restrict__ zmap,
                                                      const double2 *__restrict__ tw3g, FftGeom g, int ntiles,
                                                      int nb) {
    constexpr int M = N3 / G, CPB = COLR_T / G;
    constexpr u64mask ZB = PR ? box_slots(N3, G, M, N3 / 4) : full_mask(M);
    __shared__ cplx<R> twN[N3];
    int band, tile;
    if (!xcd_band_tile(blockIdx.x, ntiles, nb, band, tile)) return;  // whole block
    load_twiddles(twN, tw3g, N3);
    __syncthreads();
    const Lanes<R, G> L;
    const int c = tile * CPB + (int)threadIdx.x / G;
    const bool live = c < g.ncol;
    const int *zm = zmap + (long)(live ? c : g.ncol - 1) * N3;
    const double2 *pb = psi + (long)band * ldp;
    int idx[M];
#pragma unroll
    for (int j = 0; j < M; j++)
        if (has_bit(ZB, j)) idx[j] = zm[L.g + G * j];
    cplx<R> v[M];
#pragma unroll
    for (int j = 0; j < M; j++)
        if (has_bit(ZB, j)) {
            const double2 p = idx[j] >= 0 ? pb[idx[j]] : make_double2(0.0, 0.0);
            v[j] = mk<R>((R)p.x, (R)p.y);
        }
    line_dif<+1, G, M, R, ZB>(v, L, twN);
    if (live) {
        // v[k] = column value at z = k + M * bitrev(g)
        cplx<R> *ob = colbuf + (long)band * N3 * g.ncolp + c + (long)M * L.k2 * g.ncolp;
#pragma unroll
        for (int k = 0; k < M; k++) ob[(long)k * g.ncolp] = v[k];
    }
}

// colbuf -> forward z transform in registers -> hpsi = local/N (+ ekpg psi)
// (+ hpsi) at the sphere points of each column.
template <typename R, int N3, int G, bool PR>
__global__ __launch_bounds__(COLR_T) void k_col_fwd_reg(const double2 *__restrict__ psi, long ldp,
                                                      double2 *__restrict__ hpsi, long ldh,
                                                      const cplx<R> *__restrict__ colbuf,
                                                      const int *__restrict__ zmap,
                                                      const double *__restrict__ ekpg,
                                                      const double2 *__restrict__ tw3g, FftGeom g, int ntiles,
                                                      int nb, double inv_n, int mode) {
    constexpr int M = N3 / G, CPB = COLR_T / G;
    constexpr u64mask ZB = PR ? box_slots(N3, G, M, N3 / 4) : full_mask(M);
    __shared__ cplx<R> twN[N3];
    int band, tile;
    if (!xcd_band_tile(blockIdx.x, ntiles, nb, band, tile)) return;  // whole block
    load_twiddles(twN, tw3g, N3);
    const Lanes<R, G> L;
    const int c = tile * CPB + (int)threadIdx.x / G;
    const bool live = c < g.ncol;
    const int cc = live ? c : g.ncol - 1;
    const int *zm = zmap + (long)cc * N3;
    const double2 *pb = psi + (long)band * ldp;
    double2 *hb = hpsi + (long)band * ldh;
    const bool kin = mode & COLF_KIN, acc = mode & COLF_ACC;
    // the column (z = k + M bitrev(g) in register k), and, before the
    // transform so that the round trips overlap it, the sphere rows of the
    // outputs and what is added to local/N there
    cplx<R> v[M];
    const cplx<R> *ib = colbuf + (long)band * N3 * g.ncolp + cc + (long)M * L.k2 * g.ncolp;
#pragma unroll
    for (int k = 0; k < M; k++) v[k] = ib[(long)k * g.ncolp];
    int idx[M];
    double2 add[M];
#pragma unroll
    for (int m = 0; m < M; m++)
        if (has_bit(ZB, m)) {
            idx[m] = live ? zm[L.g + G * m] : -1;
            add[m] = make_double2(0.0, 0.0);
            if (idx[m] >= 0) {
                if (acc) add[m] = hb[idx[m]];
                if (kin) {
                    const double2 p = pb[idx[m]];
                    const double ek = ekpg[idx[m]];
                    add[m].x += ek * p.x;
                    add[m].y += ek * p.y;
                }
            }
        }
    __syncthreads();  // twiddles
    line_dit<-1, G, M, R, ZB>(v, L, twN);
#pragma unroll
    for (int m = 0; m < M; m++)
        if (has_bit(ZB, m) && idx[m] >= 0)
            hb[idx[m]] = make_double2((double)v[m].x * inv_n + add[m].x, (double)v[m].y * inv_n + add[m].y);
}

}  // namespace cpwgpu
