// Two-component (spinor) H|psi>, Src/hk_psi/reference/hk_psi_spin_c16.f90.
// psi_sp(2 mxddim, neig): row 2i is the spin-up (s = +1/2) coefficient of
// plane wave i, row 2i+1 the spin-down one.  The components are split into
// scalar bands [up 0..nb) | [down 0..nb) so the scalar column kernels apply;
// the potential acts per plane (k_plane_spin couples the two components for
// the non-collinear case); the results are merged back into the interleaved
// rows.
#pragma once

#include "fft_fast.hpp"
#include "fft_reg.hpp"

namespace cpwgpu {

constexpr int SP_THREADS = 256;

// s(:, n) = up component of band n, s(:, nb + n) = down component (ld = mtxd)
__global__ __launch_bounds__(SP_THREADS) void k_spin_split(const double2 *__restrict__ psi, long ldp, int mtxd,
                                                           int nb, double2 *__restrict__ s) {
    const int i = blockIdx.x * SP_THREADS + threadIdx.x, n = blockIdx.y;
    if (i >= mtxd || n >= nb) return;
    const double2 *c = psi + (long)n * ldp + 2 * (long)i;
    s[(long)n * mtxd + i] = c[0];
    s[(long)(nb + n) * mtxd + i] = c[1];
}

// hpsi rows 2i / 2i+1 of band n (+)= h(:, n) / h(:, nb + n); add != 0 keeps
// what hpsi holds (the non-local term)
__global__ __launch_bounds__(SP_THREADS) void k_spin_merge(const double2 *__restrict__ h, int mtxd, int nb,
                                                           double2 *__restrict__ hpsi, long ldh, int add) {
    const int i = blockIdx.x * SP_THREADS + threadIdx.x, n = blockIdx.y;
    if (i >= mtxd || n >= nb) return;
    double2 *c = hpsi + (long)n * ldh + 2 * (long)i;
    double2 u = h[(long)n * mtxd + i], d = h[(long)(nb + n) * mtxd + i];
    if (add) {
        const double2 u0 = c[0], d0 = c[1];
        u = make_double2(u.x + u0.x, u.y + u0.y);
        d = make_double2(d.x + d0.x, d.y + d0.y);
    }
    c[0] = u;
    c[1] = d;
}

// Potential components in the layout the plane kernels read (per component
// the vscr layout, ntot values each), from vscr_sp(mxdscr, nsp):
//   nsp = 1: [V1]        nsp = 2: [V1 + V2, V1 - V2]
//   nsp = 4: [V1 + V2, V1 - V2, V3, V4]   (up-up, down-down, sigma_x, sigma_y parts)
template <typename R>
__global__ __launch_bounds__(SP_THREADS) void k_spin_potential(const double *__restrict__ vsp, long ldv, int nsp,
                                                               long ntot, R *__restrict__ out) {
    const long k = (long)blockIdx.x * SP_THREADS + threadIdx.x;
    if (k >= ntot) return;
    const double v1 = vsp[k];
    if (nsp == 1) {
        out[k] = (R)v1;
        return;
    }
    const double v2 = vsp[ldv + k];
    out[k] = (R)(v1 + v2);
    out[ntot + k] = (R)(v1 - v2);
    if (nsp == 4) {
        out[2 * ntot + k] = (R)vsp[2 * ldv + k];
        out[3 * ntot + k] = (R)vsp[3 * ldv + k];
    }
}

// Plane stage of the non-collinear case: per (z, band pair) the up plane
// (colbuf band j) and the down plane (band j + npair) go through the inverse
// y and x transforms; at each r
//   up'   = (V1 + V2) up   + (V3 - i V4) down
//   down' = (V1 - V2) down + (V3 + i V4) up
// (hk_psi_spin_c16.f90:253-262); then the forward transforms.  Same two-pass
// line transforms, zero/trash columns and row handling as k_plane_fast; the
// two components share one row buffer and are taken one after the other.
template <typename R, int N1, int N2>
__global__ __launch_bounds__(FT) void k_plane_spin(cplx<R> *__restrict__ colbuf, const int *__restrict__ col_uy,
                                                   const R *__restrict__ vsp, long vstride,
                                                   const double2 *__restrict__ tw1g,
                                                   const double2 *__restrict__ tw2g, FftGeom g, int npair) {
    constexpr int A1 = split_a(N1), B1 = N1 / A1;
    constexpr int RB = FT / B1;
    constexpr int RS = row_stride<N1>();
    constexpr int NV = B1 > A1 ? B1 : A1;
    __shared__ cplx<R> tw1[N1];
    __shared__ cplx<R> tw2[N2];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = g.nx, kd1 = g.kd[0];
    const int nxp = plane_sub_stride(nx);
    cplx<R> *sub[2];
    sub[0] = reinterpret_cast<cplx<R> *>(smem_raw);  // [N2][nxp] up
    sub[1] = sub[0] + N2 * nxp;                       // [N2][nxp] down
    cplx<R> *rows = sub[1] + N2 * nxp;                // [RB][RS]
    const int n3 = g.n[2];
    const int z = blockIdx.x % n3;
    const int pair = blockIdx.x / n3;

    load_twiddles(tw1, tw1g, N1);
    load_twiddles(tw2, tw2g, N2);
    for (int t = threadIdx.x; t < 2 * N2 * nxp; t += FT) sub[0][t] = mk<R>(0, 0);
    __syncthreads();
    cplx<R> *cb[2];
    cb[0] = colbuf + ((long)pair * n3 + z) * g.ncolp;
    cb[1] = colbuf + ((long)(pair + npair) * n3 + z) * g.ncolp;
    for (int c = threadIdx.x; c < g.ncol; c += FT) {
        const int uy = col_uy[c];
        const int o = (uy >> 16) * nxp + (uy & 0xffff);
        sub[0][o] = cb[0][c];
        sub[1][o] = cb[1][c];
    }
    __syncthreads();
    line_fft<R, N2, +1>(sub[0], nx, 1, nxp, tw2);
    line_fft<R, N2, +1>(sub[1], nx, 1, nxp, tw2);

    const long zoff = (long)z * N2 * g.id;
    const int xhi = N1 - kd1, ushift = N1 - nx;
    const int r = threadIdx.x / B1, t = threadIdx.x - (threadIdx.x / B1) * B1;
    cplx<R> *row = rows + r * RS;
    int offl[B1], offs[B1];
#pragma unroll
    for (int j = 0; j < B1; j++) {
        const int x = t + A1 * j;
        const bool lo = x <= kd1, hi = x >= xhi;
        offl[j] = lo ? x : hi ? x - ushift : nx;
        offs[j] = lo ? x : hi ? x - ushift : nx + 1;
    }
    for (int y0 = 0; y0 < N2; y0 += RB) {
        const int y = y0 + r;
        const bool act = r < RB && y < N2;
        cplx<R> v[2][NV];
        // inverse x transforms of both components: pass 1 through the row buffer, pass 2 in registers
#pragma unroll
        for (int c = 0; c < 2; c++) {
            if (act && t < A1) {
                const cplx<R> *srow = sub[c] + y * nxp;
#pragma unroll
                for (int j = 0; j < B1; j++) v[c][j] = srow[offl[j]];
                dft<B1, +1, R>(v[c]);
#pragma unroll
                for (int k = 0; k < B1; k++)
                    row[rpos<N1>(t + A1 * k)] = k == 0 ? v[c][0] : cmul(v[c][k], tw1[t * k]);
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int s = 0; s < A1; s++) v[c][s] = row[rpos<N1>(s + A1 * t)];
                dft<A1, +1, R>(v[c]);
            }
            __syncthreads();
        }
        if (act) {
            // the 2x2 potential at x = t + B1 q, then pass 1 of the forward transforms
            const long o = zoff + (long)y * g.id + t;
#pragma unroll
            for (int q = 0; q < A1; q++) {
                const long k = o + B1 * q;
                const R vu = vsp[k], vd = vsp[vstride + k], vx = vsp[2 * vstride + k], vy = vsp[3 * vstride + k];
                const cplx<R> u = v[0][q], d = v[1][q];
                v[0][q] = mk<R>(vu * u.x + vx * d.x + vy * d.y, vu * u.y + vx * d.y - vy * d.x);
                v[1][q] = mk<R>(vd * d.x + vx * u.x - vy * u.y, vd * d.y + vx * u.y + vy * u.x);
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                dft<A1, -1, R>(v[c]);
#pragma unroll
                for (int m = 1; m < A1; m++) {
                    cplx<R> w1 = tw1[t * m];
                    w1.y = -w1.y;
                    v[c][m] = cmul(v[c][m], w1);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            if (act) {
#pragma unroll
                for (int m = 0; m < A1; m++) row[rpos<N1>(t + B1 * m)] = v[c][m];
            }
            __syncthreads();
            if (act && t < A1) {
                cplx<R> w[NV];
#pragma unroll
                for (int k = 0; k < B1; k++) w[k] = row[rpos<N1>(k + B1 * t)];
                dft<B1, -1, R>(w);
                cplx<R> *wrow = sub[c] + y * nxp;
#pragma unroll
                for (int p = 0; p < B1; p++) wrow[offs[p]] = w[p];
            }
            __syncthreads();
        }
    }
    line_fft<R, N2, -1>(sub[0], nx, 1, nxp, tw2);
    line_fft<R, N2, -1>(sub[1], nx, 1, nxp, tw2);
    for (int c = threadIdx.x; c < g.ncol; c += FT) {
        const int uy = col_uy[c];
        const int o = (uy >> 16) * nxp + (uy & 0xffff);
        cb[0][c] = sub[0][o];
        cb[1][c] = sub[1][o];
    }
}

// Register variant of k_plane_spin, with the line layout of k_plane_reg: a
// line of N = G*M points lives in M registers of G adjacent lanes, the
// cross-lane stages are DPP butterflies.  LDS holds the twiddles twN[N] and the
// y/x-box sub-planes of both components, sub[2][N][S] (S = nx + 2: box columns,
// a zero column and a trash column).  The y lines of the two planes are
// independent tasks; an x-row task takes row y of the up and of the down plane
// in the same lanes, so after both inverse transforms every lane holds the same
// M consecutive x points of the two components and applies the 2x2 potential
// there (vsp = [V1 + V2, V1 - V2, V3, V4], stride vstride), then both forward
// transforms.  Grid: n3 * npair blocks of T threads, one (z, band pair) each.
// CS: the sphere box lies inside N/4 (checked by the caller) and the row
// stride is the compile-time pr_sub_stride<N>() instead of the argument.
template <typename R, int N, int GX, int GY, int T, int W, bool CS = false>
__global__ __launch_bounds__(T, W) void k_plane_spin_reg(cplx<R> *__restrict__ colbuf,
                                                      const int *__restrict__ col_uy,
                                                      const R *__restrict__ vsp, long vstride,
                                                      const double2 *__restrict__ twg, FftGeom g,
                                                      int npair, int S_arg) {
    static_assert(N % GX == 0 && N % GY == 0 && T % 64 == 0, "bad plane shape");
    constexpr int MX = N / GX, MY = N / GY;
    const int S = CS ? pr_sub_stride<N>() : S_arg;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R> *twN = reinterpret_cast<cplx<R> *>(smem_raw);  // [N]
    cplx<R> *sub = twN + N;                                 // [2][N][S]: up, down
    const int n3 = g.n[2];
    const int z = blockIdx.x % n3, pair = blockIdx.x / n3;
    const int nx = g.nx, kd1 = g.kd[0], ushift = N - nx, xhi = N - kd1;
    const int tid = threadIdx.x;

    load_twiddles(twN, twg, N);
    for (int q = tid; q < 2 * N * S; q += T) sub[q] = mk<R>(0, 0);
    __syncthreads();
    cplx<R> *cb[2];
    cb[0] = colbuf + ((long)pair * n3 + z) * g.ncolp;
    cb[1] = colbuf + ((long)(pair + npair) * n3 + z) * g.ncolp;
    gather_plane<R, T>(sub, cb[0], col_uy, g.ncol, S, 0);
    gather_plane<R, T>(sub + N * S, cb[1], col_uy, g.ncol, S, 0);
    __syncthreads();

    {  // inverse y on the 2 * nx box lines
        const Lanes<R, GY> L;
        for (int task = tid / GY; task < 2 * nx; task += T / GY) {
            const int p = task >= nx ? 1 : 0, u = task - p * nx;
            cplx<R> *col = sub + p * N * S + u;
            cplx<R> v[MY];
#pragma unroll
            for (int j = 0; j < MY; j++) v[j] = col[(L.g + GY * j) * S];
            line_dif<+1, GY, MY>(v, L, twN);
#pragma unroll
            for (int k = 0; k < MY; k++) col[(k + MY * L.k2) * S] = v[k];
        }
    }
    __syncthreads();

    {  // rows: inverse x of both components, the 2x2 potential, forward x
        const Lanes<R, GX> L;
        for (int y = tid / GX; y < N; y += T / GX) {
            cplx<R> *ru = sub + y * S, *rd = sub + (N + y) * S;
            cplx<R> u[MX], d[MX];
#pragma unroll
            for (int j = 0; j < MX; j++) {
                const int x = L.g + GX * j;
                const int o = x <= kd1 ? x : x >= xhi ? x - ushift : nx;
                u[j] = ru[o];
                d[j] = rd[o];
            }
            // (with the one-instruction FP32 butterflies the odd positions of
            // both components carry negated values until the forward x: the
            // potential below is linear)
            line_dif<+1, GX, MX, R, full_mask(MX), true>(u, L, twN);
            line_dif<+1, GX, MX, R, full_mask(MX), true>(d, L, twN);
            const R *v0 = vsp + ((long)z * g.n[1] + y) * g.id + MX * L.k2;
#pragma unroll
            for (int k = 0; k < MX; k++) {
                const R vu = v0[k], vd = v0[vstride + k], vx = v0[2 * vstride + k], vy = v0[3 * vstride + k];
                const cplx<R> a = u[k], b = d[k];
                u[k] = mk<R>(vu * a.x + vx * b.x + vy * b.y, vu * a.y + vx * b.y - vy * b.x);
                d[k] = mk<R>(vd * b.x + vx * a.x - vy * a.y, vd * b.y + vx * a.y + vy * a.x);
            }
            line_dit<-1, GX, MX, R, full_mask(MX), true>(u, L, twN);
            line_dit<-1, GX, MX, R, full_mask(MX), true>(d, L, twN);
#pragma unroll
            for (int m = 0; m < MX; m++) {
                const int x = L.g + GX * m;
                const int o = x <= kd1 ? x : x >= xhi ? x - ushift : nx + 1;
                ru[o] = u[m];
                rd[o] = d[m];
            }
        }
    }
    __syncthreads();

    {  // forward y
        const Lanes<R, GY> L;
        for (int task = tid / GY; task < 2 * nx; task += T / GY) {
            const int p = task >= nx ? 1 : 0, u = task - p * nx;
            cplx<R> *col = sub + p * N * S + u;
            cplx<R> v[MY];
#pragma unroll
            for (int k = 0; k < MY; k++) v[k] = col[(k + MY * L.k2) * S];
            line_dit<-1, GY, MY>(v, L, twN);
#pragma unroll
            for (int m = 0; m < MY; m++) col[(L.g + GY * m) * S] = v[m];
        }
    }
    __syncthreads();
    scatter_plane<R, T>(cb[0], sub, col_uy, g.ncol, S);
    scatter_plane<R, T>(cb[1], sub + N * S, col_uy, g.ncol, S);
}

}  // namespace cpwgpu
