// Specialised local-potential kernels for mesh lengths with a two-pass split
// N = A*B (split_a in fft_core.hpp).  Same products as the generic kernels in
// fft_local.hpp; every 1-D transform is two register passes around one LDS
// exchange, and the x-row work fuses inverse FFT -> V(r) -> forward FFT with
// the row kept in registers between the two transforms.
#pragma once

#include "fft_core.hpp"
#include "fft_local.hpp"

namespace cpwgpu {

constexpr int FT = 256;  // threads per block of the specialised kernels

// Block -> (band, column tile) so that the tiles of one band run on one XCD
// (workgroups are dealt to the 8 XCDs round robin): psi of a band is read
// through a single L2.  Returns false for the padding blocks.
__device__ __forceinline__ bool xcd_band_tile(int b, int ntiles, int nb, int &band, int &tile) {
    const int x = b & 7, i = b >> 3;
    band = x + 8 * (i / ntiles);
    tile = i - (i / ntiles) * ntiles;
    return band < nb;
}

// Two-pass transform of nl lines in LDS: line l at buf + l*ldist, element j at
// + j*estr.  Lines go in rounds of whole lines (all B work items of a line in
// the same round), LPW = 64/B lines per wave: in wave w, lane t*LPW + i is work
// item t of line w*LPW + i, so neighbouring lanes touch the same position of
// neighbouring lines, and a round that is not full occupies the first waves
// only (the others skip it).  The result is stored back in natural order.
template <int N> constexpr int lines_per_wave() { return 64 / (N / split_a(N)); }

template <typename R, int N, int DIR>
__device__ __forceinline__ void line_fft(cplx<R> *buf, int nl, int ldist, int estr,
                                         const cplx<R> *tw) {
    constexpr int A = split_a(N), B = N / A;
    constexpr int LPW = lines_per_wave<N>();
    constexpr int LPR = (FT / 64) * LPW;  // lines per round
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int t = lane / LPW, r = wave * LPW + lane - (lane / LPW) * LPW;
    for (int l0 = 0; l0 < nl; l0 += LPR) {
        const int l = l0 + r;
        const bool act = t < B && l < nl;
        cplx<R> *line = buf + l * ldist;
        cplx<R> v[B > A ? B : A];
        if (act && t < A) {
#pragma unroll
            for (int j = 0; j < B; j++) v[j] = line[(t + A * j) * estr];
            dft<B, DIR, R>(v);
#pragma unroll
            for (int k = 1; k < B; k++) {
                cplx<R> w1 = tw[t * k];
                if (DIR < 0) w1.y = -w1.y;
                v[k] = cmul(v[k], w1);
            }
#pragma unroll
            for (int k = 0; k < B; k++) line[(t + A * k) * estr] = v[k];
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int s = 0; s < A; s++) v[s] = line[(s + A * t) * estr];
            dft<A, DIR, R>(v);
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < A; q++) line[(t + B * q) * estr] = v[q];
        }
        __syncthreads();
    }
}

// z columns per block of the column kernels: one round of line_fft
template <int N> constexpr int col_tile() { return (FT / 64) * lines_per_wave<N>(); }

template <typename R>
__device__ __forceinline__ void load_twiddles(cplx<R> *dst, const double2 *src, int n) {
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double2 t = src[j];
        dst[j] = mk<R>((R)t.x, (R)t.y);
    }
}

// ---------------------------------------------------------------------------
// psi(G) -> colbuf[band][z][col]: scatter + inverse FFT along z
// ---------------------------------------------------------------------------
template <typename R, int N3>
__global__ __launch_bounds__(FT) void k_col_bwd_fast(
    const double2 *__restrict__ psi, long ldp, cplx<R> *__restrict__ colbuf,
    const int *__restrict__ col_ptr, const int *__restrict__ ent_i, const int *__restrict__ ent_cz,
    const double2 *__restrict__ tw3, FftGeom g, int ntiles, int nb) {
    constexpr int CT = col_tile<N3>();
    constexpr int LS = N3 + 1;
    __shared__ cplx<R> tw[N3];
    __shared__ cplx<R> lines[CT * LS];
    int band, tile;
    if (!xcd_band_tile(blockIdx.x, ntiles, nb, band, tile)) return;
    const int c0 = tile * CT;
    const int nc = min(CT, g.ncol - c0);
    const double2 *pb = psi + (long)band * ldp;
    const int e0 = col_ptr[c0], e1 = col_ptr[c0 + nc];
    // The first CB_PRE sphere entries of each thread (index, then value: two
    // dependent global round trips) are loaded before the twiddles and the
    // clearing of the lines, which they overlap; each tile has at least one
    // entry, so a clamped entry index is always valid.
    constexpr int CB_PRE = 4;
    int pcz[CB_PRE];
    double2 pv[CB_PRE];
    {
        int pi[CB_PRE];
#pragma unroll
        for (int q = 0; q < CB_PRE; q++) {
            const int e = e0 + (int)threadIdx.x + q * FT;
            const int ec = e < e1 ? e : e0;
            pcz[q] = ent_cz[ec];
            pi[q] = ent_i[ec];
        }
#pragma unroll
        for (int q = 0; q < CB_PRE; q++) pv[q] = pb[pi[q]];
    }
    load_twiddles(tw, tw3, N3);
    for (int t = threadIdx.x; t < CT * LS; t += FT) lines[t] = mk<R>(0, 0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CB_PRE; q++)
        if (e0 + (int)threadIdx.x + q * FT < e1)
            lines[((pcz[q] >> 12) - c0) * LS + (pcz[q] & 0xfff)] = mk<R>((R)pv[q].x, (R)pv[q].y);
    for (int e = e0 + (int)threadIdx.x + CB_PRE * FT; e < e1; e += FT) {
        const int cz = ent_cz[e];
        const double2 p = pb[ent_i[e]];
        lines[((cz >> 12) - c0) * LS + (cz & 0xfff)] = mk<R>((R)p.x, (R)p.y);
    }
    __syncthreads();
    line_fft<R, N3, +1>(lines, nc, LS, 1, tw);
    cplx<R> *ob = colbuf + (long)band * N3 * g.ncolp + c0;
    for (int t = threadIdx.x; t < N3 * CT; t += FT) {
        const int z = t / CT, cc = t - z * CT;
        if (cc < nc) ob[(long)z * g.ncolp + cc] = lines[cc * LS + z];
    }
}

// ---------------------------------------------------------------------------
// colbuf -> forward FFT along z -> hpsi = local/N (+ ekpg psi) (+ hpsi)
// ---------------------------------------------------------------------------
template <typename R, int N3>
__global__ __launch_bounds__(FT) void k_col_fwd_fast(
    const double2 *__restrict__ psi, long ldp, double2 *__restrict__ hpsi, long ldh,
    const cplx<R> *__restrict__ colbuf, const int *__restrict__ col_ptr,
    const int *__restrict__ ent_i, const int *__restrict__ ent_cz, const double *__restrict__ ekpg,
    const double2 *__restrict__ tw3, FftGeom g, int ntiles, int nb, double inv_n, int mode) {
    constexpr int CT = col_tile<N3>();
    constexpr int LS = N3 + 1;
    __shared__ cplx<R> tw[N3];
    __shared__ cplx<R> lines[CT * LS];
    int band, tile;
    if (!xcd_band_tile(blockIdx.x, ntiles, nb, band, tile)) return;
    const int c0 = tile * CT;
    const int nc = min(CT, g.ncol - c0);
    load_twiddles(tw, tw3, N3);
    const double2 *pb = psi + (long)band * ldp;
    double2 *hb = hpsi + (long)band * ldh;
    const int e0 = col_ptr[c0], e1 = col_ptr[c0 + nc];
    // The first CF_PRE sphere entries of each thread: index, and the terms
    // added to local/N (ekpg*psi and/or what hpsi already holds) are loaded
    // before the transform, so their two dependent global round trips overlap it.
    const bool kin = mode & COLF_KIN, acc = mode & COLF_ACC;
    constexpr int CF_PRE = 4;
    int pcz[CF_PRE], pi[CF_PRE];
    double2 ph[CF_PRE];
#pragma unroll
    for (int q = 0; q < CF_PRE; q++) {
        const int e = e0 + (int)threadIdx.x + q * FT;
        const int ec = e < e1 ? e : e0;  // clamped: always a valid entry
        pcz[q] = ent_cz[ec];
        pi[q] = ent_i[ec];
    }
#pragma unroll
    for (int q = 0; q < CF_PRE; q++) {
        ph[q] = acc ? hb[pi[q]] : make_double2(0.0, 0.0);
        if (kin) {
            const double2 p = pb[pi[q]];
            const double ek = ekpg[pi[q]];
            ph[q].x += ek * p.x;
            ph[q].y += ek * p.y;
        }
    }
    // the tile's columns into the lines: every load of a thread issued before
    // the first LDS store, so the round trips overlap (loads at clamped,
    // valid positions; only the positions that exist are stored)
    const cplx<R> *ib = colbuf + (long)band * N3 * g.ncolp + c0;
    constexpr int NPT = (N3 * CT + FT - 1) / FT;
    cplx<R> cv[NPT];
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        const int t = min((int)threadIdx.x + u * FT, N3 * CT - 1);
        const int z = t / CT, cc = t - z * CT;
        cv[u] = ib[(long)z * g.ncolp + min(cc, nc - 1)];
    }
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        const int t = (int)threadIdx.x + u * FT;
        const int z = t / CT, cc = t - z * CT;
        if (t < N3 * CT && cc < nc) lines[cc * LS + z] = cv[u];
    }
    __syncthreads();
    line_fft<R, N3, -1>(lines, nc, LS, 1, tw);
    // hpsi = local/N + ekpg psi (+ the non-local term already in hpsi)
#pragma unroll
    for (int q = 0; q < CF_PRE; q++) {
        if (e0 + (int)threadIdx.x + q * FT < e1) {
            const cplx<R> v = lines[((pcz[q] >> 12) - c0) * LS + (pcz[q] & 0xfff)];
            hb[pi[q]] = make_double2((double)v.x * inv_n + ph[q].x, (double)v.y * inv_n + ph[q].y);
        }
    }
    for (int e = e0 + threadIdx.x + CF_PRE * FT; e < e1; e += FT) {
        const int cz = ent_cz[e];
        const int i = ent_i[e];
        const cplx<R> v = lines[((cz >> 12) - c0) * LS + (cz & 0xfff)];
        double2 h = make_double2((double)v.x * inv_n, (double)v.y * inv_n);
        if (kin) {
            const double2 p = pb[i];
            const double ek = ekpg[i];
            h.x += ek * p.x;
            h.y += ek * p.y;
        }
        if (acc) {
            const double2 o = hb[i];
            h.x += o.x;
            h.y += o.y;
        }
        hb[i] = h;
    }
}

// ---------------------------------------------------------------------------
// plane (z, band): inverse y FFT on the box lines, per row inverse x FFT,
// V(r), forward x FFT, forward y FFT.  LDS: tw1[N1] tw2[N2] sub[N2][nxp]
// rows[RB][N1+1] (sub/rows in dynamic shared memory, nxp depends on kd1).
// ---------------------------------------------------------------------------
// Row stride of the plane kernels' y/x-box sub-plane: nx box columns, then a
// zero column and a trash column (nx is odd, so the stride is odd too).  No
// further padding: at 72^3 four FP32 blocks (two FP64) fill the 160 KB of LDS
// of a CU with a few hundred bytes to spare.
__host__ __device__ __forceinline__ int plane_sub_stride(int nx) { return nx + 2; }

// Row scratch of the x transforms: element p of a row at p + p/A1 (one pad per
// A1 elements), so that the stride-A1 and stride-B1 accesses of the two passes
// are odd strides in 16-byte units; rows RS apart.
template <int N1> constexpr int row_stride() { return N1 + N1 / split_a(N1) + 1; }
template <int N1> __device__ __forceinline__ int rpos(int p) { return p + p / split_a(N1); }

template <typename R, int N1, int N2>
__global__ __launch_bounds__(FT) void k_plane_fast(cplx<R> *__restrict__ colbuf,
                                                   const int *__restrict__ col_uy,
                                                   const R *__restrict__ vr,
                                                   const double2 *__restrict__ tw1g,
                                                   const double2 *__restrict__ tw2g, FftGeom g) {
    constexpr int A1 = split_a(N1), B1 = N1 / A1;
    constexpr int RB = FT / B1;
    constexpr int RS = row_stride<N1>();
    __shared__ cplx<R> tw1[N1];
    __shared__ cplx<R> tw2[N2];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = g.nx, kd1 = g.kd[0];
    // sub rows: nx box columns, a column that stays zero (loads of x in the
    // gap of the box) and a trash column (stores of x in the gap)
    const int nxp = plane_sub_stride(nx);
    cplx<R> *sub = reinterpret_cast<cplx<R> *>(smem_raw);  // [N2][nxp]
    cplx<R> *rows = sub + N2 * nxp;                        // [RB][RS]
    const int z = blockIdx.x % g.n[2];
    const int band = blockIdx.x / g.n[2];

    load_twiddles(tw1, tw1g, N1);
    load_twiddles(tw2, tw2g, N2);
    for (int t = threadIdx.x; t < N2 * nxp; t += FT) sub[t] = mk<R>(0, 0);
    __syncthreads();
    cplx<R> *cb = colbuf + ((long)band * g.n[2] + z) * g.ncolp;
    for (int c = threadIdx.x; c < g.ncol; c += FT) {
        const int uy = col_uy[c];
        sub[(uy >> 16) * nxp + (uy & 0xffff)] = cb[c];
    }
    __syncthreads();
    line_fft<R, N2, +1>(sub, nx, 1, nxp, tw2);

    const R *vz = vr + (long)z * N2 * g.id;
    const int xhi = N1 - kd1;            // x >= xhi maps to u = x - (N1 - nx)
    const int ushift = N1 - nx;
    const int r = threadIdx.x / B1, t = threadIdx.x - (threadIdx.x / B1) * B1;
    cplx<R> *row = rows + r * RS;
    // sub column of x = t + A1 j for the loads (gap -> zero column) and the
    // stores (gap -> trash column): fixed per thread for every row
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
        cplx<R> v[B1 > A1 ? B1 : A1];
        // V(r) of this thread's points, loaded before the first pass so that
        // the global round trip overlaps it
        R vv[A1];
        if (act) {
            const R *vy = vz + (long)y * g.id;
#pragma unroll
            for (int q = 0; q < A1; q++) vv[q] = vy[t + B1 * q];
        }
        // inverse x transform, pass 1: x = t + A1 j, only box x are nonzero
        const cplx<R> *srow = sub + y * nxp;
        if (act && t < A1) {
#pragma unroll
            for (int j = 0; j < B1; j++) v[j] = srow[offl[j]];
            dft<B1, +1, R>(v);
#pragma unroll
            for (int k = 0; k < B1; k++) row[rpos<N1>(t + A1 * k)] = k == 0 ? v[0] : cmul(v[k], tw1[t * k]);
        }
        __syncthreads();
        if (act) {
            // pass 2 -> psi(r) at x = t + B1 q, times V(r), then the forward
            // transform's first pass on the same registers (split B1 x A1)
#pragma unroll
            for (int s = 0; s < A1; s++) v[s] = row[rpos<N1>(s + A1 * t)];
            dft<A1, +1, R>(v);
#pragma unroll
            for (int q = 0; q < A1; q++) v[q] = scale(v[q], vv[q]);
            dft<A1, -1, R>(v);
#pragma unroll
            for (int m = 1; m < A1; m++) {
                cplx<R> w1 = tw1[t * m];
                w1.y = -w1.y;
                v[m] = cmul(v[m], w1);
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int m = 0; m < A1; m++) row[rpos<N1>(t + B1 * m)] = v[m];
        }
        __syncthreads();
        if (act && t < A1) {
#pragma unroll
            for (int k = 0; k < B1; k++) v[k] = row[rpos<N1>(k + B1 * t)];
            dft<B1, -1, R>(v);
            cplx<R> *wrow = sub + y * nxp;
#pragma unroll
            for (int p = 0; p < B1; p++) wrow[offs[p]] = v[p];
        }
        __syncthreads();
    }
    line_fft<R, N2, -1>(sub, nx, 1, nxp, tw2);
    for (int c = threadIdx.x; c < g.ncol; c += FT) {
        const int uy = col_uy[c];
        cb[c] = sub[(uy >> 16) * nxp + (uy & 0xffff)];
    }
}

template <int N1> constexpr int plane_rows_fast() { return FT / (N1 / split_a(N1)); }

}  // namespace cpwgpu
