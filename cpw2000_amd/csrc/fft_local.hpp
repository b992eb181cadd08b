// Local part of H|psi> on gfx950: psi(G) -> psi(r) -> V(r) psi(r) -> (V psi)(G),
// plus the kinetic term, for a block of bands.
//
// The reference (Src/hk_psi/reference/hk_psi_c16.f90:176-222) loops over bands
// and, per band, zeroes a full mesh, scatters psi, runs a pruned 3-D FFT
// (Src/fft/_cfft/cfft_wf_c16.f), multiplies by vscr, transforms back and
// gathers.  Here the same product is computed band-batched in three kernels
// that never materialise the full mesh in HBM:
//
//   k_col_bwd  : per (tile of z-columns, band): scatter the sphere entries of
//                each (x,y) column into LDS lines, inverse FFT along z, write
//                the column buffer colbuf[band][z][col].
//   k_plane    : per (z, band): load the plane's columns into an LDS sub-plane
//                that holds only the x values of the sphere box, inverse FFT
//                along y on those lines, then row by row: inverse FFT along
//                x, multiply by V(r), forward FFT along x, keep the box x
//                values; forward FFT along y; write the columns back.
//   k_col_fwd  : per (tile of columns, band): forward FFT along z, 1/N,
//                gather the sphere entries and add 1/2|k+G|^2 psi in FP64.
//
// The mathematics equals the reference: the product it returns is
//   (V psi)(G_i) = (1/N) sum_r V(r) e^{-i G_i r} sum_j psi_j e^{+i G_j r}.
// Only the order of the three 1-D passes and the rounding differ.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_core.hpp"

namespace cpwgpu {

// what the gather of the forward column kernels adds to local/N
constexpr int COLF_ACC = 1;  // the value already in hpsi (non-local + kinetic terms)
constexpr int COLF_KIN = 2;  // ekpg * psi

// Geometry of one k-point's FFT problem (built on the host).
constexpr int MAXRAD = 12;
struct FftGeom {
    int n[3];          // n1, n2, n3
    int id;            // leading dimension of vscr (x stride of the mesh)
    int kd[3];         // max |G_j| of the wave-function sphere
    int nx;            // number of x values of the sphere box (2*kd1+1)
    int ncol, ncolp;   // distinct (x,y) columns and their padded stride in colbuf
    int nrad[3];
    int rad[3][MAXRAD];  // radices of n1, n2, n3 (4s, then 2, then odd)
    int mtxd;
};

// x of the compressed box index u (0..nx-1): 0..kd1, then n1-kd1..n1-1
__device__ __forceinline__ int box_x(int u, int kd1, int nx, int n1) { return u <= kd1 ? u : n1 - nx + u; }

// radix-r DFT with runtime r and direction (generic kernels)
template <typename R> __device__ __forceinline__ void dft_r(cplx<R> *v, int r, int dir) {
    if (dir > 0) {
        switch (r) {
        case 2: dft<2, 1, R>(v); break;
        case 3: dft<3, 1, R>(v); break;
        case 4: dft<4, 1, R>(v); break;
        default: dft<5, 1, R>(v); break;
        }
    } else {
        switch (r) {
        case 2: dft<2, -1, R>(v); break;
        case 3: dft<3, -1, R>(v); break;
        case 4: dft<4, -1, R>(v); break;
        default: dft<5, -1, R>(v); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Batched in-place Stockham FFT of lines held in LDS.
//   nl lines; line l starts at buf + l*ldist; element j at + j*estr.
//   tw[j] = exp(+2 pi i j / n), j < n (in LDS).
// Every thread of the block must call it (it contains barriers).  Each
// butterfly's inputs are held in registers across the barrier that separates
// reads from writes, so the transform needs no second buffer.
// KMAX bounds the butterflies per thread per stage: nl*n/R <= KMAX*blockDim.
// ---------------------------------------------------------------------------
template <typename R, int KMAX>
__device__ void lds_fft_batch(cplx<R> *buf, int nl, int ldist, int estr, int n, const int *rad,
                              int nrad, const cplx<R> *tw, int dir) {
    int L = 1;
    for (int s = 0; s < nrad; s++) {
        const int r = rad[s];
        const int m = n / r;
        const int items = nl * m;
        const int tstep = n / (L * r);
        cplx<R> v[KMAX][5];
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int w = threadIdx.x + k * blockDim.x;
            if (w < items) {
                const int l = w / m, j = w - l * m, kk = j % L;
                cplx<R> *line = buf + l * ldist;
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    if (q < r) {
                        cplx<R> x = line[(j + q * m) * estr];
                        int e = (kk * q * tstep) % n;
                        cplx<R> t = tw[e];
                        if (dir < 0) t.y = -t.y;
                        v[k][q] = q == 0 ? x : cmul(x, t);
                    }
                }
                dft_r(v[k], r, dir);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int w = threadIdx.x + k * blockDim.x;
            if (w < items) {
                const int l = w / m, j = w - l * m, kk = j % L;
                cplx<R> *line = buf + l * ldist;
                const int base = (j / L) * L * r + kk;
#pragma unroll
                for (int q = 0; q < 5; q++)
                    if (q < r) line[(base + q * L) * estr] = v[k][q];
            }
        }
        __syncthreads();
        L *= r;
    }
}

// Splits the lines into batches small enough for KMAX butterflies per thread.
template <typename R, int KMAX>
__device__ void lds_fft(cplx<R> *buf, int nl, int ldist, int estr, int n, const int *rad, int nrad,
                        const cplx<R> *tw, int dir) {
    int rmin = 5;
    for (int s = 0; s < nrad; s++) rmin = min(rmin, rad[s]);
    const int lpb = max(1, (KMAX * (int)blockDim.x) / (n / rmin));
    for (int l0 = 0; l0 < nl; l0 += lpb)
        lds_fft_batch<R, KMAX>(buf + l0 * ldist, min(lpb, nl - l0), ldist, estr, n, rad, nrad, tw, dir);
}

template <typename R> __device__ __forceinline__ cplx<R> ld_psi(const double2 *p) {
    double2 v = *p;
    return mk<R>((R)v.x, (R)v.y);
}

template <typename R>
__device__ void load_tw(cplx<R> *dst, const double2 *src, int n) {
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double2 t = src[j];
        dst[j] = mk<R>((R)t.x, (R)t.y);
    }
}

constexpr int COL_THREADS = 256;
constexpr int PLANE_THREADS = 256;

// ---------------------------------------------------------------------------
// k_col_bwd: psi(G) -> colbuf[band][z][col] (inverse FFT along z)
// grid.x = ntiles * nbands (tile fastest), CT columns per tile.
// ---------------------------------------------------------------------------
template <typename R, int CT, int KMAX>
__global__ __launch_bounds__(COL_THREADS) void k_col_bwd(
    const double2 *__restrict__ psi, long ldp, cplx<R> *__restrict__ colbuf,
    const int *__restrict__ col_ptr, const int *__restrict__ ent_i, const int *__restrict__ ent_cz,
    const double2 *__restrict__ tw3, FftGeom g, int ntiles) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n3 = g.n[2];
    const int ls = n3 + 1;
    cplx<R> *tw = reinterpret_cast<cplx<R> *>(smem_raw);
    cplx<R> *lines = tw + n3;
    const int tile = blockIdx.x % ntiles;
    const int band = blockIdx.x / ntiles;
    const int c0 = tile * CT;
    const int nc = min(CT, g.ncol - c0);

    load_tw(tw, tw3, n3);
    for (int t = threadIdx.x; t < CT * ls; t += blockDim.x) lines[t] = mk<R>(0, 0);
    __syncthreads();
    const double2 *pb = psi + (long)band * ldp;
    const int e0 = col_ptr[c0], e1 = col_ptr[c0 + nc];
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int cz = ent_cz[e];  // (column << 12) | z, entries sorted by column
        lines[((cz >> 12) - c0) * ls + (cz & 0xfff)] = ld_psi<R>(pb + ent_i[e]);
    }
    __syncthreads();
    lds_fft<R, KMAX>(lines, nc, ls, 1, n3, g.rad[2], g.nrad[2], tw, +1);
    cplx<R> *ob = colbuf + (long)band * n3 * g.ncolp + c0;
    for (int t = threadIdx.x; t < n3 * CT; t += blockDim.x) {
        const int z = t / CT, cc = t - z * CT;
        if (cc < nc) ob[(long)z * g.ncolp + cc] = lines[cc * ls + z];
    }
}

// ---------------------------------------------------------------------------
// k_plane: per (z, band): y and x transforms of one plane, V(r) product.
// grid.x = n3 * nbands (z fastest).
// LDS: twiddles n1 + n2, sub-plane [n2][nxp], row scratch [rb][n1+1].
// ---------------------------------------------------------------------------
template <typename R, int KMAX>
__global__ __launch_bounds__(PLANE_THREADS) void k_plane(
    cplx<R> *__restrict__ colbuf, const int *__restrict__ col_uy, const R *__restrict__ vr,
    const double2 *__restrict__ tw1, const double2 *__restrict__ tw2, FftGeom g, int rb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
    const int nx = g.nx, kd1 = g.kd[0];
    const int nxp = nx | 1;
    const int rs = n1 + 1;
    cplx<R> *t1 = reinterpret_cast<cplx<R> *>(smem_raw);
    cplx<R> *t2 = t1 + n1;
    cplx<R> *sub = t2 + n2;                // [n2][nxp]
    cplx<R> *rows = sub + n2 * nxp;        // [rb][rs]
    const int z = blockIdx.x % n3;
    const int band = blockIdx.x / n3;

    load_tw(t1, tw1, n1);
    load_tw(t2, tw2, n2);
    for (int t = threadIdx.x; t < n2 * nxp; t += blockDim.x) sub[t] = mk<R>(0, 0);
    __syncthreads();
    cplx<R> *cb = colbuf + ((long)band * n3 + z) * g.ncolp;
    for (int c = threadIdx.x; c < g.ncol; c += blockDim.x) {
        const int uy = col_uy[c];
        sub[(uy >> 16) * nxp + (uy & 0xffff)] = cb[c];
    }
    __syncthreads();
    // inverse FFT along y on the nx box lines
    lds_fft<R, KMAX>(sub, nx, 1, nxp, n2, g.rad[1], g.nrad[1], t2, +1);
    // rows: inverse x FFT, V(r), forward x FFT
    const R *vz = vr + (long)z * n2 * g.id;
    for (int y0 = 0; y0 < n2; y0 += rb) {
        const int nr = min(rb, n2 - y0);
        for (int t = threadIdx.x; t < nr * n1; t += blockDim.x) {
            const int r = t / n1, x = t - r * n1;
            // x -> box index, or zero outside the box
            int u = -1;
            if (x <= kd1) u = x;
            else if (x >= n1 - nx + kd1 + 1) u = x - (n1 - nx);
            rows[r * rs + x] = u >= 0 ? sub[(y0 + r) * nxp + u] : mk<R>(0, 0);
        }
        __syncthreads();
        lds_fft<R, KMAX>(rows, nr, rs, 1, n1, g.rad[0], g.nrad[0], t1, +1);
        for (int t = threadIdx.x; t < nr * n1; t += blockDim.x) {
            const int r = t / n1, x = t - r * n1;
            rows[r * rs + x] = scale(rows[r * rs + x], vz[(long)(y0 + r) * g.id + x]);
        }
        __syncthreads();
        lds_fft<R, KMAX>(rows, nr, rs, 1, n1, g.rad[0], g.nrad[0], t1, -1);
        for (int t = threadIdx.x; t < nr * nx; t += blockDim.x) {
            const int r = t / nx, u = t - r * nx;
            sub[(y0 + r) * nxp + u] = rows[r * rs + box_x(u, kd1, nx, n1)];
        }
        __syncthreads();
    }
    // forward FFT along y on the box lines
    lds_fft<R, KMAX>(sub, nx, 1, nxp, n2, g.rad[1], g.nrad[1], t2, -1);
    for (int c = threadIdx.x; c < g.ncol; c += blockDim.x) {
        const int uy = col_uy[c];
        cb[c] = sub[(uy >> 16) * nxp + (uy & 0xffff)];
    }
}

// ---------------------------------------------------------------------------
// k_col_fwd: colbuf -> forward FFT along z -> hpsi(G) = local/N (+ ekpg*psi if
// mode & COLF_KIN) (+ the term already in hpsi if mode & COLF_ACC)
// ---------------------------------------------------------------------------
template <typename R, int CT, int KMAX>
__global__ __launch_bounds__(COL_THREADS) void k_col_fwd(
    const double2 *__restrict__ psi, long ldp, double2 *__restrict__ hpsi, long ldh,
    const cplx<R> *__restrict__ colbuf, const int *__restrict__ col_ptr,
    const int *__restrict__ ent_i, const int *__restrict__ ent_cz, const double *__restrict__ ekpg,
    const double2 *__restrict__ tw3, FftGeom g, int ntiles, double inv_n, int mode) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n3 = g.n[2];
    const int ls = n3 + 1;
    cplx<R> *tw = reinterpret_cast<cplx<R> *>(smem_raw);
    cplx<R> *lines = tw + n3;
    const int tile = blockIdx.x % ntiles;
    const int band = blockIdx.x / ntiles;
    const int c0 = tile * CT;
    const int nc = min(CT, g.ncol - c0);

    load_tw(tw, tw3, n3);
    const cplx<R> *ib = colbuf + (long)band * n3 * g.ncolp + c0;
    for (int t = threadIdx.x; t < n3 * CT; t += blockDim.x) {
        const int z = t / CT, cc = t - z * CT;
        if (cc < nc) lines[cc * ls + z] = ib[(long)z * g.ncolp + cc];
    }
    __syncthreads();
    lds_fft<R, KMAX>(lines, nc, ls, 1, n3, g.rad[2], g.nrad[2], tw, -1);
    const double2 *pb = psi + (long)band * ldp;
    double2 *hb = hpsi + (long)band * ldh;
    const int e0 = col_ptr[c0], e1 = col_ptr[c0 + nc];
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int cz = ent_cz[e];
        const int i = ent_i[e];
        const cplx<R> v = lines[((cz >> 12) - c0) * ls + (cz & 0xfff)];
        double2 h = make_double2((double)v.x * inv_n, (double)v.y * inv_n);
        if (mode & COLF_KIN) {
            const double2 p = pb[i];
            const double ek = ekpg[i];
            h.x += ek * p.x;
            h.y += ek * p.y;
        }
        if (mode & COLF_ACC) {  // the non-local (and kinetic) term is already in hpsi
            const double2 o = hb[i];
            h.x += o.x;
            h.y += o.y;
        }
        hb[i] = h;
    }
}

}  // namespace cpwgpu
