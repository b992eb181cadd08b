// Device kernels of the block iterative eigensolver (the device-resident
// counterpart of Src/ditsp/ditsp_c16.f90).  All arrays are complex128
// column-major with a leading dimension, as psi/hpsi at the boundary; the
// small matrices are dense column-major with ld = their row count.
#pragma once

#include <hip/hip_runtime.h>

namespace cpwgpu {

constexpr int DT_THREADS = 256;

// X[:, j] = e_j for j < m (rows 0..mtxd-1): the unit vectors whose images
// under H give the leading block of H (the small Hamiltonian hamsm that
// Src/ALFA/hamilt_kb.f90 builds analytically for Src/ditsp/ditsp_c16_start.f90:116).
__global__ __launch_bounds__(DT_THREADS) void k_unit_cols(double2 *__restrict__ x, long ldx, int mtxd,
                                                          int m) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= mtxd || j >= m) return;
    x[(long)j * ldx + i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
}

// hs(ms, ms) = he[0:ms, 0:ms] (ld of hs = ms)
__global__ __launch_bounds__(DT_THREADS) void k_leading_block(const double2 *__restrict__ he, long ldh,
                                                              int ms, double2 *__restrict__ hs) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= ms || j >= ms) return;
    hs[(long)j * ms + i] = he[(long)j * ldh + i];
}

// out[:, j] = (u[0:ms, j], 0, ..., 0) for j < ncols: eigenvectors of the
// small matrix padded with zeros up to mtxd (ditsp_c16_start.f90:209-217)
__global__ __launch_bounds__(DT_THREADS) void k_pad_rows(const double2 *__restrict__ u, long ldu, int ms,
                                                         double2 *__restrict__ out, long ldo, int mtxd,
                                                         int ncols) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= mtxd || j >= ncols) return;
    out[(long)j * ldo + i] = i < ms ? u[(long)j * ldu + i] : make_double2(0.0, 0.0);
}

// Diagonal of H in the plane-wave basis: ekpg_i + V(G=0) + sum_a xnlkb_a |anlga_ia|^2
// (the quantity Src/ALFA/hamilt_diag_kb.f90:93-124 accumulates)
__global__ __launch_bounds__(DT_THREADS) void k_hdiag(const double *__restrict__ ekpg,
                                                      const double2 *__restrict__ anl, long lda,
                                                      const double *__restrict__ xnl, int nanl, double v0,
                                                      double *__restrict__ hdiag, int mtxd) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x;
    if (i >= mtxd) return;
    double s = ekpg[i] + v0;
    for (int a = 0; a < nanl; a++) {
        const double2 b = anl[(long)a * lda + i];
        s += xnl[a] * (b.x * b.x + b.y * b.y);
    }
    hdiag[i] = s;
}

// Jacobi relaxation of the residuals of the active vectors
// (Src/omp/diag_rq_jac_c16.f90:67-79):
//   t[:, j] = -w .* (hx_n - e_n x_n),  w_i = 1 if hdiag_i - e_n < 1 else 1/(hdiag_i - e_n),
// with n = act[j].
__global__ __launch_bounds__(DT_THREADS) void k_precondition(
    const double2 *__restrict__ x, long ldx, const double2 *__restrict__ hx, long ldh,
    const double *__restrict__ eig, const double *__restrict__ hdiag, const int *__restrict__ act,
    double2 *__restrict__ t, long ldt, int mtxd, int m) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= mtxd || j >= m) return;
    const int n = act[j];
    const double e = eig[n];
    const double d = hdiag[i] - e;
    const double w = d < 1.0 ? 1.0 : 1.0 / d;
    const double2 a = x[(long)n * ldx + i], b = hx[(long)n * ldh + i];
    t[(long)j * ldt + i] = make_double2(-w * (b.x - e * a.x), -w * (b.y - e * a.y));
}

// t[:, j] <- t[:, j] / ||t[:, j]||, one block per column (a zero column stays zero)
__global__ __launch_bounds__(DT_THREADS) void k_normalize_cols(double2 *__restrict__ t, long ldt, int mtxd) {
    __shared__ double red[DT_THREADS];
    double2 *c = t + (long)blockIdx.x * ldt;
    double s = 0.0;
    for (int i = threadIdx.x; i < mtxd; i += DT_THREADS) s += c[i].x * c[i].x + c[i].y * c[i].y;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = DT_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const double nrm = red[0];
    const double f = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
    for (int i = threadIdx.x; i < mtxd; i += DT_THREADS) c[i] = make_double2(f * c[i].x, f * c[i].y);
}

// Deterministic random start vectors for columns [c0, c0 + ncols): uniform
// (-1/2, 1/2) real and imaginary parts from a counter hash (splitmix64 of
// seed, row, column), damped by 1/(1 + ekpg_i) so that the high plane waves
// start small, as in a preconditioned random guess.
__device__ inline unsigned long long dt_mix(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ __launch_bounds__(DT_THREADS) void k_random_cols(double2 *__restrict__ x, long ldx, int mtxd,
                                                            int c0, int ncols,
                                                            const double *__restrict__ ekpg,
                                                            unsigned long long seed) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= mtxd || j >= ncols) return;
    const int c = c0 + j;
    const unsigned long long h = dt_mix(seed ^ dt_mix(((unsigned long long)c << 32) | (unsigned)i));
    const double u1 = (double)(h >> 40) * (1.0 / 16777216.0) - 0.5;
    const double u2 = (double)((h >> 16) & 0xFFFFFF) * (1.0 / 16777216.0) - 0.5;
    const double w = 1.0 / (1.0 + ekpg[i]);
    x[(long)c * ldx + i] = make_double2(w * u1, w * u2);
}

// d[k] = |a(k, k)|^2: the squared Cholesky pivots, read by the host to
// truncate a nearly dependent block
__global__ __launch_bounds__(DT_THREADS) void k_diag_sq(const double2 *__restrict__ a, long lda, int n,
                                                        double *__restrict__ d) {
    const int k = blockIdx.x * DT_THREADS + threadIdx.x;
    if (k >= n) return;
    const double2 v = a[(long)k * lda + k];
    d[k] = v.x * v.x + v.y * v.y;
}

// dst[:, j] = src[:, idx[j]] for j < m (rows 0..mtxd-1)
__global__ __launch_bounds__(DT_THREADS) void k_gather_cols(const double2 *__restrict__ src, long lds,
                                                            const int *__restrict__ idx,
                                                            double2 *__restrict__ dst, long ldd, int mtxd,
                                                            int m) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= mtxd || j >= m) return;
    dst[(long)j * ldd + i] = src[(long)idx[j] * lds + i];
}

// dst(nr, nc) = src(r0:r0+nr, 0:nc) packed with ld nr (src ld lds)
__global__ __launch_bounds__(DT_THREADS) void k_pack_rows(const double2 *__restrict__ src, long lds, int r0,
                                                          int nr, int nc, double2 *__restrict__ dst) {
    const int r = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (r >= nr || j >= nc) return;
    dst[(long)j * nr + r] = src[(long)j * lds + r0 + r];
}

__global__ __launch_bounds__(DT_THREADS) void k_fill(double *__restrict__ d, int n, double v) {
    const int k = blockIdx.x * DT_THREADS + threadIdx.x;
    if (k < n) d[k] = v;
}

// One step of the three-term Chebyshev recurrence on n columns (ld ldy):
//   y2 = alpha (hy1 - c y1) + beta y0
// elementwise, so y2 may be hy1 (the Lanczos update w -= a v + b v_prev uses
// it with alpha = 1, c = a, beta = -b).
__global__ __launch_bounds__(DT_THREADS) void k_cheb(double2 *y2, const double2 *hy1, const double2 *y1,
                                                     const double2 *y0, long ldy, int mtxd, double c,
                                                     double alpha, double beta) {
    const int i = blockIdx.x * DT_THREADS + threadIdx.x, j = blockIdx.y;
    if (i >= mtxd) return;
    const long o = (long)j * ldy + i;
    const double2 h = hy1[o], a = y1[o];
    double2 r = make_double2(alpha * (h.x - c * a.x), alpha * (h.y - c * a.y));
    if (beta != 0.0) {
        const double2 b = y0[o];
        r.x += beta * b.x;
        r.y += beta * b.y;
    }
    y2[o] = r;
}

}  // namespace cpwgpu
