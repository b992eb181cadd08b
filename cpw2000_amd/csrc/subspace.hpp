// FP64 subspace step on the GPU: Rayleigh-Ritz with Cholesky
// orthonormalisation, a Hermitian eigensolver, the rotation of psi / H psi and
// the residual norms.
//
// Reference semantics (the FP64 steps of the diagonaliser that calls hk_psi):
//   Src/ditsp/ditsp_c16_ritz.f90:195-200  H_sub = hbas^H bas, diagonalised
//   Src/ditsp/diag_c16.f90:65-86           zheevd of the subspace matrix
//   Src/ditsp/ditsp_c16_ritz.f90:284-287  rotation of the basis
//   Src/ditsp/grsch_loop_c16.f90:100-130  orthonormalisation (Gram-Schmidt;
//                                          here Cholesky of the overlap)
//   Src/ditsp/ditsp_error.f90:70-76        residual ||H psi - eps psi||^2
//
// The nb x nb matrices are column-major complex128 (element (i,j) at i + j*n).
// The dense steps run in one workgroup of SS_THREADS threads each (nb is at
// most a few hundred): no grid-wide synchronisation, deterministic order.
#pragma once

#include <hip/hip_runtime.h>

namespace cpwgpu {

constexpr int SS_THREADS = 1024;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// In place: lower Cholesky factor of the Hermitian positive definite a.
// info = 0, or j+1 if the j-th pivot is not positive.
__global__ __launch_bounds__(SS_THREADS) void k_cholesky(double2 *a, int n, int *info) {
    __shared__ double piv;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        if (threadIdx.x == 0) {
            const double d = a[j + (long)j * n].x;
            if (!(d > 0.0) && !bad) bad = j + 1;
            piv = d > 0.0 ? sqrt(d) : 1.0;
            a[j + (long)j * n] = make_double2(piv, 0.0);
        }
        __syncthreads();
        const double inv = 1.0 / piv;
        for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) {
            double2 v = a[i + (long)j * n];
            a[i + (long)j * n] = make_double2(v.x * inv, v.y * inv);
        }
        __syncthreads();
        const int m = n - j - 1;
        for (long idx = threadIdx.x; idx < (long)m * m; idx += blockDim.x) {
            const int i = j + 1 + (int)(idx % m), k = j + 1 + (int)(idx / m);
            if (i >= k) {
                const double2 lij = a[i + (long)j * n], lkj = a[k + (long)j * n];
                double2 v = a[i + (long)k * n];
                const double2 p = zmul(lij, zconj(lkj));
                a[i + (long)k * n] = make_double2(v.x - p.x, v.y - p.y);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *info = bad;
}

// In place on h: h <- L^-1 h L^-H, with L the lower factor in l.  w is n*n
// scratch.  One thread per column for each triangular solve.
__global__ __launch_bounds__(SS_THREADS) void k_to_standard(const double2 *l, double2 *h, double2 *w,
                                                            int n) {
    // w = L^-1 h  (forward substitution on each column of h)
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        for (int i = 0; i < n; i++) {
            double2 s = h[i + (long)c * n];
            for (int k = 0; k < i; k++) {
                const double2 p = zmul(l[i + (long)k * n], w[k + (long)c * n]);
                s.x -= p.x;
                s.y -= p.y;
            }
            const double d = 1.0 / l[i + (long)i * n].x;
            w[i + (long)c * n] = make_double2(s.x * d, s.y * d);
        }
    }
    __syncthreads();
    // h = L^-1 w^H, so that h^H = w L^-H = L^-1 h_in L^-H
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        for (int i = 0; i < n; i++) {
            double2 s = zconj(w[c + (long)i * n]);
            for (int k = 0; k < i; k++) {
                const double2 p = zmul(l[i + (long)k * n], h[k + (long)c * n]);
                s.x -= p.x;
                s.y -= p.y;
            }
            const double d = 1.0 / l[i + (long)i * n].x;
            h[i + (long)c * n] = make_double2(s.x * d, s.y * d);
        }
    }
    __syncthreads();
    // h <- h^H (the result is Hermitian: symmetrise exactly)
    for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        if (i < j) {
            const double2 a = h[i + (long)j * n], b = h[j + (long)i * n];
            // result_ij = conj(h_ji) = h_ij for an exactly Hermitian result: average
            const double2 m = make_double2(0.5 * (b.x + a.x), 0.5 * (-b.y + a.y));
            h[i + (long)j * n] = m;
            h[j + (long)i * n] = make_double2(m.x, -m.y);
        } else if (i == j) {
            h[i + (long)i * n].y = 0.0;
        }
    }
}

// Cyclic two-sided Jacobi for the Hermitian matrix a (destroyed): on exit the
// eigenvalues ascend in eig and v holds the eigenvectors (columns).  Each
// round applies n/2 disjoint rotations at once (round-robin pairing); a sweep
// is n-1 rounds; stops when off(a) <= tol*|a| or after max_sweeps sweeps;
// elements below tol*|a|/n are not rotated.  Round r pairs the positions
// (slot k, slot np-1-k) of the rotating schedule: slot 0 holds index 0 and
// slot i >= 1 holds index 1 + (i - 1 - r) mod (np - 1).  pos is n ints of
// scratch (the final ordering), rot is unused (kept for the workspace layout).
__device__ __forceinline__ int rr_slot(int i, int r, int np) {
    if (i == 0) return 0;
    int m = (i - 1 - r) % (np - 1);
    if (m < 0) m += np - 1;
    return 1 + m;
}

constexpr int JAC_MAXPAIR = 1024;  // n <= 2048

__global__ __launch_bounds__(SS_THREADS) void k_jacobi(double2 *a, double2 *v, double *eig, int n,
                                                       int max_sweeps, double tol, int *pos,
                                                       double *rot, int *sweeps_out) {
    (void)rot;
    __shared__ double red[SS_THREADS];
    __shared__ int stop;
    __shared__ double rc[JAC_MAXPAIR], rs[JAC_MAXPAIR];
    __shared__ double2 ru[JAC_MAXPAIR];
    __shared__ short rp[JAC_MAXPAIR], rq[JAC_MAXPAIR];
    const int np = (n + 1) & ~1;  // even number of positions (one dummy if n odd)
    for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        v[idx] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    int sweep = 0;
    for (; sweep < max_sweeps; sweep++) {
        // convergence test: off-diagonal vs total Frobenius norm
        double off = 0.0, tot = 0.0;
        for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) {
            const int i = (int)(idx % n), j = (int)(idx / n);
            const double2 x = a[idx];
            const double m = x.x * x.x + x.y * x.y;
            tot += m;
            if (i != j) off += m;
        }
        red[threadIdx.x] = off;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        const double off_all = red[0];
        __syncthreads();
        red[threadIdx.x] = tot;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) stop = off_all <= tol * tol * red[0] || off_all == 0.0;
        // pairs with |a_pq| below thr are left alone: together they stay below
        // the stopping bound (n^2 thr^2 = tol^2 |a|^2), and rotating them only
        // stirs rounding noise that keeps the sweeps from ever meeting it
        const double thr = fmax(1e-300, tol * sqrt(red[0]) / n);
        __syncthreads();
        if (stop) break;
        for (int round = 0; round < np - 1; round++) {
            // rotation parameters of the pairs of this round
            for (int k = threadIdx.x; k < np / 2; k += blockDim.x) {
                int p = rr_slot(k, round, np), q = rr_slot(np - 1 - k, round, np);
                if (p > q) { const int t = p; p = q; q = t; }
                double c = 1.0, sn = 0.0;
                double2 u = make_double2(1.0, 0.0);
                if (q < n) {
                    const double app = a[p + (long)p * n].x, aqq = a[q + (long)q * n].x;
                    const double2 apq = a[p + (long)q * n];
                    // |apq| without squaring: once the element has decayed
                    // below ~1e-154, apq.x^2 + apq.y^2 is denormal and keeps
                    // only a few digits, |u| drifts from 1 and the phase
                    // factor would rescale row and column q
                    const double g = hypot(apq.x, apq.y);
                    if (g > thr) {
                        const double zeta = (aqq - app) / (2.0 * g);
                        const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                                         (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        if (t != 0.0) {  // a zero angle is the identity: no phase either
                            u = make_double2(apq.x / g, -apq.y / g);  // exp(-i phi)
                            c = 1.0 / sqrt(1.0 + t * t);
                            sn = t * c;
                        }
                    }
                }
                rc[k] = c;
                rs[k] = sn;
                ru[k] = u;
                rp[k] = (short)p;
                rq[k] = (short)q;
            }
            __syncthreads();
            // rows: p' = c p - conj(u) s q ; q' = s p + conj(u) c q
            for (long idx = threadIdx.x; idx < (long)(np / 2) * n; idx += blockDim.x) {
                const int k = (int)(idx / n), j = (int)(idx % n);
                const int p = rp[k], q = rq[k];
                const double s = rs[k];
                if (q >= n || (s == 0.0 && ru[k].y == 0.0 && ru[k].x == 1.0)) continue;  // identity
                const double c = rc[k];
                const double2 uc = zconj(ru[k]);
                const double2 ap = a[p + (long)j * n], aq = a[q + (long)j * n];
                const double2 t = zmul(uc, aq);
                a[p + (long)j * n] = make_double2(c * ap.x - s * t.x, c * ap.y - s * t.y);
                a[q + (long)j * n] = make_double2(s * ap.x + c * t.x, s * ap.y + c * t.y);
            }
            __syncthreads();
            // columns (of a and v): p' = c p - u s q ; q' = s p + u c q
            for (long idx = threadIdx.x; idx < (long)(np / 2) * n; idx += blockDim.x) {
                const int k = (int)(idx / n), i = (int)(idx % n);
                const int p = rp[k], q = rq[k];
                const double s = rs[k];
                if (q >= n || (s == 0.0 && ru[k].y == 0.0 && ru[k].x == 1.0)) continue;  // identity
                const double c = rc[k];
                const double2 u = ru[k];
                double2 ap = a[i + (long)p * n], aq = a[i + (long)q * n];
                double2 t = zmul(u, aq);
                a[i + (long)p * n] = make_double2(c * ap.x - s * t.x, c * ap.y - s * t.y);
                a[i + (long)q * n] = make_double2(s * ap.x + c * t.x, s * ap.y + c * t.y);
                ap = v[i + (long)p * n];
                aq = v[i + (long)q * n];
                t = zmul(u, aq);
                v[i + (long)p * n] = make_double2(c * ap.x - s * t.x, c * ap.y - s * t.y);
                v[i + (long)q * n] = make_double2(s * ap.x + c * t.x, s * ap.y + c * t.y);
            }
            __syncthreads();
        }
    }
    // eigenvalues in ascending order (rank by value, ties by index), vectors with them
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double e = a[i + (long)i * n].x;
        int r = 0;
        for (int j = 0; j < n; j++) {
            const double f = a[j + (long)j * n].x;
            r += (f < e) || (f == e && j < i);
        }
        pos[i] = r;
        eig[r] = e;
    }
    __syncthreads();
    // permute eigenvector columns into a (a is free now): a[:, rank] = v[:, i]
    for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        a[i + (long)pos[j] * n] = v[idx];
    }
    __syncthreads();
    for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) v[idx] = a[idx];
    if (threadIdx.x == 0) *sweeps_out = sweep;
}

// In place on v: v <- L^-H v (backward substitution, one thread per column)
__global__ __launch_bounds__(SS_THREADS) void k_back_substitute(const double2 *l, double2 *v, int n) {
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        for (int i = n - 1; i >= 0; i--) {
            double2 s = v[i + (long)c * n];
            for (int k = i + 1; k < n; k++) {
                // (L^H)_{ik} = conj(L_{ki})
                const double2 p = zmulc(l[k + (long)i * n], v[k + (long)c * n]);
                s.x -= p.x;
                s.y -= p.y;
            }
            const double d = 1.0 / l[i + (long)i * n].x;
            v[i + (long)c * n] = make_double2(s.x * d, s.y * d);
        }
    }
}

// res[n] = sum_i |hpsi_n[i] - eig[n] psi_n[i]|^2  (one block per band, fixed-order tree)
__global__ __launch_bounds__(256) void k_residual(const double2 *__restrict__ psi, long ldp,
                                                  const double2 *__restrict__ hpsi, long ldh,
                                                  const double *__restrict__ eig,
                                                  double *__restrict__ res, int mtxd) {
    __shared__ double red[256];
    const int n = blockIdx.x;
    const double e = eig[n];
    double s = 0.0;
    for (int i = threadIdx.x; i < mtxd; i += blockDim.x) {
        const double2 p = psi[(long)n * ldp + i], h = hpsi[(long)n * ldh + i];
        const double rx = h.x - e * p.x, ry = h.y - e * p.y;
        s += rx * rx + ry * ry;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) res[n] = red[0];
}

// copy nb columns of mtxd rows (ld src -> ld dst)
__global__ void k_copy_cols(const double2 *__restrict__ src, long lds, double2 *__restrict__ dst,
                            long ldd, int rows, int cols) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const int c = (int)(t / rows), r = (int)(t - (long)c * rows);
    dst[(long)c * ldd + r] = src[(long)c * lds + r];
}

}  // namespace cpwgpu
