// Atomic forces of E = sum_n occ_n <psi_n| V_loc + V_NL |psi_n> at fixed psi
// (FP64), the adjoint of the local and non-local parts of H|psi> with respect
// to the atomic positions.
//
// Local (Src/ALFA/for_str_local_force.f90:102-145):
//   f_j(atom) = 2 pi sum_q q_j Re[i exp(-2 pi i q.rat) f_s(q) conj(rho(q))]
// with rho(q) = (1/N) sum_r rho(r) exp(-i q.r), rho(r) = sum_n occ_n |psi_n(r)|^2
// (Src/omp/charge_by_fft.f90:191-214).  rho(r) is formed per z-plane from the
// z-column buffer of k_col_bwd (same inverse transforms as H|psi>), and only
// its Fourier components on the box of the local potential's q vectors are
// kept: per plane a partial transform along x then y, then along z per q.
//
// Non-local (Src/ALFA/for_str_nl_kb.f90:286-362, :392-398; d beta/d rat_j =
// -2 pi i g_j beta, Src/omp/proj_nl_kb_drat.f90:218-220):
//   f_j(atom) = -sum_n occ_n sum_{a in atom} xnlkb_a 2 Re[conj(P0) 2 pi i Pj]
// with P0 = <beta|psi> and Pj = sum_G g_j conj(beta) psi from the split-K
// projection GEMM with a per-plane-wave weight g_j.
// All f are covariant (per unit fractional coordinate); the host converts to
// Cartesian with inv(avec)^T.
#pragma once

#include "fft_fast.hpp"
#include "fft_reg.hpp"

namespace cpwgpu {

// per (z, band group): rho(x,y) = sum_{bands in group} occ |psi(x,y,z)|^2, then
// T2[grp][z][q2][q1] = sum_y e^{-2 pi i q2 y/N2} sum_x e^{-2 pi i q1 x/N1} rho(x,y)
// for |q1| <= Q1, |q2| <= Q2.  LDS: tw1[N1] tw2[N2] static; dynamic
// sub[N2][nxp] (reused for the x-partial transform), rows[RB][row_stride], acc[N2][N1].
template <int N1, int N2>
__global__ __launch_bounds__(FT) void k_density_fast(const cplx<double> *__restrict__ colbuf,
                                                     const int *__restrict__ col_uy,
                                                     const double *__restrict__ occ,
                                                     const double2 *__restrict__ tw1g,
                                                     const double2 *__restrict__ tw2g, FftGeom g,
                                                     int nb, int bpg, int Q1, int Q2,
                                                     double2 *__restrict__ t2,
                                                     double *__restrict__ rhor = nullptr) {
    typedef double R;
    constexpr int A1 = split_a(N1), B1 = N1 / A1;
    constexpr int RB = FT / B1;
    constexpr int RS = row_stride<N1>();
    __shared__ cplx<R> tw1[N1];
    __shared__ cplx<R> tw2[N2];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = g.nx, kd1 = g.kd[0];
    const int nxp = nx | 1;
    cplx<R> *sub = reinterpret_cast<cplx<R> *>(smem_raw);  // [N2][nxp]
    cplx<R> *rows = sub + N2 * nxp;                        // [RB][RS]
    R *acc = reinterpret_cast<R *>(rows + RB * RS);        // [N2][N1]
    const int n3 = g.n[2];
    const int z = blockIdx.x % n3, grp = blockIdx.x / n3;
    const int b0 = grp * bpg, b1 = min(nb, b0 + bpg);
    load_twiddles(tw1, tw1g, N1);
    load_twiddles(tw2, tw2g, N2);
    for (int t = threadIdx.x; t < N2 * N1; t += FT) acc[t] = 0.0;
    const int xhi = N1 - kd1, ushift = N1 - nx;
    const int r = threadIdx.x / B1, t = threadIdx.x - (threadIdx.x / B1) * B1;
    cplx<R> *row = rows + r * RS;
    for (int b = b0; b < b1; b++) {
        for (int q = threadIdx.x; q < N2 * nxp; q += FT) sub[q] = mk<R>(0, 0);
        __syncthreads();
        const cplx<R> *cb = colbuf + ((long)b * n3 + z) * g.ncolp;
        for (int c = threadIdx.x; c < g.ncol; c += FT) {
            const int uy = col_uy[c];
            sub[(uy >> 16) * nxp + (uy & 0xffff)] = cb[c];
        }
        __syncthreads();
        line_fft<R, N2, +1>(sub, nx, 1, nxp, tw2);
        const R w = occ ? occ[b] : 1.0;
        for (int y0 = 0; y0 < N2; y0 += RB) {
            const int y = y0 + r;
            const bool act = r < RB && y < N2;
            cplx<R> v[B1 > A1 ? B1 : A1];
            if (act && t < A1) {
#pragma unroll
                for (int j = 0; j < B1; j++) {
                    const int x = t + A1 * j;
                    v[j] = mk<R>(0, 0);
                    if (x <= kd1) v[j] = sub[y * nxp + x];
                    else if (x >= xhi) v[j] = sub[y * nxp + x - ushift];
                }
                dft<B1, +1, R>(v);
#pragma unroll
                for (int k = 0; k < B1; k++) row[rpos<N1>(t + A1 * k)] = k == 0 ? v[0] : cmul(v[k], tw1[t * k]);
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int s = 0; s < A1; s++) v[s] = row[rpos<N1>(s + A1 * t)];
                dft<A1, +1, R>(v);
                // each (y, x = t + B1 q) belongs to this thread alone
#pragma unroll
                for (int q = 0; q < A1; q++) acc[y * N1 + t + B1 * q] += w * (v[q].x * v[q].x + v[q].y * v[q].y);
            }
            __syncthreads();
        }
    }
    if (rhor) {
        // rho(x, y) of this plane and band group, for a fixed-order sum over groups
        __syncthreads();
        double *o = rhor + ((long)grp * n3 + z) * (N2 * N1);
        for (int q = threadIdx.x; q < N2 * N1; q += FT) o[q] = acc[q];
        return;
    }
    // partial transform along x into sub: T1[y][q1 + Q1]
    const int W1 = 2 * Q1 + 1, W2 = 2 * Q2 + 1;
    for (int idx = threadIdx.x; idx < N2 * W1; idx += FT) {
        const int y = idx / W1, q1 = idx - y * W1 - Q1;
        double sr = 0.0, si = 0.0;
        for (int x = 0; x < N1; x++) {
            int e = (q1 * x) % N1;
            if (e < 0) e += N1;
            const cplx<R> w1 = tw1[e];  // exp(+2 pi i e/N1); use the conjugate
            const double a = acc[y * N1 + x];
            sr += a * w1.x;
            si -= a * w1.y;
        }
        sub[idx] = mk<R>(sr, si);
    }
    __syncthreads();
    double2 *out = t2 + ((long)grp * n3 + z) * W2 * W1;
    for (int idx = threadIdx.x; idx < W2 * W1; idx += FT) {
        const int q2 = idx / W1 - Q2, q1i = idx - (idx / W1) * W1;
        double sr = 0.0, si = 0.0;
        for (int y = 0; y < N2; y++) {
            int e = (q2 * y) % N2;
            if (e < 0) e += N2;
            const cplx<R> w2 = tw2[e];
            const cplx<R> a = sub[y * W1 + q1i];
            // a * conj(w2)
            sr += a.x * w2.x + a.y * w2.y;
            si += a.y * w2.x - a.x * w2.y;
        }
        out[idx] = make_double2(sr, si);
    }
}

// Register-line variant of the density pass of k_density_fast (the line
// layout of k_plane_reg): per (z, band group) each band's y/x-box sub-plane
// (LDS, S = nx + 2 columns with a zero column) goes through the inverse y and
// x transforms in registers.  An x-row task keeps the same rows for every band
// of the group, so occ |psi(x, y)|^2 accumulates in the registers of the lane
// that holds (x, y) (x = MX k2 + k after the decimation-in-frequency
// transform), and each block writes its rho(x, y) plane once:
// rhor[grp][z][y][x], summed over groups in a fixed order afterwards.
// Grid: n3 * ceil(nb / bpg) blocks of T threads.
// CS: the sphere box lies inside N/4 (checked by the caller) and the row
// stride is the compile-time pr_sub_stride<N>() instead of the argument.
template <int N, int GX, int GY, int T, int W, bool CS = false>
__global__ __launch_bounds__(T, W) void k_density_reg(const cplx<double> *__restrict__ colbuf,
                                                   const int *__restrict__ col_uy,
                                                   const double *__restrict__ occ,
                                                   const double2 *__restrict__ twg, FftGeom g, int nb,
                                                   int bpg, int S_arg, double *__restrict__ rhor) {
    typedef double R;
    static_assert(N % GX == 0 && N % GY == 0 && T % 64 == 0, "bad plane shape");
    const int S = CS ? pr_sub_stride<N>() : S_arg;
    constexpr int MX = N / GX, MY = N / GY;
    constexpr int TX = T / GX;                 // x-row tasks per round
    constexpr int RPT = (N + TX - 1) / TX;     // rows per x task
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R> *twN = reinterpret_cast<cplx<R> *>(smem_raw);  // [N]
    cplx<R> *sub = twN + N;                                 // [N][S]
    const int n3 = g.n[2];
    const int z = blockIdx.x % n3, grp = blockIdx.x / n3;
    const int b0 = grp * bpg, b1 = min(nb, b0 + bpg);
    const int nx = g.nx, kd1 = g.kd[0], ushift = N - nx, xhi = N - kd1;
    const int tid = threadIdx.x;
    load_twiddles(twN, twg, N);
    const Lanes<R, GX> LX;
    const Lanes<R, GY> LY;
    R racc[RPT][MX];
#pragma unroll
    for (int r = 0; r < RPT; r++)
#pragma unroll
        for (int k = 0; k < MX; k++) racc[r][k] = 0.0;
    for (int b = b0; b < b1; b++) {
        for (int q = tid; q < N * S; q += T) sub[q] = mk<R>(0, 0);
        __syncthreads();
        gather_plane<R, T>(sub, colbuf + ((long)b * n3 + z) * g.ncolp, col_uy, g.ncol, S, 0);
        __syncthreads();
        // inverse y on the nx box lines
        for (int u = tid / GY; u < nx; u += T / GY) {
            cplx<R> *col = sub + u;
            cplx<R> v[MY];
#pragma unroll
            for (int j = 0; j < MY; j++) v[j] = col[(LY.g + GY * j) * S];
            line_dif<+1, GY, MY>(v, LY, twN);
#pragma unroll
            for (int k = 0; k < MY; k++) col[(k + MY * LY.k2) * S] = v[k];
        }
        __syncthreads();
        // inverse x per row, occ |psi|^2 into the row's registers
        const R w = occ ? occ[b] : 1.0;
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int y = tid / GX + r * TX;
            if (y < N) {
                const cplx<R> *row = sub + y * S;
                cplx<R> v[MX];
#pragma unroll
                for (int j = 0; j < MX; j++) {
                    const int x = LX.g + GX * j;
                    v[j] = row[x <= kd1 ? x : x >= xhi ? x - ushift : nx];
                }
                line_dif<+1, GX, MX>(v, LX, twN);
#pragma unroll
                for (int k = 0; k < MX; k++) racc[r][k] += w * (v[k].x * v[k].x + v[k].y * v[k].y);
            }
        }
        __syncthreads();  // the rows are read before the next band clears sub
    }
    double *o = rhor + ((long)grp * n3 + z) * (N * N);
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int y = tid / GX + r * TX;
        if (y < N) {
#pragma unroll
            for (int k = 0; k < MX; k++) o[y * N + MX * LX.k2 + k] = racc[r][k];
        }
    }
}

// rho[(z n2 + y) id + x] = sum over band groups of part[grp][z][y][x] in
// group order (x < n1; the padding x >= n1 of the id-wide layout is zero)
__global__ __launch_bounds__(256) void k_rho_r_reduce(const double *__restrict__ part, int ngrp, int n1,
                                                      int n2, int n3, int id, double *__restrict__ rho) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    const long tot = (long)id * n2 * n3;
    if (q >= tot) return;
    const int x = (int)(q % id);
    const long yz = q / id;
    double s = 0.0;
    if (x < n1) {
        const long nmesh = (long)n1 * n2 * n3;
        const long src = yz * n1 + x;
        for (int gq = 0; gq < ngrp; gq++) s += part[gq * nmesh + src];
    }
    rho[q] = s;
}

// Per z-plane (one block each): rho(x, y) = sum over band groups of
// rhor[grp][z][y][x] in group order, then the partial transforms of
// k_density_fast: T2[z][q2][q1] = sum_y e^{-2 pi i q2 y/n2} sum_x
// e^{-2 pi i q1 x/n1} rho(x, y), |q1| <= Q1, |q2| <= Q2.  Dynamic LDS:
// tw1[n1], tw2[n2] (exp(+2 pi i e/n)), T1[n2][2 Q1 + 1], rho[n2][n1].
__global__ __launch_bounds__(256) void k_rho_partial(const double *__restrict__ rhor, int ngrp, int n1, int n2,
                                                     int n3, int Q1, int Q2, const double2 *__restrict__ tw1g,
                                                     const double2 *__restrict__ tw2g, double2 *__restrict__ t2) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int W1 = 2 * Q1 + 1, W2 = 2 * Q2 + 1;
    double2 *tw1 = reinterpret_cast<double2 *>(smem_raw);
    double2 *tw2 = tw1 + n1;
    double2 *t1 = tw2 + n2;                                  // [n2][W1]
    double *acc = reinterpret_cast<double *>(t1 + n2 * W1);  // [n2][n1]
    const int z = blockIdx.x;
    const long plane = (long)n1 * n2;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) tw1[i] = tw1g[i];
    for (int i = threadIdx.x; i < n2; i += blockDim.x) tw2[i] = tw2g[i];
    for (long i = threadIdx.x; i < plane; i += blockDim.x) {
        double v = 0.0;
        for (int grp = 0; grp < ngrp; grp++) v += rhor[((long)grp * n3 + z) * plane + i];
        acc[i] = v;
    }
    __syncthreads();
    // along x: T1[y][q1 + Q1] = sum_x rho(x, y) conj(tw1[q1 x mod n1])
    for (int idx = threadIdx.x; idx < n2 * W1; idx += blockDim.x) {
        const int y = idx / W1, q1 = idx - y * W1 - Q1;
        const int de = ((q1 % n1) + n1) % n1;
        int e = 0;
        double sr = 0.0, si = 0.0;
        for (int x = 0; x < n1; x++) {
            const double a = acc[y * n1 + x];
            sr += a * tw1[e].x;
            si -= a * tw1[e].y;
            e += de;
            if (e >= n1) e -= n1;
        }
        t1[idx] = make_double2(sr, si);
    }
    __syncthreads();
    // along y: T2[z][q2 + Q2][q1 + Q1] = sum_y T1[y][q1] conj(tw2[q2 y mod n2])
    double2 *out = t2 + (long)z * W2 * W1;
    for (int idx = threadIdx.x; idx < W2 * W1; idx += blockDim.x) {
        const int q2 = idx / W1 - Q2, q1i = idx - (idx / W1) * W1;
        const int de = ((q2 % n2) + n2) % n2;
        int e = 0;
        double sr = 0.0, si = 0.0;
        for (int y = 0; y < n2; y++) {
            const double2 a = t1[y * W1 + q1i], w = tw2[e];
            sr += a.x * w.x + a.y * w.y;
            si += a.y * w.x - a.x * w.y;
            e += de;
            if (e >= n2) e -= n2;
        }
        out[idx] = make_double2(sr, si);
    }
}

// T2s[z][q2][q1] = sum over band groups of T2[grp][z][q2][q1], in group order
__global__ void k_t2_sum(const double2 *__restrict__ t2, long nplane, int ngrp, double2 *__restrict__ t2s) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nplane) return;
    double sr = 0.0, si = 0.0;
    for (int grp = 0; grp < ngrp; grp++) {
        const double2 a = t2[grp * nplane + i];
        sr += a.x;
        si += a.y;
    }
    t2s[i] = make_double2(sr, si);
}

// rho(q) for each listed q: (1/N) sum_z T2s exp(-2 pi i q3 z / N3)
__global__ void k_rho_q(const double2 *__restrict__ t2s, const int *__restrict__ qv, int nq,
                        const double2 *__restrict__ tw3, int n3, int Q1, int Q2, double inv_n,
                        double2 *__restrict__ rho) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const int q1 = qv[3 * k], q2 = qv[3 * k + 1], q3 = qv[3 * k + 2];
    const int W1 = 2 * Q1 + 1, W2 = 2 * Q2 + 1;
    const long off = (long)(q2 + Q2) * W1 + (q1 + Q1);
    // twiddle index e = q3 z mod n3, advanced by de per plane
    const int de = ((q3 % n3) + n3) % n3;
    int e = 0;
    double sr = 0.0, si = 0.0;
    for (int z = 0; z < n3; z++) {
        const double2 a = t2s[(long)z * W2 * W1 + off];
        const double2 w = tw3[e];
        sr += a.x * w.x + a.y * w.y;
        si += a.y * w.x - a.x * w.y;
        e += de;
        if (e >= n3) e -= n3;
    }
    rho[k] = make_double2(sr * inv_n, si * inv_n);
}

__device__ __forceinline__ void block_sum3(double *red, double v0, double v1, double v2, double *out) {
    red[threadIdx.x] = v0;
    red[blockDim.x + threadIdx.x] = v1;
    red[2 * blockDim.x + threadIdx.x] = v2;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] += red[threadIdx.x + s];
            red[blockDim.x + threadIdx.x] += red[blockDim.x + threadIdx.x + s];
            red[2 * blockDim.x + threadIdx.x] += red[2 * blockDim.x + threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = red[blockDim.x];
        out[2] = red[2 * blockDim.x];
    }
}

// covariant local force of one atom per block (256 threads, fixed-order sums)
__global__ __launch_bounds__(256) void k_force_local(const double2 *__restrict__ rho,
                                                     const int *__restrict__ qv,
                                                     const double *__restrict__ fq, int nq,
                                                     const double *__restrict__ rat,
                                                     const int *__restrict__ species,
                                                     double *__restrict__ fout) {
    __shared__ double red[3 * 256];
    const int atom = blockIdx.x;
    const int sp = species[atom];
    const double r1 = rat[3 * atom], r2 = rat[3 * atom + 1], r3 = rat[3 * atom + 2];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = threadIdx.x; k < nq; k += blockDim.x) {
        const int q1 = qv[3 * k], q2 = qv[3 * k + 1], q3 = qv[3 * k + 2];
        double sn, cs;
        sincospi(2.0 * (q1 * r1 + q2 * r2 + q3 * r3), &sn, &cs);
        // i * exp(-i phi) * f * conj(rho): real part
        const double f = fq[(long)sp * nq + k];
        const double2 rh = rho[k];
        // exp(-i phi) conj(rho) = (cs - i sn)(rx - i ry) = (cs rx - sn ry) - i (cs ry + sn rx)
        const double im = -(cs * rh.y + sn * rh.x);
        const double e1 = -f * im;  // Re[i z] = -Im z
        s0 += q1 * e1;
        s1 += q2 * e1;
        s2 += q3 * e1;
    }
    block_sum3(red, s0, s1, s2, fout + 3 * atom);
    if (threadIdx.x == 0) {
        const double tp = 6.283185307179586477;
        fout[3 * atom] *= tp;
        fout[3 * atom + 1] *= tp;
        fout[3 * atom + 2] *= tp;
    }
}

// covariant non-local force and non-local energy of one atom per block.
// p holds 4 matrices (P0, P1, P2, P3), each [band][proj] (nanl fastest).
__global__ __launch_bounds__(256) void k_force_nl(const double2 *__restrict__ p, int nanl, int nb,
                                                  const double *__restrict__ xnl,
                                                  const double *__restrict__ occ,
                                                  const int *__restrict__ proj_atom,
                                                  double *__restrict__ fout,
                                                  double *__restrict__ eout) {
    __shared__ double red[3 * 256];
    const int atom = blockIdx.x;
    const long mat = (long)nanl * nb;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, en = 0.0;
    for (long idx = threadIdx.x; idx < mat; idx += blockDim.x) {
        const int a = (int)(idx % nanl), n = (int)(idx / nanl);
        if (proj_atom[a] != atom) continue;
        const double w = 2.0 * (occ ? occ[n] : 1.0) * xnl[a];
        const double2 p0 = p[idx];
        en += 0.5 * w * (p0.x * p0.x + p0.y * p0.y);
        double s[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double2 pj = p[(j + 1) * mat + idx];
            // Re[(-i pj) conj(p0)] = Re[-i (pj conj(p0))] = Im(pj conj(p0))
            s[j] = w * (pj.y * p0.x - pj.x * p0.y);
        }
        s0 += s[0];
        s1 += s[1];
        s2 += s[2];
    }
    block_sum3(red, s0, s1, s2, fout + 3 * atom);
    __syncthreads();
    red[threadIdx.x] = en;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double tp = 6.283185307179586477;
        fout[3 * atom] *= tp;
        fout[3 * atom + 1] *= tp;
        fout[3 * atom + 2] *= tp;
        eout[atom] = red[0];
    }
}

}  // namespace cpwgpu
