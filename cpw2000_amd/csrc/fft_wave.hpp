// Plane stage with wave-local exchanges.
//
// Same transforms as k_plane_fast (two register passes around one LDS
// exchange per 1-D transform, the x row kept in registers from the inverse
// transform through V(r) into the forward one), but every line and every row
// is worked on by the lanes of a single wavefront.  The LDS exchange between
// the two passes of a line then only has to be ordered within that wave:
// a wave barrier with LDS fences instead of a block barrier.  Block
// barriers remain only between the phases (columns in, inverse y, rows,
// forward y, columns out), five per plane instead of about thirty, so the
// waves of a block run out of step and hide each other's LDS latency.
//
// Lane map (B = N / A work items per line, LPW = 64 / B lines per wave):
// lane t * LPW + i is work item t of the wave's line i; lanes beyond
// B * LPW idle.  Row scratch is private to each wave (LPW rows).
#pragma once

#include "fft_fast.hpp"

namespace cpwgpu {

// Orders the LDS accesses of one wavefront: writes before the barrier are
// seen by reads of any lane of the same wave after it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Two-pass transform of nl lines in LDS (line l at buf + l*ldist, element j
// at + j*estr) by T threads; line l belongs to wave (l mod LPR) / LPW.  The
// caller orders the phases around it with block barriers.
template <typename R, int N, int DIR, int T>
__device__ __forceinline__ void line_fft_wave(cplx<R> *buf, int nl, int ldist, int estr,
                                              const cplx<R> *tw) {
    constexpr int A = split_a(N), B = N / A;
    constexpr int LPW = 64 / B;
    constexpr int LPR = (T / 64) * LPW;  // lines per round of the block
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
        wave_sync();
        if (act) {
#pragma unroll
            for (int s = 0; s < A; s++) v[s] = line[(s + A * t) * estr];
            dft<A, DIR, R>(v);
        }
        wave_sync();
        if (act) {
#pragma unroll
            for (int q = 0; q < A; q++) line[(t + B * q) * estr] = v[q];
        }
        // the next round works on other lines: no barrier needed here
    }
}

// rows of x-row scratch per block of T threads
template <int N1, int T> constexpr int plane_rows_wave() { return (T / 64) * (64 / (N1 / split_a(N1))); }

// LDS bytes of k_plane_wave beyond its static twiddle tables
template <typename R, int N, int T> constexpr size_t plane_wave_lds(int nx) {
    return sizeof(cplx<R>) * ((size_t)N * plane_sub_stride(nx) + (size_t)plane_rows_wave<N, T>() * row_stride<N>());
}

template <typename R, int N1, int N2, int T, int W>
__global__ __launch_bounds__(T, W) void k_plane_wave(cplx<R> *__restrict__ colbuf,
                                                    const int *__restrict__ col_uy,
                                                    const R *__restrict__ vr,
                                                    const double2 *__restrict__ tw1g,
                                                    const double2 *__restrict__ tw2g, FftGeom g) {
    constexpr int A1 = split_a(N1), B1 = N1 / A1;
    constexpr int LPW = 64 / B1;                // rows per wave and round
    constexpr int RPR = (T / 64) * LPW;         // rows per round of the block
    constexpr int RS = row_stride<N1>();
    __shared__ cplx<R> tw1[N1];
    __shared__ cplx<R> tw2[N2];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = g.nx, kd1 = g.kd[0];
    const int nxp = plane_sub_stride(nx);
    cplx<R> *sub = reinterpret_cast<cplx<R> *>(smem_raw);  // [N2][nxp]
    cplx<R> *rows = sub + N2 * nxp;                        // [RPR][RS], LPW rows per wave
    const int z = blockIdx.x % g.n[2];
    const int band = blockIdx.x / g.n[2];

    load_twiddles(tw1, tw1g, N1);
    load_twiddles(tw2, tw2g, N2);
    for (int q = threadIdx.x; q < N2 * nxp; q += T) sub[q] = mk<R>(0, 0);
    __syncthreads();
    cplx<R> *cb = colbuf + ((long)band * g.n[2] + z) * g.ncolp;
    for (int c = threadIdx.x; c < g.ncol; c += T) {
        const int uy = col_uy[c];
        sub[(uy >> 16) * nxp + (uy & 0xffff)] = cb[c];
    }
    __syncthreads();
    line_fft_wave<R, N2, +1, T>(sub, nx, 1, nxp, tw2);
    __syncthreads();

    const R *vz = vr + (long)z * N2 * g.id;
    const int xhi = N1 - kd1;  // x >= xhi maps to u = x - (N1 - nx)
    const int ushift = N1 - nx;
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int t = lane / LPW, ri = lane - (lane / LPW) * LPW;
    const int r = wave * LPW + ri;  // scratch row of this lane's row
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
    for (int y0 = 0; y0 < N2; y0 += RPR) {
        const int y = y0 + r;
        const bool act = t < B1 && y < N2;
        cplx<R> v[B1 > A1 ? B1 : A1];
        const cplx<R> *srow = sub + y * nxp;
        // V(r) of this lane's points, loaded now so that the global round trip
        // overlaps the first pass
        R vv[A1];
        if (act) {
            const R *vy = vz + (long)y * g.id;
#pragma unroll
            for (int q = 0; q < A1; q++) vv[q] = vy[t + B1 * q];
        }
        if (act && t < A1) {
            // inverse x transform, pass 1: x = t + A1 j, only box x are nonzero
#pragma unroll
            for (int j = 0; j < B1; j++) v[j] = srow[offl[j]];
            dft<B1, +1, R>(v);
#pragma unroll
            for (int k = 0; k < B1; k++) row[rpos<N1>(t + A1 * k)] = k == 0 ? v[0] : cmul(v[k], tw1[t * k]);
        }
        wave_sync();
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
        wave_sync();
        if (act) {
#pragma unroll
            for (int m = 0; m < A1; m++) row[rpos<N1>(t + B1 * m)] = v[m];
        }
        wave_sync();
        if (act && t < A1) {
#pragma unroll
            for (int k = 0; k < B1; k++) v[k] = row[rpos<N1>(k + B1 * t)];
            dft<B1, -1, R>(v);
            cplx<R> *wrow = sub + y * nxp;
#pragma unroll
            for (int p = 0; p < B1; p++) wrow[offs[p]] = v[p];
        }
        // the scratch row is written again only after the next round's first
        // wave barrier, by the same wave: its reads above are done by then
        wave_sync();
    }
    __syncthreads();
    line_fft_wave<R, N2, -1, T>(sub, nx, 1, nxp, tw2);
    __syncthreads();
    for (int c = threadIdx.x; c < g.ncol; c += T) {
        const int uy = col_uy[c];
        cb[c] = sub[(uy >> 16) * nxp + (uy & 0xffff)];
    }
}

}  // namespace cpwgpu
