// Kleinman-Bylander non-local term on the matrix cores (gfx950 MFMA).
//
// Reference: Src/cpw_lib/hk_psi_nl_c16.f90:78-96
//   dhd  = anlga^H psi                 zgemm('c','n', nanl, neig, mtxd)
//   dhd  = Diag(xnlkb) dhd
//   hpsi = hpsi + anlga dhd            zgemm('n','n', mtxd, neig, nanl)
//
// Complex products are split into real MFMA products on separate real and
// imaginary accumulators (v_mfma_f64_16x16x4_f64 for FP64,
// v_mfma_f32_16x16x4_f32 for FP32, FP32 accumulation).
//
// MFMA 16x16x4 operand layout (wave64, lane l), measured on gfx950:
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   D (16x16): lane l, element r holds D[row][l%16] with
//              row = 4*(l/16) + r   for v_mfma_f32_16x16x4_f32
//              row = l/16 + 4*r     for v_mfma_f64_16x16x4_f64
#pragma once

#include <hip/hip_runtime.h>

#include "fft_local.hpp"

namespace cpwgpu {

template <typename R> struct MF;
template <> struct MF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct MF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

constexpr int GEMM_THREADS = 256;
constexpr int PJ_BM = 64;  // projectors per tile
constexpr int PJ_BN = 64;  // bands per tile
constexpr int PJ_BK = 16;  // plane waves per LDS stage

// ---------------------------------------------------------------------------
// k_proj: part[s][n][a] = sum_{i in slice s} conj(anlga[i,a]) psi[i,n]
// grid (ceil(nanl/64), ceil(neig/64), nslice), 4 waves each 32x32.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(GEMM_THREADS) void k_proj(const cplx<R> *__restrict__ anl, long lda,
                                                       const double2 *__restrict__ psi, long ldp,
                                                       cplx<R> *__restrict__ part, int mtxd,
                                                       int nanl, int neig, int kslice) {
    typedef typename MF<R>::v4 v4;
    __shared__ R sAr[PJ_BM][PJ_BK + 1], sAi[PJ_BM][PJ_BK + 1];
    __shared__ R sBr[PJ_BN][PJ_BK + 1], sBi[PJ_BN][PJ_BK + 1];
    const int a0 = blockIdx.x * PJ_BM, n0 = blockIdx.y * PJ_BN;
    const int kb = blockIdx.z * kslice, ke = min(mtxd, kb + kslice);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    v4 accR[2][2], accI[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            accR[i][j] = v4{0, 0, 0, 0};
            accI[i][j] = v4{0, 0, 0, 0};
        }
    const int lcol = threadIdx.x >> 2;       // 0..63
    const int lrow = (threadIdx.x & 3) * 4;  // 0,4,8,12
    for (int k0 = kb; k0 < ke; k0 += PJ_BK) {
        // stage A (projectors) and B (psi) tiles: 64 columns x 16 plane waves
        {
            const int a = a0 + lcol, n = n0 + lcol;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = k0 + lrow + q;
                cplx<R> va = mk<R>(0, 0);
                if (a < nanl && i < ke) va = anl[(long)a * lda + i];
                sAr[lcol][lrow + q] = va.x;
                sAi[lcol][lrow + q] = va.y;
                double2 vb = make_double2(0.0, 0.0);
                if (n < neig && i < ke) vb = psi[(long)n * ldp + i];
                sBr[lcol][lrow + q] = (R)vb.x;
                sBi[lcol][lrow + q] = (R)vb.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < PJ_BK; kk += 4) {
            R ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                ar[t] = sAr[wm * 32 + t * 16 + lr][kk + lk];
                ai[t] = sAi[wm * 32 + t * 16 + lr][kk + lk];
                br[t] = sBr[wn * 32 + t * 16 + lr][kk + lk];
                bi[t] = sBi[wn * 32 + t * 16 + lr][kk + lk];
            }
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int ni = 0; ni < 2; ni++) {
                    // conj(a) * b = (ar br + ai bi) + i (ar bi - ai br)
                    accR[mi][ni] = MF<R>::mma(ar[mi], br[ni], accR[mi][ni]);
                    accR[mi][ni] = MF<R>::mma(ai[mi], bi[ni], accR[mi][ni]);
                    accI[mi][ni] = MF<R>::mma(ar[mi], bi[ni], accI[mi][ni]);
                    accI[mi][ni] = MF<R>::mma(-ai[mi], br[ni], accI[mi][ni]);
                }
        }
        __syncthreads();
    }
    cplx<R> *ps = part + (long)blockIdx.z * neig * nanl;
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++) {
            const int n = n0 + wn * 32 + ni * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int a = a0 + wm * 32 + mi * 16 + MF<R>::row(lk, r);
                if (a < nanl && n < neig) ps[(long)n * nanl + a] = mk<R>(accR[mi][ni][r], accI[mi][ni][r]);
            }
        }
}

// dhd[n][a] = xnlkb[a] * sum_s part[s][n][a]   (fixed summation order: deterministic)
template <typename R>
__global__ void k_proj_reduce(const cplx<R> *__restrict__ part, const double *__restrict__ xnl,
                              cplx<R> *__restrict__ dhd, int nanl, int neig, int nslice) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)nanl * neig;
    if (t >= tot) return;
    const int a = (int)(t % nanl);
    R sr = 0, si = 0;
    for (int s = 0; s < nslice; s++) {
        cplx<R> v = part[(long)s * tot + t];
        sr += v.x;
        si += v.y;
    }
    const R x = (R)xnl[a];
    dhd[t] = mk<R>(x * sr, x * si);
}

constexpr int BP_BI = 64;  // plane waves per block (16 per wave)
constexpr int BP_BN = 64;  // bands per block
constexpr int BP_KC = 32;  // projectors per LDS stage

// ---------------------------------------------------------------------------
// k_back: hpsi[i,n] += sum_a anlga[i,a] dhd[a,n]
// computed as D^T = dhd^T anlga^T so that each MFMA output row is 16
// consecutive plane waves (coalesced 256-byte read-modify-write of hpsi).
// grid (ceil(mtxd/64), ceil(neig/64)); wave w owns plane waves i0+16w..+15.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(GEMM_THREADS) void k_back(const cplx<R> *__restrict__ anl, long lda,
                                                       const cplx<R> *__restrict__ dhd,
                                                       double2 *__restrict__ hpsi, long ldh,
                                                       int mtxd, int nanl, int neig) {
    typedef typename MF<R>::v4 v4;
    __shared__ R sDr[BP_BN][BP_KC + 1], sDi[BP_BN][BP_KC + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int i = blockIdx.x * BP_BI + wave * 16 + lr;
    const int n0 = blockIdx.y * BP_BN;
    v4 accR[4], accI[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        accR[t] = v4{0, 0, 0, 0};
        accI[t] = v4{0, 0, 0, 0};
    }
    for (int ka = 0; ka < nanl; ka += BP_KC) {
        for (int t = threadIdx.x; t < BP_BN * BP_KC; t += blockDim.x) {
            const int nn = t / BP_KC, aa = t - nn * BP_KC;
            cplx<R> v = mk<R>(0, 0);
            if (n0 + nn < neig && ka + aa < nanl) v = dhd[(long)(n0 + nn) * nanl + ka + aa];
            sDr[nn][aa] = v.x;
            sDi[nn][aa] = v.y;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BP_KC; kk += 4) {
            const int a = ka + kk + lk;
            cplx<R> av = mk<R>(0, 0);
            if (i < mtxd && a < nanl) av = anl[(long)a * lda + i];
#pragma unroll
            for (int ni = 0; ni < 4; ni++) {
                const R dr = sDr[ni * 16 + lr][kk + lk];
                const R di = sDi[ni * 16 + lr][kk + lk];
                // d * a = (dr ar - di ai) + i (dr ai + di ar)
                accR[ni] = MF<R>::mma(dr, av.x, accR[ni]);
                accR[ni] = MF<R>::mma(-di, av.y, accR[ni]);
                accI[ni] = MF<R>::mma(dr, av.y, accI[ni]);
                accI[ni] = MF<R>::mma(di, av.x, accI[ni]);
            }
        }
        __syncthreads();
    }
    if (i < mtxd) {
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int n = n0 + ni * 16 + MF<R>::row(lk, r);
                if (n < neig) {
                    double2 h = hpsi[(long)n * ldh + i];
                    h.x += (double)accR[ni][r];
                    h.y += (double)accI[ni][r];
                    hpsi[(long)n * ldh + i] = h;
                }
            }
    }
}

// FP64 -> FP32 copy of the projectors (FP32-mixed mode)
__global__ void k_to_c64(const double2 *__restrict__ src, long lds, cplx<float> *__restrict__ dst,
                         long ldd, int rows, int cols) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const int c = (int)(t / rows), r = (int)(t - (long)c * rows);
    double2 v = src[(long)c * lds + r];
    dst[(long)c * ldd + r] = mk<float>((float)v.x, (float)v.y);
}

__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (float)src[t];
}

}  // namespace cpwgpu
