// Kleinman-Bylander non-local term on the matrix cores (gfx950 MFMA).
//
// Reference: Src/cpw_lib/hk_psi_nl_c16.f90:78-96
//   dhd  = anlga^H psi                 zgemm('c','n', nanl, neig, mtxd)
//   dhd  = Diag(xnlkb) dhd
//   hpsi = hpsi + anlga dhd            zgemm('n','n', mtxd, neig, nanl)
//
// Complex products are split into real MFMA products
// (v_mfma_f64_16x16x4_f64 for FP64, v_mfma_f32_16x16x4_f32 for FP32, FP32
// accumulation).
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

// Complex products use three real MFMA products (the Gauss / 3M form):
//   conj(a) b :  P1 = ar br, P2 = ai bi, P3 = (ar - ai)(br + bi)
//                re = P1 + P2,  im = P3 - P1 + P2
//   a b       :  P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi)
//                re = P1 - P2,  im = P3 - P1 - P2
// K is permuted inside each LDS stage: in MFMA k-step u, lane group g = l/16
// handles K index g*KPL + u, so a lane reads KPL consecutive K values of its
// row with 128-bit LDS loads.  Rows are padded so that 16 lanes reading 16
// rows hit 16 distinct 16-byte bank groups.

template <typename R> struct GemmCfg;
template <> struct GemmCfg<float> {
    // k_proj: wave tile (WM*16) x (WN*16), block 2x2 waves; BK K per LDS stage
    static constexpr int WM = 4, WN = 4, BK = 16, KP = 20;  // KP: 5 x 16 B rows
    // k_back: KC projectors per LDS chunk
    static constexpr int KC = 64, KCP = 68;                 // 17 x 16 B rows
};
template <> struct GemmCfg<double> {
    static constexpr int WM = 4, WN = 2, BK = 8, KP = 10;   // 5 x 16 B rows
    static constexpr int KC = 32, KCP = 34;                 // 17 x 16 B rows
};

template <typename R> constexpr int proj_bm() { return 2 * GemmCfg<R>::WM * 16; }
template <typename R> constexpr int proj_bn() { return 2 * GemmCfg<R>::WN * 16; }

// ---------------------------------------------------------------------------
// k_proj: part[s][n][a] = sum_{i in slice s} w_i conj(anlga[i,a]) psi[i,n]
// (w_i = 1 without wrow) grid (ceil(nanl/BM), ceil(neig/BN), nslice); 4 waves,
// each WM*16 a x WN*16 n
// (FP32: block 128 x 128, so a K slice of anlga and psi is read at most
// twice).  Double-buffered LDS: one barrier per K stage, the next stage's
// global loads in flight during the current stage's MFMAs.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(GEMM_THREADS) void k_proj(const cplx<R> *__restrict__ anl, long lda,
                                                       const double2 *__restrict__ psi, long ldp,
                                                       cplx<R> *__restrict__ part, int mtxd,
                                                       int nanl, int neig, int kslice,
                                                       const double *__restrict__ wrow = nullptr) {
    typedef typename MF<R>::v4 v4;
    constexpr int WM = GemmCfg<R>::WM, WN = GemmCfg<R>::WN;
    constexpr int BK = GemmCfg<R>::BK, KP = GemmCfg<R>::KP;
    constexpr int BM = proj_bm<R>(), BN = proj_bn<R>();
    constexpr int KPL = BK / 4;                   // K values per lane group per stage
    constexpr int LPA = BM * BK / GEMM_THREADS;   // complex values a thread stages (A)
    constexpr int LPB = BN * BK / GEMM_THREADS;   // (psi)
    __shared__ __align__(16) R sA[2][2][BM][KP];  // [buf][re/im][a][k]
    __shared__ __align__(16) R sB[2][2][BN][KP];  // [buf][re/im][n][k]
    const int a0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kb = blockIdx.z * kslice, ke = min(mtxd, kb + kslice);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 15, lg = lane >> 4;
    v4 P1[WM][WN], P2[WM][WN], P3[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) {
            P1[i][j] = v4{0, 0, 0, 0};
            P2[i][j] = v4{0, 0, 0, 0};
            P3[i][j] = v4{0, 0, 0, 0};
        }
    // staging: thread -> (row, K offset)
    const int arow = threadIdx.x / (BK / LPA), aoff = (threadIdx.x % (BK / LPA)) * LPA;
    const int brow = threadIdx.x / (BK / LPB), boff = (threadIdx.x % (BK / LPB)) * LPB;
    const int ga = a0 + arow, gn = n0 + brow;
    // branch-free loads: clamped (always valid) address, zero selected outside
    const long abase = (long)min(ga, nanl - 1) * lda, bbase = (long)min(gn, neig - 1) * ldp;
    cplx<R> ra[LPA];
    double2 rb[LPB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < LPA; q++) {
            const int i = k0 + aoff + q;
            const cplx<R> v = anl[abase + min(i, mtxd - 1)];
            const bool ok = ga < nanl && i < ke;
            ra[q] = mk<R>(ok ? v.x : (R)0, ok ? v.y : (R)0);
            if (wrow) ra[q] = scale(ra[q], (R)wrow[min(i, mtxd - 1)]);  // per-plane-wave weight
        }
#pragma unroll
        for (int q = 0; q < LPB; q++) {
            const int i = k0 + boff + q;
            const double2 v = psi[bbase + min(i, mtxd - 1)];
            const bool ok = gn < neig && i < ke;
            rb[q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < LPA; q++) {
            sA[buf][0][arow][aoff + q] = ra[q].x;
            sA[buf][1][arow][aoff + q] = ra[q].y;
        }
#pragma unroll
        for (int q = 0; q < LPB; q++) {
            sB[buf][0][brow][boff + q] = (R)rb[q].x;
            sB[buf][1][brow][boff + q] = (R)rb[q].y;
        }
    };
    if (kb < ke) {
        gload(kb);
        sstore(0);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        const bool more = k0 + BK < ke;
        if (more) gload(k0 + BK);
        R ar[WM][KPL], ai[WM][KPL], br[WN][KPL], bi[WN][KPL];
#pragma unroll
        for (int t = 0; t < WM; t++)
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                ar[t][u] = sA[buf][0][wm * WM * 16 + t * 16 + lr][lg * KPL + u];
                ai[t][u] = sA[buf][1][wm * WM * 16 + t * 16 + lr][lg * KPL + u];
            }
#pragma unroll
        for (int t = 0; t < WN; t++)
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                br[t][u] = sB[buf][0][wn * WN * 16 + t * 16 + lr][lg * KPL + u];
                bi[t][u] = sB[buf][1][wn * WN * 16 + t * 16 + lr][lg * KPL + u];
            }
#pragma unroll
        for (int u = 0; u < KPL; u++) {
            R as[WM], bs[WN];
#pragma unroll
            for (int t = 0; t < WM; t++) as[t] = ar[t][u] - ai[t][u];
#pragma unroll
            for (int t = 0; t < WN; t++) bs[t] = br[t][u] + bi[t][u];
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
#pragma unroll
                for (int ni = 0; ni < WN; ni++) {
                    P1[mi][ni] = MF<R>::mma(ar[mi][u], br[ni][u], P1[mi][ni]);
                    P2[mi][ni] = MF<R>::mma(ai[mi][u], bi[ni][u], P2[mi][ni]);
                    P3[mi][ni] = MF<R>::mma(as[mi], bs[ni], P3[mi][ni]);
                }
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    cplx<R> *ps = part + (long)blockIdx.z * neig * nanl;
#pragma unroll
    for (int mi = 0; mi < WM; mi++)
#pragma unroll
        for (int ni = 0; ni < WN; ni++) {
            const int n = n0 + wn * WN * 16 + ni * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int a = a0 + wm * WM * 16 + mi * 16 + MF<R>::row(lg, r);
                const R p1 = P1[mi][ni][r], p2 = P2[mi][ni][r], p3 = P3[mi][ni][r];
                if (a < nanl && n < neig) ps[(long)n * nanl + a] = mk<R>(p1 + p2, p3 - p1 + p2);
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
    const R x = xnl ? (R)xnl[a] : (R)1;  // no scaling for plain Gram matrices
    dhd[t] = mk<R>(x * sr, x * si);
}

constexpr int BP_BI = 64;  // plane waves per block (16 per wave)
constexpr int BP_BN = 64;  // bands per block

// ---------------------------------------------------------------------------
// k_back: hpsi[i,n] = beta*hpsi[i,n] + sum_a anlga[i,a] dhd[a,n], beta 0 or 1
// computed as D^T = dhd^T anlga^T so that each MFMA output row is 16
// consecutive plane waves (coalesced 256-byte stores of hpsi).
// grid (ceil(mtxd/64), ceil(neig/64)); wave w owns plane waves i0+16w..+15
// and the block's 64 bands.  dhd is staged in LDS per chunk of KC projectors
// as three planes (re, im, re+im); projector values and the next dhd chunk
// are loaded into registers while the current chunk is multiplied.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(GEMM_THREADS) void k_back(const cplx<R> *__restrict__ anl, long lda,
                                                       const cplx<R> *__restrict__ dhd,
                                                       double2 *__restrict__ hpsi, long ldh,
                                                       int mtxd, int nanl, int neig, int beta) {
    typedef typename MF<R>::v4 v4;
    constexpr int KC = GemmCfg<R>::KC, KCP = GemmCfg<R>::KCP;
    constexpr int KPL = KC / 4;             // projectors per lane group per chunk
    constexpr int DPT = BP_BN * KC / GEMM_THREADS;  // dhd values a thread stages
    __shared__ __align__(16) R sD[3][BP_BN][KCP];   // re, im, re+im
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int i = blockIdx.x * BP_BI + wave * 16 + lr;
    const int n0 = blockIdx.y * BP_BN;
    v4 P1[4], P2[4], P3[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        P1[t] = v4{0, 0, 0, 0};
        P2[t] = v4{0, 0, 0, 0};
        P3[t] = v4{0, 0, 0, 0};
    }
    // staging of dhd: thread -> band row tid/4, projector offset (tid%4)*DPT
    const int srow = threadIdx.x >> 2, soff = (threadIdx.x & 3) * DPT;
    cplx<R> rd[DPT];
    // branch-free loads: clamped (always valid) address, zero selected outside
    const int nrow = n0 + srow;
    const long dbase = (long)min(nrow, neig - 1) * nanl;
    auto dload = [&](int ka) {
#pragma unroll
        for (int q = 0; q < DPT; q++) {
            const int a = ka + soff + q;
            const cplx<R> v = dhd[dbase + min(a, nanl - 1)];
            const bool ok = nrow < neig && a < nanl;
            rd[q] = mk<R>(ok ? v.x : (R)0, ok ? v.y : (R)0);
        }
    };
    cplx<R> av[KPL];
    const int ic = min(i, mtxd - 1);
    auto aload = [&](int ka) {
#pragma unroll
        for (int u = 0; u < KPL; u++) {
            const int a = ka + lg * KPL + u;
            const cplx<R> v = anl[(long)min(a, nanl - 1) * lda + ic];
            const bool ok = i < mtxd && a < nanl;
            av[u] = mk<R>(ok ? v.x : (R)0, ok ? v.y : (R)0);
        }
    };
    dload(0);
    aload(0);
    for (int ka = 0; ka < nanl; ka += KC) {
#pragma unroll
        for (int q = 0; q < DPT; q++) {
            sD[0][srow][soff + q] = rd[q].x;
            sD[1][srow][soff + q] = rd[q].y;
            sD[2][srow][soff + q] = rd[q].x + rd[q].y;
        }
        __syncthreads();
        R cr[KPL], ci[KPL], cs[KPL];
#pragma unroll
        for (int u = 0; u < KPL; u++) {
            cr[u] = av[u].x;
            ci[u] = av[u].y;
            cs[u] = av[u].x + av[u].y;
        }
        if (ka + KC < nanl) {
            dload(ka + KC);
            aload(ka + KC);
        }
#pragma unroll
        for (int ni = 0; ni < 4; ni++) {
            R dr[KPL], di[KPL], ds[KPL];
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                dr[u] = sD[0][ni * 16 + lr][lg * KPL + u];
                di[u] = sD[1][ni * 16 + lr][lg * KPL + u];
                ds[u] = sD[2][ni * 16 + lr][lg * KPL + u];
            }
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                P1[ni] = MF<R>::mma(dr[u], cr[u], P1[ni]);
                P2[ni] = MF<R>::mma(di[u], ci[u], P2[ni]);
                P3[ni] = MF<R>::mma(ds[u], cs[u], P3[ni]);
            }
        }
        __syncthreads();
    }
    if (i < mtxd) {
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int n = n0 + ni * 16 + MF<R>::row(lg, r);
                if (n < neig) {
                    const R p1 = P1[ni][r], p2 = P2[ni][r], p3 = P3[ni][r];
                    double2 h = make_double2((double)(p1 - p2), (double)(p3 - p1 - p2));
                    if (beta) {
                        const double2 o = hpsi[(long)n * ldh + i];
                        h.x += o.x;
                        h.y += o.y;
                    }
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
