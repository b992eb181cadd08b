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

// Workgroups are dealt to the 8 XCDs round robin on their linear id.
// xcd_linear renumbers the T blocks of a grid so that each XCD runs one
// contiguous range of the new ids, a bijection on [0, T): with
// per = ceil(T/8), XCDs 0 .. full-1 (full = T - 8 (per - 1)) take per ids
// each, the others per - 1.
// Tiles with neighbouring ids share operand rows, which then fill one L2
// instead of eight.
__device__ __forceinline__ int xcd_linear(int L, int T) {
    const int per = (T + 7) >> 3;
    const int full = T - (per - 1) * 8;  // 1..8
    const int x = L & 7, l = L >> 3;
    return x < full ? x * per + l : full * per + (x - full) * (per - 1) + l;
}

// Tile (x, y, z) of this block in a GEMM grid, renumbered by xcd_linear with
// x fastest: the row and band tiles of one K slice run on one XCD.
struct Tile3 {
    int x, y, z;
};
__device__ __forceinline__ Tile3 gemm_tile_xfirst() {
    const int gx = gridDim.x, gy = gridDim.y;
    const int L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int j = xcd_linear(L, gx * gy * (int)gridDim.z);
    const int r = j / gx;
    return {j - r * gx, r % gy, r / gy};
}
// The same for a 2-D grid with y fastest: the band tiles of one block of
// plane waves (which read the same projector rows) run on one XCD.
__device__ __forceinline__ Tile3 gemm_tile_yfirst() {
    const int gx = gridDim.x, gy = gridDim.y;
    const int j = xcd_linear(blockIdx.x + gx * blockIdx.y, gx * gy);
    const int x = j / gy;
    return {x, j - x * gy, 0};
}

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
// WMX x WNX: MFMA tiles per wave (variant 0 uses GemmCfg<R>; half of it in
// each direction keeps fewer accumulators per lane, for more waves per SIMD)
template <typename R, int WMX = GemmCfg<R>::WM, int WNX = GemmCfg<R>::WN>
__global__ __launch_bounds__(GEMM_THREADS) void k_proj(const cplx<R> *__restrict__ anl, long lda,
                                                       const double2 *__restrict__ psi, long ldp,
                                                       cplx<R> *__restrict__ part, int mtxd,
                                                       int nanl, int neig, int kslice,
                                                       const double *__restrict__ wrow = nullptr,
                                                       int lower = 0) {
    typedef typename MF<R>::v4 v4;
    constexpr int WM = WMX, WN = WNX;
    constexpr int BK = GemmCfg<R>::BK, KP = GemmCfg<R>::KP;
    constexpr int BM = 2 * WM * 16, BN = 2 * WN * 16;
    static_assert(BM * BK % GEMM_THREADS == 0 && BN * BK % GEMM_THREADS == 0, "bad k_proj tile");
    constexpr int KPL = BK / 4;                   // K values per lane group per stage
    constexpr int LPA = BM * BK / GEMM_THREADS;   // complex values a thread stages (A)
    constexpr int LPB = BN * BK / GEMM_THREADS;   // (psi)
    __shared__ __align__(16) R sA[2][2][BM][KP];  // [buf][re/im][a][k]
    __shared__ __align__(16) R sB[2][2][BN][KP];  // [buf][re/im][n][k]
    const Tile3 tl = gemm_tile_xfirst();
    const int a0 = tl.x * BM, n0 = tl.y * BN;
    // lower != 0: only the lower triangle (a >= n) of a Hermitian Gram matrix
    // is wanted; tiles wholly above the diagonal are skipped (whole block)
    if (lower && a0 + BM <= n0) return;
    const int kb = tl.z * kslice, ke = min(mtxd, kb + kslice);
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
    // rows beyond the matrix (a >= nanl, n >= neig) are read at clamped
    // addresses and only reach outputs that are not stored; full stages issue
    // plain loads, the slice tail is clamped with psi zero-masked
    const long abase = (long)min(ga, nanl - 1) * lda, bbase = (long)min(gn, neig - 1) * ldp;
    cplx<R> ra[LPA];
    double2 rb[LPB];
    auto gload = [&](int k0) {
        if (k0 + BK <= ke) {  // the same for the whole block
#pragma unroll
            for (int q = 0; q < LPA; q++) ra[q] = anl[abase + k0 + aoff + q];
#pragma unroll
            for (int q = 0; q < LPB; q++) rb[q] = psi[bbase + k0 + boff + q];
        } else {
#pragma unroll
            for (int q = 0; q < LPA; q++) ra[q] = anl[abase + min(k0 + aoff + q, mtxd - 1)];
#pragma unroll
            for (int q = 0; q < LPB; q++) {
                const int i = k0 + boff + q;
                const double2 v = psi[bbase + min(i, mtxd - 1)];
                rb[q] = make_double2(i < ke ? v.x : 0.0, i < ke ? v.y : 0.0);
            }
        }
        if (wrow) {  // per-plane-wave weight on the projector operand
#pragma unroll
            for (int q = 0; q < LPA; q++) ra[q] = scale(ra[q], (R)wrow[min(k0 + aoff + q, mtxd - 1)]);
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
    cplx<R> *ps = part + (long)tl.z * neig * nanl;
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

// ---------------------------------------------------------------------------
// k_proj_direct: the product and partial layout of k_proj, without LDS.  Each
// wave loads its own A (projector rows) and B (psi columns) fragments from
// global memory into registers, one K chunk of 4*KPL ahead of the MFMAs, so
// no barrier ever stalls the matrix cores; the waves of a block that share
// tiles hit the same L1/L2 lines.  Requires anl rows 16-byte aligned (lda
// even for complex64).  Full chunks use 16-byte vector loads; the slice tail
// falls back to clamped element loads with psi zero-masked.
// ---------------------------------------------------------------------------
// Shapes of k_proj_direct (autotuner choices 1..NPROJ_DIRECT): wave tile
// (WM*16) x (WN*16), KPL K values per lane per chunk, register budget for W
// waves per SIMD (accumulators plus two chunks of loads).
// register buffers in the chunk ring of the direct kernels (2: double
// buffering; a third buffer measured no faster on gfx950)
constexpr int DIRECT_D = 2;
constexpr int NPROJ_DIRECT = 4;
template <typename R, int V> struct DirectCfg;
// (all compile without register spills on gfx950)
template <> struct DirectCfg<float, 1> { static constexpr int WM = 4, WN = 2, KPL = 2, W = 2, D = DIRECT_D; };
template <> struct DirectCfg<float, 2> { static constexpr int WM = 4, WN = 2, KPL = 4, W = 1, D = DIRECT_D; };
template <> struct DirectCfg<float, 3> { static constexpr int WM = 2, WN = 2, KPL = 8, W = 1, D = DIRECT_D; };
template <> struct DirectCfg<float, 4> { static constexpr int WM = 2, WN = 2, KPL = 4, W = 2, D = DIRECT_D; };
template <> struct DirectCfg<double, 1> { static constexpr int WM = 2, WN = 2, KPL = 2, W = 2, D = DIRECT_D; };
template <> struct DirectCfg<double, 2> { static constexpr int WM = 2, WN = 2, KPL = 4, W = 1, D = DIRECT_D; };
template <> struct DirectCfg<double, 3> { static constexpr int WM = 4, WN = 1, KPL = 2, W = 2, D = DIRECT_D; };
template <> struct DirectCfg<double, 4> { static constexpr int WM = 2, WN = 1, KPL = 4, W = 2, D = DIRECT_D; };
template <typename R, int V> constexpr int direct_bm() { return 2 * DirectCfg<R, V>::WM * 16; }
template <typename R, int V> constexpr int direct_bn() { return 2 * DirectCfg<R, V>::WN * 16; }
template <typename R, int V> constexpr int direct_bk() { return 4 * DirectCfg<R, V>::KPL; }

// Chunk loads of the direct kernels.  Rows beyond the matrix (projectors
// a >= nanl, bands n >= neig) are read at a clamped, valid address and not
// masked: such a row only reaches output rows or columns that are never
// stored.  Only the K tail is masked, on one operand (psi), so that the other
// operand's finite values there multiply zeros.  A full chunk thus issues
// plain loads whose results are first used by the MFMAs one chunk later,
// and the load latency stays hidden behind the current chunk.
template <typename R, int KPL>
__device__ __forceinline__ void load_row_full(cplx<R> (&dst)[KPL], const cplx<R> *__restrict__ row, int i0) {
    if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int q = 0; q < KPL / 2; q++) {
            const float4 f = *reinterpret_cast<const float4 *>(row + i0 + 2 * q);
            dst[2 * q] = mk<R>(f.x, f.y);
            dst[2 * q + 1] = mk<R>(f.z, f.w);
        }
    } else {
#pragma unroll
        for (int q = 0; q < KPL; q++) dst[q] = row[i0 + q];
    }
}

template <typename R, int KPL>
__device__ __forceinline__ void load_row_tail(cplx<R> (&dst)[KPL], const cplx<R> *__restrict__ row, int i0,
                                              int mtxd) {
#pragma unroll
    for (int q = 0; q < KPL; q++) dst[q] = row[min(i0 + q, mtxd - 1)];
}

template <int KPL>
__device__ __forceinline__ void load_psi_full(double2 (&dst)[KPL], const double2 *__restrict__ col, int i0) {
#pragma unroll
    for (int q = 0; q < KPL; q++) dst[q] = col[i0 + q];
}

template <int KPL>
__device__ __forceinline__ void load_psi_tail(double2 (&dst)[KPL], const double2 *__restrict__ col, int i0,
                                              int ke, int mtxd) {
#pragma unroll
    for (int q = 0; q < KPL; q++) {
        const int i = i0 + q;
        const double2 v = col[min(i, mtxd - 1)];
        const bool ok = i < ke;
        dst[q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
}

template <typename R, int V>
__global__ __launch_bounds__(GEMM_THREADS, (DirectCfg<R, V>::W)) void k_proj_direct(
    const cplx<R> *__restrict__ anl, long lda, const double2 *__restrict__ psi, long ldp,
    cplx<R> *__restrict__ part, int mtxd, int nanl, int neig, int kslice,
    const double *__restrict__ wrow = nullptr, int lower = 0) {
    typedef typename MF<R>::v4 v4;
    constexpr int WM = DirectCfg<R, V>::WM, WN = DirectCfg<R, V>::WN;
    constexpr int KPL = DirectCfg<R, V>::KPL, KC = 4 * KPL;
    constexpr int BM = direct_bm<R, V>(), BN = direct_bn<R, V>();
    const Tile3 tl = gemm_tile_xfirst();
    const int a0 = tl.x * BM, n0 = tl.y * BN;
    // lower != 0: only the lower triangle (a >= n) of a Hermitian Gram matrix;
    // tiles wholly above the diagonal are skipped (the reduction zeroes them)
    if (lower && a0 + BM <= n0) return;
    const int kb = tl.z * kslice, ke = min(mtxd, kb + kslice);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 15, lg = lane >> 4;
    const cplx<R> *arow[WM];
#pragma unroll
    for (int t = 0; t < WM; t++) {
        const int a = a0 + wm * WM * 16 + t * 16 + lr;
        arow[t] = anl + (long)min(a, nanl - 1) * lda;
    }
    const double2 *bcol[WN];
#pragma unroll
    for (int t = 0; t < WN; t++) {
        const int n = n0 + wn * WN * 16 + t * 16 + lr;
        bcol[t] = psi + (long)min(n, neig - 1) * ldp;
    }
    v4 P1[WM][WN], P2[WM][WN], P3[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) {
            P1[i][j] = v4{0, 0, 0, 0};
            P2[i][j] = v4{0, 0, 0, 0};
            P3[i][j] = v4{0, 0, 0, 0};
        }
    constexpr int D = DirectCfg<R, V>::D;
    cplx<R> ra[D][WM][KPL];
    double2 rb[D][WN][KPL];
    // tail: the last, partial chunk of the slice
    auto load = [&](cplx<R>(&ra_)[WM][KPL], double2(&rb_)[WN][KPL], int k0, bool tail) {
        const int i0 = k0 + lg * KPL;
        if (!tail) {
#pragma unroll
            for (int t = 0; t < WM; t++) load_row_full<R, KPL>(ra_[t], arow[t], i0);
#pragma unroll
            for (int t = 0; t < WN; t++) load_psi_full<KPL>(rb_[t], bcol[t], i0);
        } else {
#pragma unroll
            for (int t = 0; t < WM; t++) load_row_tail<R, KPL>(ra_[t], arow[t], i0, mtxd);
#pragma unroll
            for (int t = 0; t < WN; t++) load_psi_tail<KPL>(rb_[t], bcol[t], i0, ke, mtxd);
        }
        if (wrow) {
            // per-plane-wave weight w_i on the projector operand (zeros stay zero)
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                const R w = (R)wrow[min(i0 + u, mtxd - 1)];
#pragma unroll
                for (int t = 0; t < WM; t++) ra_[t][u] = scale(ra_[t][u], w);
            }
        }
    };
    auto compute = [&](const cplx<R>(&ra_)[WM][KPL], const double2(&rb_)[WN][KPL]) {
#pragma unroll
        for (int u = 0; u < KPL; u++) {
            R ar[WM], ai[WM], as[WM], br[WN], bi[WN], bs[WN];
#pragma unroll
            for (int t = 0; t < WM; t++) {
                ar[t] = ra_[t][u].x;
                ai[t] = ra_[t][u].y;
                as[t] = ar[t] - ai[t];
            }
#pragma unroll
            for (int t = 0; t < WN; t++) {
                br[t] = (R)rb_[t][u].x;
                bi[t] = (R)rb_[t][u].y;
                bs[t] = br[t] + bi[t];
            }
#pragma unroll
            for (int mi = 0; mi < WM; mi++)
#pragma unroll
                for (int ni = 0; ni < WN; ni++) {
                    P1[mi][ni] = MF<R>::mma(ar[mi], br[ni], P1[mi][ni]);
                    P2[mi][ni] = MF<R>::mma(ai[mi], bi[ni], P2[mi][ni]);
                    P3[mi][ni] = MF<R>::mma(as[mi], bs[ni], P3[mi][ni]);
                }
        }
    };
    // Full chunks through a ring of D register buffers, D per trip: chunk
    // c + D - 1 is loaded while chunk c is multiplied, unconditionally (at
    // the last full chunk again when there is none: a valid address whose
    // values go unused), so that the loop body is one straight block and each
    // chunk's loads have D - 1 chunks of MFMAs to arrive.  Then the partial
    // chunk, if any.  Chunks are accumulated in K order for any D.
    const int nfull = max(0, ke - kb) / KC;
    if (nfull > 0) {
#pragma unroll
        for (int j = 0; j < D - 1; j++) load(ra[j], rb[j], kb + min(j, nfull - 1) * KC, false);
        int c = 0;
        for (; c + D <= nfull; c += D) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                load(ra[(j + D - 1) % D], rb[(j + D - 1) % D], kb + min(c + j + D - 1, nfull - 1) * KC, false);
                compute(ra[j], rb[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < D - 1; j++)
            if (c + j < nfull) compute(ra[j], rb[j]);
    }
    if (kb + nfull * KC < ke) {
        load(ra[0], rb[0], kb + nfull * KC, true);
        compute(ra[0], rb[0]);
    }
    cplx<R> *ps = part + (long)tl.z * neig * nanl;
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
// Block of 256 threads = 64 consecutive outputs x 4 slice groups; group g sums
// slices g, g+4, ... into four independent partial sums (16 loads in flight
// per thread), then the groups are combined through LDS in group order.
constexpr int RED_OUT = 64, RED_GRP = 4;
template <typename R>
__global__ __launch_bounds__(RED_OUT * RED_GRP) void k_proj_reduce(const cplx<R> *__restrict__ part,
                                                                  const double *__restrict__ xnl,
                                                                  cplx<R> *__restrict__ dhd,
                                                                  int nanl, int neig, int nslice,
                                                                  int lower = 0) {
    __shared__ cplx<R> red[RED_GRP][RED_OUT];
    const int o = threadIdx.x % RED_OUT, grp = threadIdx.x / RED_OUT;
    const long tot = (long)nanl * neig;
    const long t = (long)blockIdx.x * RED_OUT + o;
    const long tc = t < tot ? t : tot - 1;  // clamped, always a valid address
    R sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
    int s = grp;
    for (; s + 3 * RED_GRP < nslice; s += 4 * RED_GRP) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const cplx<R> v = part[(long)(s + u * RED_GRP) * tot + tc];
            sr[u] += v.x;
            si[u] += v.y;
        }
    }
    for (; s < nslice; s += RED_GRP) {
        const cplx<R> v = part[(long)s * tot + tc];
        sr[0] += v.x;
        si[0] += v.y;
    }
    red[grp][o] = mk<R>((sr[0] + sr[1]) + (sr[2] + sr[3]), (si[0] + si[1]) + (si[2] + si[3]));
    __syncthreads();
    if (grp == 0 && t < tot) {
        R r = 0, i = 0;
#pragma unroll
        for (int g2 = 0; g2 < RED_GRP; g2++) {
            r += red[g2][o].x;
            i += red[g2][o].y;
        }
        const R x = xnl ? (R)xnl[t % nanl] : (R)1;  // no scaling for plain Gram matrices
        // lower: the strict upper triangle (partials of skipped tiles) is zero
        const bool up = lower && (t % nanl) < (t / nanl);
        dhd[t] = up ? mk<R>((R)0, (R)0) : mk<R>(x * r, x * i);
    }
}

// Launches k_proj_reduce over nanl * neig outputs (nothing to do when empty).
template <typename R>
inline void launch_proj_reduce(const cplx<R> *part, const double *xnl, cplx<R> *dhd, int nanl,
                               int neig, int nslice, hipStream_t s, int lower = 0) {
    const long tot = (long)nanl * neig;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_proj_reduce<R>, dim3((unsigned)((tot + RED_OUT - 1) / RED_OUT)),
                       dim3(RED_OUT * RED_GRP), 0, s, part, xnl, dhd, nanl, neig, nslice, lower);
}

constexpr int BP_BI = 64;  // plane waves per block (16 per wave)
constexpr int BP_BN = 64;  // bands per block

// ---------------------------------------------------------------------------
// k_back: hpsi[i,n] = beta*hpsi[i,n] + sum_a anlga[i,a] dhd[a,n]
// (+ ekpg[i] kpsi[i,n] when kpsi is given: the kinetic term), beta 0 or 1
// computed as D^T = dhd^T anlga^T so that each MFMA output row is 16
// consecutive plane waves (coalesced 256-byte stores of hpsi).
// grid (ceil(mtxd/64), ceil(neig/64)); wave w owns plane waves i0+16w..+15
// and the block's 64 bands.  dhd is staged in LDS per chunk of KC projectors
// as three planes (re, im, re+im); projector values and the next dhd chunk
// are loaded into registers while the current chunk is multiplied.
// ---------------------------------------------------------------------------
// KCX: projectors per LDS chunk (variant 0 uses GemmCfg<R>::KC; half of it
// keeps fewer values in flight per lane, for more waves per SIMD)
template <typename R, int KCX = GemmCfg<R>::KC>
__global__ __launch_bounds__(GEMM_THREADS) void k_back(const cplx<R> *__restrict__ anl, long lda,
                                                       const cplx<R> *__restrict__ dhd,
                                                       double2 *__restrict__ hpsi, long ldh,
                                                       int mtxd, int nanl, int neig, int beta,
                                                       const double2 *__restrict__ kpsi = nullptr,
                                                       long ldk = 0,
                                                       const double *__restrict__ ekpg = nullptr) {
    typedef typename MF<R>::v4 v4;
    // LDS rows padded by 16 bytes: an odd number of 16-byte units per row
    constexpr int KC = KCX, KCP = KCX + 16 / (int)sizeof(R);
    static_assert((KCP * sizeof(R) / 16) % 2 == 1 && KC % 16 == 0, "bad k_back chunk");
    constexpr int KPL = KC / 4;             // projectors per lane group per chunk
    constexpr int DPT = BP_BN * KC / GEMM_THREADS;  // dhd values a thread stages
    __shared__ __align__(16) R sD[3][BP_BN][KCP];   // re, im, re+im
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const Tile3 tl = gemm_tile_yfirst();
    const int i = tl.x * BP_BI + wave * 16 + lr;
    const int n0 = tl.y * BP_BN;
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
    // rows beyond the matrix (bands n >= neig, plane waves i >= mtxd) are
    // read at clamped addresses and only reach outputs that are not stored;
    // full chunks issue plain loads (first used after the current chunk's
    // MFMAs), the projector tail is clamped with dhd zero-masked
    const long dbase = (long)min(n0 + srow, neig - 1) * nanl;
    auto dload = [&](int ka) {
        if (ka + KC <= nanl) {  // the same for the whole block
#pragma unroll
            for (int q = 0; q < DPT; q++) rd[q] = dhd[dbase + ka + soff + q];
        } else {
#pragma unroll
            for (int q = 0; q < DPT; q++) {
                const int a = ka + soff + q;
                const cplx<R> v = dhd[dbase + min(a, nanl - 1)];
                rd[q] = mk<R>(a < nanl ? v.x : (R)0, a < nanl ? v.y : (R)0);
            }
        }
    };
    cplx<R> av[KPL];
    const cplx<R> *acol = anl + min(i, mtxd - 1);
    auto aload = [&](int ka) {
        if (ka + KC <= nanl) {
#pragma unroll
            for (int u = 0; u < KPL; u++) av[u] = acol[(long)(ka + lg * KPL + u) * lda];
        } else {
#pragma unroll
            for (int u = 0; u < KPL; u++) av[u] = acol[(long)min(ka + lg * KPL + u, nanl - 1) * lda];
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
        // the kinetic term in FP64 (coalesced rows of psi) and the term already
        // in hpsi: the loads of four outputs issued together (at clamped rows:
        // a row beyond neig is loaded and not used), then their stores
        const double ek = kpsi ? ekpg[i] : 0.0;
#pragma unroll
        for (int ni = 0; ni < 4; ni++) {
            double2 kp[4], ho[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long nc = min(n0 + ni * 16 + MF<R>::row(lg, r), neig - 1);
                kp[r] = kpsi ? kpsi[nc * ldk + i] : make_double2(0.0, 0.0);
                ho[r] = beta ? hpsi[nc * ldh + i] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int n = n0 + ni * 16 + MF<R>::row(lg, r);
                if (n < neig) {
                    const R p1 = P1[ni][r], p2 = P2[ni][r], p3 = P3[ni][r];
                    hpsi[(long)n * ldh + i] = make_double2((double)(p1 - p2) + ek * kp[r].x + ho[r].x,
                                                           (double)(p3 - p1 - p2) + ek * kp[r].y + ho[r].y);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_back_direct: the product of k_back without LDS.  Each wave owns 16 plane
// waves and 64 bands (D^T = dhd^T anlga^T as in k_back, so an MFMA output row
// is 16 consecutive plane waves); it loads its dhd fragments (small, L2
// resident) and its projector values straight into registers, one chunk of
// 4*BKPL projectors ahead of the MFMAs.  No barrier.
// ---------------------------------------------------------------------------
// Shapes of k_back_direct (autotuner choices 1..NBACK_DIRECT): a wave owns NT
// tiles of 16 bands and IT tiles of 16 plane waves, loads KPL projectors per
// lane per chunk, register budget for W waves per SIMD.  A block is 4 waves
// along the plane waves: 64*IT plane waves x 16*NT bands.
constexpr int NBACK_DIRECT = 3;
template <typename R, int V> struct BackDirectCfg;
// (all compile without register spills on gfx950)
template <> struct BackDirectCfg<float, 1> { static constexpr int NT = 4, IT = 1, KPL = 4, W = 2, D = DIRECT_D; };
template <> struct BackDirectCfg<float, 2> { static constexpr int NT = 4, IT = 2, KPL = 4, W = 2, D = DIRECT_D; };
template <> struct BackDirectCfg<float, 3> { static constexpr int NT = 2, IT = 2, KPL = 4, W = 3, D = DIRECT_D; };
template <> struct BackDirectCfg<double, 1> { static constexpr int NT = 4, IT = 1, KPL = 2, W = 2, D = DIRECT_D; };
template <> struct BackDirectCfg<double, 2> { static constexpr int NT = 2, IT = 2, KPL = 2, W = 2, D = DIRECT_D; };
template <> struct BackDirectCfg<double, 3> { static constexpr int NT = 2, IT = 1, KPL = 2, W = 3, D = DIRECT_D; };
template <typename R, int V> constexpr int back_bi() { return 4 * 16 * BackDirectCfg<R, V>::IT; }
template <typename R, int V> constexpr int back_bn() { return 16 * BackDirectCfg<R, V>::NT; }

template <typename R, int V>
__global__ __launch_bounds__(GEMM_THREADS, (BackDirectCfg<R, V>::W)) void k_back_direct(
    const cplx<R> *__restrict__ anl, long lda, const cplx<R> *__restrict__ dhd,
    double2 *__restrict__ hpsi, long ldh, int mtxd, int nanl, int neig, int beta,
    const double2 *__restrict__ kpsi, long ldk, const double *__restrict__ ekpg) {
    typedef typename MF<R>::v4 v4;
    constexpr int NT = BackDirectCfg<R, V>::NT, IT = BackDirectCfg<R, V>::IT;
    constexpr int KPL = BackDirectCfg<R, V>::KPL, KC = 4 * KPL;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const Tile3 tl = gemm_tile_yfirst();
    const int i0 = tl.x * back_bi<R, V>() + wave * 16 * IT;
    const int n0 = tl.y * back_bn<R, V>();
    // projector values of this lane's plane waves (column i of anlga^T) and
    // dhd rows (bands) of its NT row tiles, at clamped addresses: a clamped
    // row only reaches outputs that are not stored (see load_row_full)
    bool iok[IT];
    const cplx<R> *acol[IT];
#pragma unroll
    for (int q = 0; q < IT; q++) {
        const int i = i0 + q * 16 + lr;
        iok[q] = i < mtxd;
        acol[q] = anl + min(i, mtxd - 1);
    }
    const cplx<R> *drow[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) drow[t] = dhd + (long)min(n0 + t * 16 + lr, neig - 1) * nanl;
    v4 P1[NT][IT], P2[NT][IT], P3[NT][IT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int q = 0; q < IT; q++) {
            P1[t][q] = v4{0, 0, 0, 0};
            P2[t][q] = v4{0, 0, 0, 0};
            P3[t][q] = v4{0, 0, 0, 0};
        }
    constexpr int D = BackDirectCfg<R, V>::D;
    cplx<R> dv[D][NT][KPL], av[D][IT][KPL];
    // full chunks: plain loads; the projector tail: clamped, dhd zero-masked
    auto load = [&](cplx<R>(&d)[NT][KPL], cplx<R>(&a_)[IT][KPL], int ka, bool tail) {
        if (!tail) {
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                const int a = ka + lg * KPL + u;
#pragma unroll
                for (int t = 0; t < NT; t++) d[t][u] = drow[t][a];
#pragma unroll
                for (int q = 0; q < IT; q++) a_[q][u] = acol[q][(long)a * lda];
            }
        } else {
#pragma unroll
            for (int u = 0; u < KPL; u++) {
                const int a = ka + lg * KPL + u;
                const int acl = min(a, nanl - 1);
                const bool aok = a < nanl;
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const cplx<R> v = drow[t][acl];
                    d[t][u] = mk<R>(aok ? v.x : (R)0, aok ? v.y : (R)0);
                }
#pragma unroll
                for (int q = 0; q < IT; q++) a_[q][u] = acol[q][(long)acl * lda];
            }
        }
    };
    auto compute = [&](const cplx<R>(&d)[NT][KPL], const cplx<R>(&a_)[IT][KPL]) {
#pragma unroll
        for (int u = 0; u < KPL; u++) {
            R cr[IT], ci[IT], cs[IT];
#pragma unroll
            for (int q = 0; q < IT; q++) {
                cr[q] = a_[q][u].x;
                ci[q] = a_[q][u].y;
                cs[q] = a_[q][u].x + a_[q][u].y;
            }
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const R ds = d[t][u].x + d[t][u].y;
#pragma unroll
                for (int q = 0; q < IT; q++) {
                    P1[t][q] = MF<R>::mma(d[t][u].x, cr[q], P1[t][q]);
                    P2[t][q] = MF<R>::mma(d[t][u].y, ci[q], P2[t][q]);
                    P3[t][q] = MF<R>::mma(ds, cs[q], P3[t][q]);
                }
            }
        }
    };
    // full chunks through a ring of D register buffers with unconditional
    // (clamped) prefetch, as in k_proj_direct, then the partial chunk
    const int nfull = nanl / KC;
    if (nfull > 0) {
#pragma unroll
        for (int j = 0; j < D - 1; j++) load(dv[j], av[j], min(j, nfull - 1) * KC, false);
        int c = 0;
        for (; c + D <= nfull; c += D) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                load(dv[(j + D - 1) % D], av[(j + D - 1) % D], min(c + j + D - 1, nfull - 1) * KC, false);
                compute(dv[j], av[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < D - 1; j++)
            if (c + j < nfull) compute(dv[j], av[j]);
    }
    if (nfull * KC < nanl) {
        load(dv[0], av[0], nfull * KC, true);
        compute(dv[0], av[0]);
    }
#pragma unroll
    for (int q = 0; q < IT; q++) {
        if (!iok[q]) continue;
        const int i = i0 + q * 16 + lr;
        // as in k_back: the loads of four outputs issued together, then the stores
        const double ek = kpsi ? ekpg[i] : 0.0;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double2 kp[4], ho[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long nc = min(n0 + t * 16 + MF<R>::row(lg, r), neig - 1);
                kp[r] = kpsi ? kpsi[nc * ldk + i] : make_double2(0.0, 0.0);
                ho[r] = beta ? hpsi[nc * ldh + i] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int n = n0 + t * 16 + MF<R>::row(lg, r);
                if (n < neig) {
                    const R p1 = P1[t][q][r], p2 = P2[t][q][r], p3 = P3[t][q][r];
                    hpsi[(long)n * ldh + i] = make_double2((double)(p1 - p2) + ek * kp[r].x + ho[r].x,
                                                           (double)(p3 - p1 - p2) + ek * kp[r].y + ho[r].y);
                }
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

// complex64 (rows, cols; ld lds) -> complex128 (ld ldd): the inputs of the
// single-precision drop-in
__global__ void k_to_c128(const cplx<float> *__restrict__ src, long lds, double2 *__restrict__ dst, long ldd,
                          int rows, int cols) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const int c = (int)(t / rows), r = (int)(t - (long)c * rows);
    const cplx<float> v = src[(long)c * lds + r];
    dst[(long)c * ldd + r] = make_double2((double)v.x, (double)v.y);
}

__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (float)src[t];
}

}  // namespace cpwgpu
