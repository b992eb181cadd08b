// cpwgpu: MI355X-native H|psi> for cpw2000 -- host side and C ABI.
//
// Drop-in for Src/hk_psi/reference/hk_psi_c16.f90 (subroutine hk_psi_c16).
// The per-k set-up that the reference redoes on every call (index map
// ipoint, :141-172) is kept in a plan; the band loop (:176-227) becomes three
// band-batched FFT kernels (fft_local.hpp) and the non-local term
// (Src/cpw_lib/hk_psi_nl_c16.f90) two MFMA GEMMs (nl_gemm.hpp).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpwgpu.h"
#include "fft_fast.hpp"
#include "fft_local.hpp"
#include "fft_reg.hpp"
#include "fft_wave.hpp"
#include "fft_col_reg.hpp"
#include "nl_gemm.hpp"
#include "subspace.hpp"
#include "forces.hpp"
#include "ditsp.hpp"
#include "spin.hpp"
#include "proj_kb.hpp"

using namespace cpwgpu;

namespace {

thread_local std::string g_err;
std::mutex g_mu;
int g_device = 0;
int g_prec = CPWGPU_PREC_FP64;
int g_ditsp_method = CPWGPU_DITSP_CHEB;
int g_overlap = 1;
hipStream_t g_stream = nullptr;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(e_ == hipErrorOutOfMemory ? CPWGPU_ERR_NOMEM : CPWGPU_ERR_HIP,     \
                        "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);  \
    } while (0)

int ensure_stream() {
    if (!g_stream) {
        HIPCHK(hipSetDevice(g_device));
        HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    }
    return CPWGPU_OK;
}

// grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    int reserve(size_t bytes) {
        if (bytes <= n) return CPWGPU_OK;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
        n = bytes;
        return CPWGPU_OK;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Src/ALFA/size_fft.f90:53-76
void size_fft(const int *km, int *ns, int *mfft, int *mwrk) {
    static const int inprim[11] = {2, 4, 6, 8, 10, 12, 16, 18, 20, 24, 32};
    static const int insec[6] = {36, 40, 48, 50, 54, 64};
    for (int n = 0; n < 3; n++) {
        int m = 2 * km[n] + 1;
        if (m < 32) {
            for (int j = 0; j < 11; j++) {
                ns[n] = inprim[j];
                if (m <= ns[n]) break;
            }
        } else {
            int k = (int)(logf((float)m) / logf(2.0f));
            for (int j = 0; j < 6; j++) {
                ns[n] = (1 << (k - 5)) * insec[j];
                if (m <= ns[n]) break;
            }
        }
    }
    *mfft = (ns[0] + 1) * ns[1] * ns[2];
    *mwrk = 4 * std::max({ns[0] * ns[1], ns[0] * ns[2], ns[1] * ns[2]});
}

// radices 4..4, 2, 3.., 5..; false if n has another prime factor
bool radices(int n, int *rad, int *nrad) {
    int k = 0;
    while (n % 4 == 0 && k < MAXRAD) { rad[k++] = 4; n /= 4; }
    while (n % 2 == 0 && k < MAXRAD) { rad[k++] = 2; n /= 2; }
    while (n % 3 == 0 && k < MAXRAD) { rad[k++] = 3; n /= 3; }
    while (n % 5 == 0 && k < MAXRAD) { rad[k++] = 5; n /= 5; }
    *nrad = k;
    return n == 1;
}

int min_radix(const int *rad, int nrad) {
    int r = 5;
    for (int s = 0; s < nrad; s++) r = std::min(r, rad[s]);
    return r;
}

// ---- per-kernel event timing (measurement hook) ----------------------------
struct Prof {
    bool on = false;
    unsigned mask = 0;
    struct Rec { int k; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        // timing only (read after the whole stream has finished): no
        // system-scope fence, so recording them does not write back and
        // invalidate the caches in the middle of the work they time
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        return e;
    }
} g_prof;

// Launches kernel f on stream s.  With the measurement hook on for kernel id
// k, the start and stop events travel with the dispatch itself (the kernel's
// own timestamps, hipExtLaunchKernelGGL) instead of as marker packets
// recorded around it, which would idle the stream between the kernels.
template <typename... Args, typename F = void (*)(Args...)>
void launch_timed(int k, F f, dim3 grid, dim3 block, size_t shmem, hipStream_t s, Args... args) {
    if (g_prof.on && (g_prof.mask >> k & 1u)) {
        hipEvent_t a = g_prof.get(), b = a ? g_prof.get() : nullptr;
        if (a && b) {
            hipExtLaunchKernelGGL(f, grid, block, (std::uint32_t)shmem, s, a, b, 0u, args...);
            g_prof.recs.push_back({k, a, b});
            return;
        }
        if (a) g_prof.pool.push_back(a);
    }
    hipLaunchKernelGGL(f, grid, block, shmem, s, args...);
}

constexpr int COL_CT = 16;   // z-columns per block of the column kernels
constexpr int KMAX = 8;      // butterflies per thread per FFT stage

}  // namespace

struct cpwgpu_plan_s {
    FftGeom g{};
    int mtxd = 0, ng = 0, nanl = 0;
    int kdmax[3] = {0, 0, 0};
    double inv_n = 1.0;
    size_t ntot = 0;          // id*n2*n3
    int mwrk = 0;
    // host copies used to recognise an unchanged k-point in the drop-in
    std::vector<long> key_pts;
    std::vector<double> h_ekpg;
    // device arrays
    DevBuf col_ptr, col_uy, ent_i, ent_cz, ekpg, tw[3];
    DevBuf zmap;              // [col][z] -> psi row or -1 (register column kernels)
    DevBuf v64, v32;          // V(r) in FP64 and FP32
    double v0 = 0.0;          // mean of V(r) over the mesh = V(G=0), the local part of diag(H)
    DevBuf anl64, anl32, xnl; // projectors
    long lda = 0;             // leading dimension of anl64
    long lda32 = 0;           // of anl32: lda rounded up to even (16-byte aligned rows)
    bool anl32_valid = false;
    // plane-stage kernel per precision (index 0 FP64, 1 FP32): -1 = default for
    // the mesh, else a variant number (set by the tuner or by the caller)
    int plane_var[2] = {-1, -1};
    // column kernels: 0 LDS lines, 1 register lines; -1 default
    int col_var[2] = {-1, -1};
    // stream schedule of the two-stream apply: 0 (default) the whole
    // non-local term beside the local one, the gather adding to it; 1 the
    // projection beside the local term, the back-projection after the gather
    int sched_var[2] = {-1, -1};
    // split-K target blocks of the projection GEMM per precision (-1 = 256)
    int splitk[2] = {-1, -1};
    // projection GEMM kernel per precision: 0 LDS-staged k_proj, 1 k_proj_direct
    // (-1 = default)
    int proj_var[2] = {-1, -1};
    // back-projection GEMM kernel per precision: 0 LDS-staged k_back, 1
    // k_back_direct (-1 = default)
    int back_var[2] = {-1, -1};
    // where the kinetic term ekpg*psi is added when the non-local term is on:
    // 0 the back-projection epilogue, 1 the gather of the forward column kernel
    // (-1 = default, 0)
    int kin_var[2] = {-1, -1};
    // plane kernel of the coupled two-component case (nsp = 4): 0 LDS
    // two-pass, 1 register lines (-1 = default)
    int spin_var[2] = {-1, -1};
    // workspaces
    DevBuf colbuf, part, dhd;
    DevBuf gw;                // g_j of each plane wave as doubles [3][mtxd] (force projections)
    // spinor path (hk_psi_spin_c16): potential components in the layout of
    // k_spin_potential, projectors on the 2*mtxd interleaved rows, split bands
    int nsp = 0;
    DevBuf vsp64, vsp32;
    DevBuf sanl64, sanl32, sxnl;
    long slda = 0, slda32 = 0;
    int snanl = 0;
    bool sanl32_valid = false;
    DevBuf spsi, shpsi;
    ~cpwgpu_plan_s() {
        col_ptr.release(); col_uy.release(); ent_i.release(); ent_cz.release(); ekpg.release();
        for (auto &t : tw) t.release();
        v64.release(); v32.release(); anl64.release(); anl32.release(); xnl.release();
        colbuf.release(); part.release(); dhd.release(); gw.release();
        vsp64.release(); vsp32.release(); sanl64.release(); sanl32.release(); sxnl.release();
        spsi.release(); shpsi.release();
    }
};

namespace {

// lines of the G list of each plane wave, wrapped into the mesh
// (Src/hk_psi/reference/hk_psi_c16.f90:143-158)
int plane_wave_points(int mtxd, int ng, const int *kgv, const int *isort, const int *ns,
                      std::vector<int> &kx, std::vector<int> &ky, std::vector<int> &kz, int *kd) {
    kd[0] = kd[1] = kd[2] = 0;
    kx.resize(mtxd); ky.resize(mtxd); kz.resize(mtxd);
    for (int i = 0; i < mtxd; i++) {
        const int it = isort[i] - 1;
        if (it < 0 || it >= ng) return fail(CPWGPU_ERR_ARG, "isort(%d) = %d outside 1..ng=%d", i + 1, it + 1, ng);
        int k[3];
        for (int j = 0; j < 3; j++) {
            k[j] = kgv[3L * it + j];
            kd[j] = std::max(kd[j], std::abs(k[j]));
            if (k[j] < 0) k[j] += ns[j];
        }
        kx[i] = k[0]; ky[i] = k[1]; kz[i] = k[2];
    }
    return CPWGPU_OK;
}

template <typename T> int upload(DevBuf &b, const T *h, size_t n) {
    int st = b.reserve(sizeof(T) * std::max<size_t>(n, 1));
    if (st) return st;
    if (n) HIPCHK(hipMemcpyAsync(b.p, h, sizeof(T) * n, hipMemcpyHostToDevice, g_stream));
    return CPWGPU_OK;
}

int build_plan(cpwgpu_plan_s *p, int mtxd, int ng, const int *kgv, const double *ekpg,
               const int *isort, const int *kmscr, int mxddim, int mxdgve) {
    if (ng > mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", ng, mxdgve);
    if (mtxd < 0 || mtxd > mxddim) return fail(CPWGPU_ERR_MTXD, "mtxd = %d, mxddim = %d", mtxd, mxddim);
    int ns[3] = {kmscr[3], kmscr[4], kmscr[5]};
    const int id = kmscr[6];
    for (int j = 0; j < 3; j++)
        if (ns[j] < 2 || ns[j] > 4095)
            return fail(CPWGPU_ERR_SIZE, "FFT mesh %d %d %d outside 2..4095", ns[0], ns[1], ns[2]);
    if (id < ns[0]) return fail(CPWGPU_ERR_SIZE, "kmscr(7) = %d < n1 = %d", id, ns[0]);
    FftGeom &g = p->g;
    for (int j = 0; j < 3; j++) {
        g.n[j] = ns[j];
        if (!radices(ns[j], g.rad[j], &g.nrad[j]))
            return fail(CPWGPU_ERR_SIZE, "FFT length %d is not of the form 2^a 3^b 5^c", ns[j]);
    }
    g.id = id;
    int nsz[3], mfft;
    size_fft(kmscr, nsz, &mfft, &p->mwrk);
    std::vector<int> kx, ky, kz;
    int kd[3];
    int st = plane_wave_points(mtxd, ng, kgv, isort, ns, kx, ky, kz, kd);
    if (st) return st;
    // no wrap-around (Src/hk_psi/reference/hk_psi_c16.f90:162-172)
    if (kd[0] > (ns[0] - 1) / 2 || kd[1] > (ns[1] - 1) / 2 || kd[2] > (ns[2] - 1) / 2)
        return fail(CPWGPU_ERR_WRAP, "dimension of k index %d %d %d exceeds %d %d %d", kd[0], kd[1],
                    kd[2], (ns[0] - 1) / 2, (ns[1] - 1) / 2, (ns[2] - 1) / 2);
    for (int j = 0; j < 3; j++) { g.kd[j] = kd[j]; p->kdmax[j] = kd[j]; }
    g.nx = 2 * kd[0] + 1;
    g.mtxd = mtxd;
    p->mtxd = mtxd;
    p->ng = ng;
    p->ntot = (size_t)id * ns[1] * ns[2];
    p->inv_n = 1.0 / ((double)ns[0] * ns[1] * ns[2]);

    // columns = distinct (x, y), numbered in (y, x) order; CSR of plane waves
    const long nxy = (long)ns[0] * ns[1];
    std::vector<int> colid(nxy, -1), count;
    for (int i = 0; i < mtxd; i++) colid[(long)ky[i] * ns[0] + kx[i]] = 0;
    int ncol = 0;
    std::vector<int> uy;
    for (long q = 0; q < nxy; q++)
        if (colid[q] == 0) {
            colid[q] = ncol++;
            const int x = (int)(q % ns[0]), y = (int)(q / ns[0]);
            const int u = x <= kd[0] ? x : x - (ns[0] - g.nx);
            uy.push_back((y << 16) | u);
        }
    if (ncol >= (1 << 19)) return fail(CPWGPU_ERR_SIZE, "too many z-columns (%d)", ncol);
    g.ncol = ncol;
    g.ncolp = ((std::max(ncol, 1) + 31) / 32) * 32;
    std::vector<int> cptr(ncol + 1, 0), ei(mtxd), ecz(mtxd);
    for (int i = 0; i < mtxd; i++) cptr[colid[(long)ky[i] * ns[0] + kx[i]] + 1]++;
    for (int c = 0; c < ncol; c++) cptr[c + 1] += cptr[c];
    std::vector<int> fillp(cptr.begin(), cptr.end() - 1);
    for (int i = 0; i < mtxd; i++) {
        const int c = colid[(long)ky[i] * ns[0] + kx[i]];
        const int e = fillp[c]++;
        ei[e] = i;
        ecz[e] = (c << 12) | kz[i];
    }
    {
        std::vector<double> gwh(3 * (size_t)std::max(mtxd, 1));
        for (int i = 0; i < mtxd; i++)
            for (int j = 0; j < 3; j++) gwh[(size_t)j * mtxd + i] = (double)kgv[3L * (isort[i] - 1) + j];
        if ((st = ensure_stream()) || (st = upload(p->gw, gwh.data(), gwh.size()))) return st;
    }
    p->key_pts.resize(mtxd);
    for (int i = 0; i < mtxd; i++) p->key_pts[i] = ((long)kz[i] * ns[1] + ky[i]) * id + kx[i];

    if ((st = upload(p->col_ptr, cptr.data(), cptr.size()))) return st;
    if ((st = upload(p->col_uy, uy.data(), uy.size()))) return st;
    if ((st = upload(p->ent_i, ei.data(), ei.size()))) return st;
    if ((st = upload(p->ent_cz, ecz.data(), ecz.size()))) return st;
    {
        std::vector<int> zm((size_t)std::max(ncol, 1) * ns[2], -1);
        for (int i = 0; i < mtxd; i++) zm[(size_t)colid[(long)ky[i] * ns[0] + kx[i]] * ns[2] + kz[i]] = i;
        if ((st = upload(p->zmap, zm.data(), zm.size()))) return st;
    }
    p->h_ekpg.assign(ekpg, ekpg + mtxd);
    if ((st = upload(p->ekpg, ekpg, mtxd))) return st;
    for (int j = 0; j < 3; j++) {
        std::vector<double2> t(ns[j]);
        for (int k = 0; k < ns[j]; k++) {
            long double a = 2.0L * 3.141592653589793238462643383279502884L * k / ns[j];
            t[k] = make_double2((double)cosl(a), (double)sinl(a));
        }
        if ((st = upload(p->tw[j], t.data(), t.size()))) return st;
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

// LDS bytes of the kernels for element size cs
size_t lds_col(const FftGeom &g, size_t cs) { return cs * (g.n[2] + (size_t)COL_CT * (g.n[2] + 1)); }
int plane_rows(const FftGeom &g, size_t cs, size_t *bytes) {
    const int nxp = g.nx | 1;
    const size_t fixed = cs * ((size_t)g.n[0] + g.n[1] + (size_t)g.n[1] * nxp);
    const size_t budget = 80 * 1024;  // two blocks per CU when it fits
    int rb = std::min(g.n[1], 32);
    while (rb > 1 && fixed + cs * (size_t)rb * (g.n[0] + 1) > budget) rb--;
    *bytes = fixed + cs * (size_t)rb * (g.n[0] + 1);
    return rb;
}

// ---- specialised kernels: mesh lengths with a two-pass split ---------------
#define CPW_FAST_SIZES(X) X(24) X(48) X(64) X(72) X(80) X(96) X(108) X(128)

enum LocalStage { ST_COL_BWD, ST_PLANE, ST_COL_FWD };

struct LocalArgs {
    const double2 *psi;
    long ldp;
    double2 *hpsi;
    long ldh;
    int nb;
    int mode;  // COLF_ACC / COLF_KIN: what the gather adds to local/N
};

template <typename R, int N>
int launch_fast(cpwgpu_plan_s *p, const LocalArgs &a, cplx<R> *colbuf, const R *vr, int stage,
                hipStream_t s) {
    const FftGeom &g = p->g;
    constexpr int CT = col_tile<N>();
    const int ntiles = (g.ncol + CT - 1) / CT;
    const dim3 grid(8 * ((a.nb + 7) / 8) * ntiles);
    if (stage == ST_COL_BWD) {
        launch_timed(CPWGPU_K_COL_BWD, (k_col_bwd_fast<R, N>), grid, dim3(FT), 0, s, a.psi, a.ldp, colbuf,
                           (const int *)p->col_ptr.p, (const int *)p->ent_i.p,
                           (const int *)p->ent_cz.p, (const double2 *)p->tw[2].p, g, ntiles, a.nb);
    } else if (stage == ST_COL_FWD) {
        launch_timed(CPWGPU_K_COL_FWD, (k_col_fwd_fast<R, N>), grid, dim3(FT), 0, s, a.psi, a.ldp, a.hpsi, a.ldh,
                           (const cplx<R> *)colbuf, (const int *)p->col_ptr.p,
                           (const int *)p->ent_i.p, (const int *)p->ent_cz.p,
                           (const double *)p->ekpg.p, (const double2 *)p->tw[2].p, g, ntiles, a.nb,
                           p->inv_n, a.mode);
    } else {
        const size_t dyn =
            sizeof(cplx<R>) * ((size_t)N * plane_sub_stride(g.nx) + (size_t)plane_rows_fast<N>() * row_stride<N>());
        const size_t stat = sizeof(cplx<R>) * 2 * N;
        if (dyn + stat > 160 * 1024)
            return fail(CPWGPU_ERR_SIZE, "plane of %d x %d does not fit in LDS", N, N);
        static bool attr = false;
        if (!attr) {
            HIPCHK(hipFuncSetAttribute((const void *)k_plane_fast<R, N, N>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(160 * 1024 - stat)));
            attr = true;
        }
        launch_timed(CPWGPU_K_PLANE, (k_plane_fast<R, N, N>), dim3(g.n[2] * a.nb), dim3(FT), dyn, s, colbuf,
                           (const int *)p->col_uy.p, vr, (const double2 *)p->tw[0].p,
                           (const double2 *)p->tw[1].p, g);
    }
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// ---- plane stage variants ---------------------------------------------------
// 0: k_plane_fast (LDS two-pass lines); 4..6: k_plane_wave (LDS two-pass
// lines, wave-local exchanges) with 256, 192 and 128 threads per block; the
// others: k_plane_reg shapes (x lanes GX, y lanes GY, planes per block P,
// threads T, waves per SIMD W, y lines through the LDS wave transform YW,
// box pruning PR) per precision.
struct RegShapeT {
    int GX, GY, P, T, W;
    bool YW, PR;
};
template <typename R> constexpr RegShapeT reg_shape(int v);
template <> constexpr RegShapeT reg_shape<float>(int v) {
    // 1-3 spill-free at 72: 67, 74 and 127 VGPRs
    return v == 1 ? RegShapeT{8, 8, 1, 256, 5, false, false}
         : v == 2 ? RegShapeT{8, 8, 2, 256, 4, false, false}
         : v == 3 ? RegShapeT{4, 4, 1, 256, 3, false, false}
         : v == 7 ? RegShapeT{8, 8, 1, 256, 5, true, false}
         : v == 8 ? RegShapeT{2, 2, 1, 256, 3, true, true}
         : v == 9 ? RegShapeT{4, 4, 1, 256, 4, true, true}
         : v == 10 ? RegShapeT{8, 8, 1, 256, 5, true, true}
         : v == 11 ? RegShapeT{2, 4, 1, 256, 3, false, true}
                   : RegShapeT{0, 0, 0, 0, 0, false, false};
}
template <> constexpr RegShapeT reg_shape<double>(int v) {
    return v == 1 ? RegShapeT{8, 8, 1, 256, 3, false, false}
         : v == 2 ? RegShapeT{8, 8, 1, 320, 4, false, false}
         : v == 3 ? RegShapeT{4, 8, 1, 256, 3, false, false}
         : v == 7 ? RegShapeT{8, 8, 1, 256, 3, true, false}
         : v == 8 ? RegShapeT{4, 4, 1, 256, 3, true, true}
         : v == 9 ? RegShapeT{8, 8, 1, 256, 3, true, true}
         : v == 10 ? RegShapeT{4, 4, 1, 256, 3, false, true}
         : v == 11 ? RegShapeT{8, 8, 1, 256, 3, false, true}
                   : RegShapeT{0, 0, 0, 0, 0, false, false};
}
constexpr int NPLANE_VAR = 12;
constexpr int WAVE_T[3] = {256, 192, 128};
constexpr bool is_reg_variant(int v) { return (v >= 1 && v <= 3) || (v >= 7 && v < NPLANE_VAR); }
// back-projection kernels: 0 LDS-staged, 1..NBACK_DIRECT direct, NBACK_DIRECT+1 LDS-staged with half chunks
constexpr int NBACK_VAR = NBACK_DIRECT + 2;
// projection kernels: 0 LDS-staged, 1..NPROJ_DIRECT direct, NPROJ_DIRECT+1 LDS-staged with half wave tiles
constexpr int NPROJ_VAR = NPROJ_DIRECT + 2;
template <typename R> constexpr int proj_half_wm() { return GemmCfg<R>::WM / 2; }
template <typename R> constexpr int proj_half_wn() { return GemmCfg<R>::WN > 1 ? GemmCfg<R>::WN / 2 : 1; }

// a register shape is built for a mesh length when both splits divide N and a
// lane holds at most 288 bytes of one line
template <typename R, int N, int GX, int GY> constexpr bool reg_ok() {
    return GX > 0 && GY > 0 && N % GX == 0 && N % GY == 0 && first_factor(N / GX) <= 5 &&
           first_factor(N / GY) <= 5 && 2 * sizeof(R) * (N / GX) <= 288 && 2 * sizeof(R) * (N / GY) <= 288 &&
           N / GX <= 64 && N / GY <= 64;
}
template <typename R, int N> constexpr bool reg_var_ok(int v) {
#define CPW_RV(V) (v == V && reg_ok<R, N, reg_shape<R>(V).GX, reg_shape<R>(V).GY>())
    return CPW_RV(1) || CPW_RV(2) || CPW_RV(3) || CPW_RV(7) || CPW_RV(8) || CPW_RV(9) || CPW_RV(10) ||
           CPW_RV(11);
#undef CPW_RV
}
// a pruned shape needs the sphere box inside |x|, |y| <= N/4 (true on the
// meshes size_fft chooses for the density, n >= 4 kd + 1)
inline bool prune_fits(const FftGeom &g) { return 4 * g.kd[0] <= g.n[0] && 4 * g.kd[1] <= g.n[1]; }

// LDS row stride of the register variant's sub-plane for nx box columns
inline int reg_stride(int nx) { return plane_sub_stride(nx); }

template <typename R, int N, int V>
int launch_plane_reg(cpwgpu_plan_s *p, const LocalArgs &a, cplx<R> *colbuf, const R *vr,
                     hipStream_t s) {
    constexpr RegShapeT SH = reg_shape<R>(V);
    constexpr int GX = SH.GX, GY = SH.GY, P = SH.P, T = SH.T, W = SH.W;
    constexpr bool YW = SH.YW, PR = SH.PR;
    if constexpr (!reg_ok<R, N, GX, GY>()) {
        return fail(CPWGPU_ERR_MODE, "plane variant %d is not built for N = %d", V, N);
    } else {
        const FftGeom &g = p->g;
        if (PR && !prune_fits(g))
            return fail(CPWGPU_ERR_MODE, "plane variant %d needs the sphere box inside N/4", V);
        // pruned shapes use their compile-time row stride (the box has at
        // most N/2 + 1 columns there)
        const int S = PR ? pr_sub_stride<N>() : reg_stride(g.nx);
        if (reg_stride(g.nx) > S)
            return fail(CPWGPU_ERR_MODE, "plane variant %d: %d box columns exceed its row stride", V, g.nx);
        const size_t dyn = plane_reg_lds<R, N, P>(S);
        if (dyn > 160 * 1024)
            return fail(CPWGPU_ERR_SIZE, "plane variant %d: %zu bytes of LDS", V, dyn);
        static bool attr = false;
        if (!attr) {
            HIPCHK(hipFuncSetAttribute((const void *)k_plane_reg<R, N, GX, GY, P, T, W, YW, PR>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = true;
        }
        launch_timed(CPWGPU_K_PLANE, k_plane_reg<R, N, GX, GY, P, T, W, YW, PR>,
                     dim3(((a.nb + P - 1) / P) * g.n[2]), dim3(T), dyn, s, colbuf,
                     (const int *)p->col_uy.p, vr, (const double2 *)p->tw[0].p, g, a.nb, S);
        HIPCHK(hipGetLastError());
        return CPWGPU_OK;
    }
}

template <typename R, int N, int T>
int launch_plane_wave(cpwgpu_plan_s *p, const LocalArgs &a, cplx<R> *colbuf, const R *vr, hipStream_t s) {
    const FftGeom &g = p->g;
    const size_t dyn = plane_wave_lds<R, N, T>(g.nx);
    const size_t stat = sizeof(cplx<R>) * 2 * N;
    if (dyn + stat > 160 * 1024)
        return fail(CPWGPU_ERR_SIZE, "plane of %d x %d does not fit in LDS with %d threads", N, N, T);
    static bool attr = false;
    if (!attr) {
        HIPCHK(hipFuncSetAttribute((const void *)k_plane_wave<R, N, N, T, 2>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(160 * 1024 - stat)));
        attr = true;
    }
    launch_timed(CPWGPU_K_PLANE, (k_plane_wave<R, N, N, T, 2>), dim3(g.n[2] * a.nb), dim3(T), dyn, s, colbuf,
                       (const int *)p->col_uy.p, vr, (const double2 *)p->tw[0].p, (const double2 *)p->tw[1].p, g);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// ---- register column kernels -------------------------------------------------
// lanes per z column: 8 where the column fits in registers, else 4
template <typename R, int N> constexpr int col_reg_g() {
    return reg_ok<R, N, 8, 8>() ? 8 : reg_ok<R, N, 4, 4>() ? 4 : 0;
}
template <typename R, int N> constexpr bool col_reg_built() { return col_reg_g<R, N>() > 0; }
// the sphere box inside |z| <= N3/4: the pruned register column kernels
inline bool col_prune_fits(const FftGeom &g) { return 4 * g.kd[2] <= g.n[2]; }

template <typename R, int N>
int launch_col_reg(cpwgpu_plan_s *p, const LocalArgs &a, cplx<R> *colbuf, int stage, hipStream_t s) {
    constexpr int G = col_reg_g<R, N>();
    if constexpr (G == 0) {
        return fail(CPWGPU_ERR_MODE, "register column kernels are not built for N = %d", N);
    } else {
        const FftGeom &g = p->g;
        constexpr int CPB = COLR_T / G;
        const int ntiles = (g.ncol + CPB - 1) / CPB;
        const dim3 grid(8 * ((a.nb + 7) / 8) * ntiles);
        const bool pr = col_prune_fits(g);
        const double2 *tw3 = (const double2 *)p->tw[2].p;
        const int *zm = (const int *)p->zmap.p;
        if (stage == ST_COL_BWD) {
            if (pr)
                launch_timed(CPWGPU_K_COL_BWD, (k_col_bwd_reg<R, N, G, true>), grid, dim3(COLR_T), 0, s, a.psi, a.ldp, colbuf,
                                   zm, tw3, g, ntiles, a.nb);
            else
                launch_timed(CPWGPU_K_COL_BWD, (k_col_bwd_reg<R, N, G, false>), grid, dim3(COLR_T), 0, s, a.psi, a.ldp, colbuf,
                                   zm, tw3, g, ntiles, a.nb);
        } else {
            if (pr)
                launch_timed(CPWGPU_K_COL_FWD, (k_col_fwd_reg<R, N, G, true>), grid, dim3(COLR_T), 0, s, a.psi, a.ldp, a.hpsi,
                                   a.ldh, (const cplx<R> *)colbuf, zm, (const double *)p->ekpg.p, tw3, g, ntiles,
                                   a.nb, p->inv_n, a.mode);
            else
                launch_timed(CPWGPU_K_COL_FWD, (k_col_fwd_reg<R, N, G, false>), grid, dim3(COLR_T), 0, s, a.psi, a.ldp, a.hpsi,
                                   a.ldh, (const cplx<R> *)colbuf, zm, (const double *)p->ekpg.p, tw3, g, ntiles,
                                   a.nb, p->inv_n, a.mode);
        }
        HIPCHK(hipGetLastError());
        return CPWGPU_OK;
    }
}

// whether the wave-exchange variant v (4..6) fits in LDS for nx box columns
template <typename R, int N> bool plane_wave_fits(int v, int nx) {
    const size_t stat = sizeof(cplx<R>) * 2 * N;
    const size_t dyn = v == 4 ? plane_wave_lds<R, N, WAVE_T[0]>(nx)
                       : v == 5 ? plane_wave_lds<R, N, WAVE_T[1]>(nx) : plane_wave_lds<R, N, WAVE_T[2]>(nx);
    return dyn + stat <= 160 * 1024;
}

template <typename R, int N> constexpr bool plane_var_ok(int v) {
    return v == 0 || (v >= 4 && v <= 6) || reg_var_ok<R, N>(v);
}

// default plane variant of a mesh length
template <typename R, int N> constexpr int plane_var_default() { return 0; }

template <typename R, int N>
int launch_plane(cpwgpu_plan_s *p, const LocalArgs &a, cplx<R> *colbuf, const R *vr, hipStream_t s) {
    int v = p->plane_var[sizeof(R) == 4 ? 1 : 0];
    if (v < 0 || !plane_var_ok<R, N>(v)) v = plane_var_default<R, N>();
#define CPW_LR(V) if (v == V) return launch_plane_reg<R, N, V>(p, a, colbuf, vr, s);
    CPW_LR(1) CPW_LR(2) CPW_LR(3) CPW_LR(7) CPW_LR(8) CPW_LR(9) CPW_LR(10) CPW_LR(11)
#undef CPW_LR
    if (v == 4) return launch_plane_wave<R, N, WAVE_T[0]>(p, a, colbuf, vr, s);
    if (v == 5) return launch_plane_wave<R, N, WAVE_T[1]>(p, a, colbuf, vr, s);
    if (v == 6) return launch_plane_wave<R, N, WAVE_T[2]>(p, a, colbuf, vr, s);
    return launch_fast<R, N>(p, a, colbuf, vr, ST_PLANE, s);
}

bool fast_sizes(const FftGeom &g) {
    bool col_ok = false, plane_ok = false;
#define CPW_HAS(N) col_ok |= g.n[2] == N; plane_ok |= (g.n[0] == N && g.n[1] == N);
    CPW_FAST_SIZES(CPW_HAS)
#undef CPW_HAS
    return col_ok && plane_ok;
}

// plane variant the next apply uses (defaults resolved), and whether v is
// built for the plan's mesh
template <typename R> int plane_var_resolve(const cpwgpu_plan_s *p, int v, bool *ok) {
    const FftGeom &g = p->g;
    *ok = v == 0 || v == -1;
    if (!fast_sizes(g)) return 0;
    int def = 0;
#define CPW_PV(N)                                      \
    if (g.n[0] == N) {                                 \
        def = plane_var_default<R, N>();               \
        if (v > 0)                                     \
            *ok = v < NPLANE_VAR && plane_var_ok<R, N>(v) && (v < 4 || v > 6 || plane_wave_fits<R, N>(v, g.nx)) && \
                  (!is_reg_variant(v) || !reg_shape<R>(v).PR || prune_fits(g)); \
    }
    CPW_FAST_SIZES(CPW_PV)
#undef CPW_PV
    return v < 0 || !*ok ? def : v;
}

// column kernel the next apply uses (default resolved: 0, the LDS lines,
// which measured as fast as the register lines in FP64 and faster in FP32 at
// 72^3); *ok: whether v is usable (1 needs the register kernels built for n3)
template <typename R> int col_var_resolve(const cpwgpu_plan_s *p, int v, bool *ok) {
    const FftGeom &g = p->g;
    bool built = false;
    if (fast_sizes(g)) {
#define CPW_CB(N) if (g.n[2] == N) built = col_reg_built<R, N>();
        CPW_FAST_SIZES(CPW_CB)
#undef CPW_CB
    }
    if (ok) *ok = v <= 0 || (v == 1 && built);
    return v == 1 && built ? 1 : 0;
}

// One stage of the local term: specialised kernels when the mesh allows,
// the generic runtime-radix kernels otherwise.
template <typename R>
int local_stage(cpwgpu_plan_s *p, const LocalArgs &a, cplx<R> *colbuf, const R *vr, int stage,
                hipStream_t s) {
    const FftGeom &g = p->g;
    const size_t cs = sizeof(cplx<R>);
    if (fast_sizes(g)) {
        int st = CPWGPU_OK;
        if (stage == ST_PLANE) {
#define CPW_PL(N) if (g.n[0] == N) st = launch_plane<R, N>(p, a, colbuf, vr, s);
            CPW_FAST_SIZES(CPW_PL)
#undef CPW_PL
        } else {
            const bool reg = col_var_resolve<R>(p, p->col_var[sizeof(R) == 4 ? 1 : 0], nullptr) == 1;
#define CPW_CO(N)                                                                \
    if (g.n[2] == N)                                                             \
        st = reg ? launch_col_reg<R, N>(p, a, colbuf, stage, s) : launch_fast<R, N>(p, a, colbuf, vr, stage, s);
            CPW_FAST_SIZES(CPW_CO)
#undef CPW_CO
        }
        return st;
    }
    const int ntiles = (g.ncol + COL_CT - 1) / COL_CT;
    const size_t lcol = lds_col(g, cs);
    size_t lplane;
    const int rb = plane_rows(g, cs, &lplane);
    if (lplane > 160 * 1024 || lcol > 160 * 1024)
        return fail(CPWGPU_ERR_SIZE, "FFT mesh too large for LDS (%zu / %zu bytes)", lplane, lcol);
    static bool attr_done = false;
    if (!attr_done) {
        HIPCHK(hipFuncSetAttribute((const void *)k_col_bwd<R, COL_CT, KMAX>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        HIPCHK(hipFuncSetAttribute((const void *)k_col_fwd<R, COL_CT, KMAX>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        HIPCHK(hipFuncSetAttribute((const void *)k_plane<R, KMAX>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_done = true;
    }
    const double2 *tw1 = (const double2 *)p->tw[0].p, *tw2 = (const double2 *)p->tw[1].p,
                  *tw3 = (const double2 *)p->tw[2].p;
    if (stage == ST_COL_BWD) {
        launch_timed(CPWGPU_K_COL_BWD, (k_col_bwd<R, COL_CT, KMAX>), dim3(ntiles * a.nb), dim3(COL_THREADS), lcol, s,
                           a.psi, a.ldp, colbuf, (const int *)p->col_ptr.p, (const int *)p->ent_i.p,
                           (const int *)p->ent_cz.p, tw3, g, ntiles);
    } else if (stage == ST_PLANE) {
        launch_timed(CPWGPU_K_PLANE, (k_plane<R, KMAX>), dim3(g.n[2] * a.nb), dim3(PLANE_THREADS), lplane, s,
                           colbuf, (const int *)p->col_uy.p, vr, tw1, tw2, g, rb);
    } else {
        launch_timed(CPWGPU_K_COL_FWD, (k_col_fwd<R, COL_CT, KMAX>), dim3(ntiles * a.nb), dim3(COL_THREADS), lcol, s,
                           a.psi, a.ldp, a.hpsi, a.ldh, (const cplx<R> *)colbuf,
                           (const int *)p->col_ptr.p, (const int *)p->ent_i.p,
                           (const int *)p->ent_cz.p, (const double *)p->ekpg.p, tw3, g, ntiles,
                           p->inv_n, a.mode);
    }
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// Kleinman-Bylander term: hpsi = beta*hpsi + anlga Diag(xnlkb) anlga^H psi
// (Src/cpw_lib/hk_psi_nl_c16.f90:78-96) on stream s.
// One set of projectors: the plan's (mtxd rows) or the spinor ones (2*mtxd
// interleaved rows, Src/hk_psi/reference/hk_psi_spin_c16.f90:302-303).
struct ProjView {
    DevBuf *anl64, *anl32;
    bool *anl32_valid;
    long lda;
    long *lda32;
    int nanl;
    const double *xnl;
    int rows;
};

ProjView scalar_proj(cpwgpu_plan_s *p) {
    return {&p->anl64, &p->anl32, &p->anl32_valid, p->lda, &p->lda32, p->nanl, (const double *)p->xnl.p,
            p->g.mtxd};
}

ProjView spin_proj(cpwgpu_plan_s *p) {
    return {&p->sanl64, &p->sanl32, &p->sanl32_valid, p->slda, &p->slda32, p->snanl,
            (const double *)p->sxnl.p, 2 * p->g.mtxd};
}

// Caller-provided device workspace of one apply (the allocation-free entry
// point): the z-column buffer of the local stage and the split-K partials and
// D<beta|psi> of the non-local stage, each with its size in bytes.  Null:
// the plan's own grow-only buffers.
struct Work {
    void *colbuf = nullptr;
    size_t colbuf_n = 0;
    void *part = nullptr;
    size_t part_n = 0;
    void *dhd = nullptr;
    size_t dhd_n = 0;
};

// Block tile (BM x BN) and K granule of the projection kernel variant pv.
template <typename R> void proj_tile(int pv, int *BM, int *BN, int *BK) {
    *BM = proj_bm<R>();
    *BN = proj_bn<R>();
    *BK = GemmCfg<R>::BK;
    if (pv == NPROJ_DIRECT + 1) {
        *BM = 2 * proj_half_wm<R>() * 16;
        *BN = 2 * proj_half_wn<R>() * 16;
    }
#define CPW_PD(V) if (pv == V) { *BM = direct_bm<R, V>(); *BN = direct_bn<R, V>(); *BK = direct_bk<R, V>(); }
    CPW_PD(1) CPW_PD(2) CPW_PD(3) CPW_PD(4)
#undef CPW_PD
}

// projection kernel variant of the plan for precision R
template <typename R> int proj_variant(const cpwgpu_plan_s *p) {
    const int ix = sizeof(R) == 4 ? 1 : 0;
    return p->proj_var[ix] >= 1 && p->proj_var[ix] < NPROJ_VAR ? p->proj_var[ix] : 0;
}

// Split-K shape of the projection GEMM over rows plane waves: number of K
// slices and rows per slice (whole LDS stages of k_proj / whole chunks of
// k_proj_direct), for nanl projectors and neig bands.
template <typename R> void splitk_shape(const cpwgpu_plan_s *p, int rows, int nanl, int neig, int *nsl_out,
                                        int *kslice_out) {
    int BM, BN, BK;
    proj_tile<R>(proj_variant<R>(p), &BM, &BN, &BK);
    const int ix = sizeof(R) == 4 ? 1 : 0;
    const int tm = (nanl + BM - 1) / BM, tn = (neig + BN - 1) / BN;
    const int ksteps = (rows + BK - 1) / BK;
    const int target = p->splitk[ix] > 0 ? p->splitk[ix] : 256;
    int nsl = std::max(1, std::min(ksteps, (target + tm * tn - 1) / (tm * tn)));
    const int kslice = ((ksteps + nsl - 1) / nsl) * BK;
    *nsl_out = (rows + kslice - 1) / kslice;
    *kslice_out = kslice;
}

// hpsi = beta*hpsi + sum_a beta_a xnlkb_a <beta_a|psi> (Src/cpw_lib/hk_psi_nl_c16.f90:82-96),
// plus ekpg*psi when kin (the kinetic term, only for the plan's own rows)
// phase: NL_ALL, or NL_PROJ (projection and D scaling only: D<beta|psi> left
// in the workspace) then NL_BACK (the back-projection from it) on the same
// workspace, possibly on another stream ordered after the first.
enum NlPhase { NL_ALL, NL_PROJ, NL_BACK };
template <typename R>
int nonlocal(cpwgpu_plan_s *p, const ProjView &pv_, int neig, const double *psi, long ldp, double *hpsi,
             long ldh, int beta, bool kin, hipStream_t s, const Work *w = nullptr, int phase = NL_ALL) {
    const size_t cs = sizeof(cplx<R>);
    const int nanl = pv_.nanl;
    const int rows = pv_.rows;
    int st;
    const cplx<R> *anl;
    if (sizeof(R) == 8) {
        anl = (const cplx<R> *)pv_.anl64->p;
    } else {
        if (!*pv_.anl32_valid && w)
            return fail(CPWGPU_ERR_ARG, "no FP32 copy of the projectors: it is made by the first FP32 apply "
                                        "(cpwgpu_plan_apply) after the projectors are set");
        if (!*pv_.anl32_valid) {
            *pv_.lda32 = (pv_.lda + 1) & ~1L;
            if ((st = pv_.anl32->reserve(sizeof(cplx<float>) * (size_t)*pv_.lda32 * nanl))) return st;
            const long tot = (long)rows * nanl;
            hipLaunchKernelGGL(k_to_c64, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                               (const double2 *)pv_.anl64->p, pv_.lda, (cplx<float> *)pv_.anl32->p,
                               *pv_.lda32, rows, nanl);
            HIPCHK(hipGetLastError());
            *pv_.anl32_valid = true;
        }
        anl = (const cplx<R> *)pv_.anl32->p;
    }
    const long lda = sizeof(R) == 8 ? pv_.lda : *pv_.lda32;
    const int ix = sizeof(R) == 4 ? 1 : 0;
    const int pv = proj_variant<R>(p);
    // block tile and K granule of the chosen projection kernel
    int BM, BN, BK;
    proj_tile<R>(pv, &BM, &BN, &BK);
    const int tm = (nanl + BM - 1) / BM, tn = (neig + BN - 1) / BN;
    int nsl, kslice;
    splitk_shape<R>(p, rows, nanl, neig, &nsl, &kslice);
    cplx<R> *part, *dhd;
    if (w) {
        if (cs * (size_t)nsl * neig * nanl > w->part_n || cs * (size_t)neig * nanl > w->dhd_n)
            return fail(CPWGPU_ERR_SIZE, "non-local workspace too small");
        part = (cplx<R> *)w->part;
        dhd = (cplx<R> *)w->dhd;
    } else {
        if ((st = p->part.reserve(cs * (size_t)nsl * neig * nanl))) return st;
        if ((st = p->dhd.reserve(cs * (size_t)neig * nanl))) return st;
        part = (cplx<R> *)p->part.p;
        dhd = (cplx<R> *)p->dhd.p;
    }
    if (phase != NL_BACK) {
        const dim3 grid(tm, tn, nsl);
        const double2 *ps = (const double2 *)psi;
        cplx<R> *pt = part;
        if (pv == 1)
            launch_timed(CPWGPU_K_PROJ, (k_proj_direct<R, 1>), grid, dim3(GEMM_THREADS), 0, s, anl, lda, ps, ldp, pt, rows, nanl, neig, kslice, nullptr, 0);
        else if (pv == 2)
            launch_timed(CPWGPU_K_PROJ, (k_proj_direct<R, 2>), grid, dim3(GEMM_THREADS), 0, s, anl, lda, ps, ldp, pt, rows, nanl, neig, kslice, nullptr, 0);
        else if (pv == 3)
            launch_timed(CPWGPU_K_PROJ, (k_proj_direct<R, 3>), grid, dim3(GEMM_THREADS), 0, s, anl, lda, ps, ldp, pt, rows, nanl, neig, kslice, nullptr, 0);
        else if (pv == 4)
            launch_timed(CPWGPU_K_PROJ, (k_proj_direct<R, 4>), grid, dim3(GEMM_THREADS), 0, s, anl, lda, ps, ldp, pt, rows, nanl, neig, kslice, nullptr, 0);
        else if (pv == NPROJ_DIRECT + 1)
            launch_timed(CPWGPU_K_PROJ, (k_proj<R, proj_half_wm<R>(), proj_half_wn<R>()>), dim3(tm, tn, nsl),
                               dim3(GEMM_THREADS), 0, s, anl, lda, (const double2 *)psi, ldp,
                               pt, rows, nanl, neig, kslice, nullptr, 0);
        else
            launch_timed(CPWGPU_K_PROJ, k_proj<R>, dim3(tm, tn, nsl), dim3(GEMM_THREADS), 0, s, anl, lda,
                               (const double2 *)psi, ldp, pt, rows, nanl, neig, kslice, nullptr, 0);
    }
    if (phase != NL_BACK) {
        // launch_proj_reduce's launch, with the measurement hook
        const long tot = (long)nanl * neig;
        if (tot > 0)
            launch_timed(CPWGPU_K_PROJ_REDUCE, k_proj_reduce<R>, dim3((unsigned)((tot + RED_OUT - 1) / RED_OUT)),
                         dim3(RED_OUT * RED_GRP), 0, s, (const cplx<R> *)part, pv_.xnl, dhd, nanl, neig, nsl, 0);
    }
    if (phase != NL_PROJ) {
        const int bv = p->back_var[ix] >= 1 && p->back_var[ix] <= NBACK_VAR - 1 ? p->back_var[ix] : 0;
        const cplx<R> *dh = dhd;
        double2 *hp = (double2 *)hpsi;
        // the kinetic term ekpg * psi is added here (coalesced), not in the gather
        const double2 *kp = kin ? (const double2 *)psi : nullptr;
        const double *ek = kin ? (const double *)p->ekpg.p : nullptr;
#define CPW_BD(V)                                                                                  \
    if (bv == V)                                                                                   \
        launch_timed(CPWGPU_K_BACK, (k_back_direct<R, V>),                                                  \
                           dim3((rows + back_bi<R, V>() - 1) / back_bi<R, V>(),                  \
                                (neig + back_bn<R, V>() - 1) / back_bn<R, V>()),                   \
                           dim3(GEMM_THREADS), 0, s, anl, lda, dh, hp, ldh, rows, nanl, neig, beta, kp,   \
                           ldp, ek);
        CPW_BD(1) CPW_BD(2) CPW_BD(3)
#undef CPW_BD
        if (bv == 0)
            launch_timed(CPWGPU_K_BACK, k_back<R>, dim3((rows + BP_BI - 1) / BP_BI, (neig + BP_BN - 1) / BP_BN),
                               dim3(GEMM_THREADS), 0, s, anl, lda, dh, hp, ldh, rows, nanl, neig, beta,
                               kp, ldp, ek);
        else if (bv == NBACK_DIRECT + 1)
            launch_timed(CPWGPU_K_BACK, (k_back<R, GemmCfg<R>::KC / 2>),
                               dim3((rows + BP_BI - 1) / BP_BI, (neig + BP_BN - 1) / BP_BN),
                               dim3(GEMM_THREADS), 0, s, anl, lda, dh, hp, ldh, rows, nanl, neig, beta,
                               kp, ldp, ek);
    }
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// second stream and fork/join events: the projector GEMMs (matrix cores) run
// beside the FFT kernels (LDS / HBM)
hipStream_t g_aux = nullptr;
hipEvent_t g_fork = nullptr, g_join = nullptr;

int ensure_aux() {
    if (!g_aux) {
        HIPCHK(hipStreamCreateWithFlags(&g_aux, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&g_join, hipEventDisableTiming));
    }
    return CPWGPU_OK;
}

// Local and kinetic terms of one k-point on stream s, in band chunks that keep
// colbuf below 4 GiB.  acc: the gather adds to the non-local term already in
// hpsi; join (may be null): event the first gather waits for (that term).
template <typename R>
int local_part(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh,
               bool acc, hipEvent_t join, hipStream_t s, bool kin_gather = false, const Work *w = nullptr,
               bool write_nokin = false) {
    const FftGeom &g = p->g;
    const size_t cs = sizeof(cplx<R>);
    const R *vr = sizeof(R) == 8 ? (const R *)p->v64.p : (const R *)p->v32.p;
    if (!vr) return fail(CPWGPU_ERR_ARG, "potential not set");
    const size_t per_band = cs * (size_t)g.n[2] * g.ncolp;
    const size_t cap = w ? w->colbuf_n : (4ull << 30);
    if (per_band > cap) return fail(CPWGPU_ERR_SIZE, "z-column workspace too small for one band");
    const int nbc = (int)std::max<size_t>(1, std::min<size_t>(neig, cap / per_band));
    int st;
    cplx<R> *colbuf;
    if (w) {
        colbuf = (cplx<R> *)w->colbuf;
    } else {
        if ((st = p->colbuf.reserve(per_band * nbc))) return st;
        colbuf = (cplx<R> *)p->colbuf.p;
    }
    bool joined = join == nullptr;
    for (int b0 = 0; b0 < neig; b0 += nbc) {
        const int nb = std::min(nbc, neig - b0);
        // with the non-local term the back-projection already wrote it and the
        // kinetic term into hpsi; without it the gather adds the kinetic term
        // write_nokin: the gather writes local/N (+ ekpg psi if kin_gather)
        // and the back-projection adds the rest afterwards
        const int mode = acc ? (COLF_ACC | (kin_gather ? COLF_KIN : 0))
                             : write_nokin ? (kin_gather ? COLF_KIN : 0) : COLF_KIN;
        const LocalArgs la{(const double2 *)psi + (long)b0 * ldp, ldp,
                           (double2 *)hpsi + (long)b0 * ldh, ldh, nb, mode};
        if (g.ncol > 0) {
            if ((st = local_stage<R>(p, la, colbuf, vr, ST_COL_BWD, s))) return st;
            if ((st = local_stage<R>(p, la, colbuf, vr, ST_PLANE, s))) return st;
        }
        if (!joined) {
            // join: the gather adds the local and kinetic terms to the non-local one
            HIPCHK(hipStreamWaitEvent(s, join, 0));
            joined = true;
        }
        if (g.ncol > 0 && (st = local_stage<R>(p, la, colbuf, vr, ST_COL_FWD, s))) return st;
    }
    return CPWGPU_OK;
}

// Local and kinetic terms with the inverse column and plane kernels on stream
// a and the gather on stream s, after the non-local term the caller enqueued
// on s (written into hpsi): per band chunk, a runs the inverse columns and the
// plane stage, s waits for them and gathers (adding to hpsi), and a waits for
// that gather before the next chunk reuses colbuf.  The caller has ordered a
// after s's earlier work.
template <typename R>
int local_beside(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh,
                 hipStream_t s, hipStream_t a, bool kin_gather, const Work *w) {
    const FftGeom &g = p->g;
    if (g.ncol <= 0) return CPWGPU_OK;
    const size_t cs = sizeof(cplx<R>);
    const R *vr = sizeof(R) == 8 ? (const R *)p->v64.p : (const R *)p->v32.p;
    if (!vr) return fail(CPWGPU_ERR_ARG, "potential not set");
    const size_t per_band = cs * (size_t)g.n[2] * g.ncolp;
    const size_t cap = w ? w->colbuf_n : (4ull << 30);
    if (per_band > cap) return fail(CPWGPU_ERR_SIZE, "z-column workspace too small for one band");
    const int nbc = (int)std::max<size_t>(1, std::min<size_t>(neig, cap / per_band));
    int st;
    cplx<R> *colbuf;
    if (w) {
        colbuf = (cplx<R> *)w->colbuf;
    } else {
        if ((st = p->colbuf.reserve(per_band * nbc))) return st;
        colbuf = (cplx<R> *)p->colbuf.p;
    }
    const int mode = COLF_ACC | (kin_gather ? COLF_KIN : 0);
    for (int b0 = 0; b0 < neig; b0 += nbc) {
        const int nb = std::min(nbc, neig - b0);
        const LocalArgs la{(const double2 *)psi + (long)b0 * ldp, ldp,
                           (double2 *)hpsi + (long)b0 * ldh, ldh, nb, mode};
        if (b0 > 0) {  // colbuf is free again after the previous chunk's gather
            HIPCHK(hipEventRecord(g_fork, s));
            HIPCHK(hipStreamWaitEvent(a, g_fork, 0));
        }
        if ((st = local_stage<R>(p, la, colbuf, vr, ST_COL_BWD, a))) return st;
        if ((st = local_stage<R>(p, la, colbuf, vr, ST_PLANE, a))) return st;
        HIPCHK(hipEventRecord(g_join, a));
        HIPCHK(hipStreamWaitEvent(s, g_join, 0));
        if ((st = local_stage<R>(p, la, colbuf, vr, ST_COL_FWD, s))) return st;
    }
    return CPWGPU_OK;
}

// w: caller workspace (see Work), or null for the plan's buffers.  With a
// workspace the second stream is used only when it already exists.
template <typename R>
int apply_t(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh,
            hipStream_t s, const Work *w = nullptr) {
    const bool nl = p->nanl > 0 && p->g.mtxd > 0;
    // kinetic term in the gather (kg) or in the back-projection epilogue
    const bool kg = p->kin_var[sizeof(R) == 4 ? 1 : 0] == 1;
    int st;
    if (nl && (!g_overlap || (w && !g_aux))) {
        // serial: non-local term first (writes hpsi), then the local term adds
        if ((st = nonlocal<R>(p, scalar_proj(p), neig, psi, ldp, hpsi, ldh, 0, !kg, s, w))) return st;
        return local_part<R>(p, neig, psi, ldp, hpsi, ldh, true, nullptr, s, kg, w);
    }
    if (nl && p->sched_var[sizeof(R) == 4 ? 1 : 0] == 1) {
        // fork: projection and D scaling on the second stream, beside the
        // column and plane kernels; the gather writes the local term, then the
        // back-projection adds the non-local one on the caller's stream
        if ((st = ensure_aux())) return st;
        HIPCHK(hipEventRecord(g_fork, s));
        HIPCHK(hipStreamWaitEvent(g_aux, g_fork, 0));
        if ((st = nonlocal<R>(p, scalar_proj(p), neig, psi, ldp, hpsi, ldh, 1, !kg, g_aux, w, NL_PROJ)))
            return st;
        HIPCHK(hipEventRecord(g_join, g_aux));
        if ((st = local_part<R>(p, neig, psi, ldp, hpsi, ldh, false, nullptr, s, kg, w, true))) return st;
        HIPCHK(hipStreamWaitEvent(s, g_join, 0));
        return nonlocal<R>(p, scalar_proj(p), neig, psi, ldp, hpsi, ldh, 1, !kg, s, w, NL_BACK);
    }
    if (nl && p->sched_var[sizeof(R) == 4 ? 1 : 0] == 2) {
        // fork the other way round: the inverse column and plane kernels on
        // the second stream, the non-local term on the caller's stream (written
        // into hpsi), then the gather there once the plane stage is done.  The
        // plane stage usually ends before the back-projection, so the gather's
        // wait finds its event complete.
        if ((st = ensure_aux())) return st;
        HIPCHK(hipEventRecord(g_fork, s));
        HIPCHK(hipStreamWaitEvent(g_aux, g_fork, 0));
        if ((st = nonlocal<R>(p, scalar_proj(p), neig, psi, ldp, hpsi, ldh, 0, !kg, s, w))) return st;
        return local_beside<R>(p, neig, psi, ldp, hpsi, ldh, s, g_aux, kg, w);
    }
    if (nl) {
        // fork: non-local term on the second stream, written into hpsi
        if ((st = ensure_aux())) return st;
        HIPCHK(hipEventRecord(g_fork, s));
        HIPCHK(hipStreamWaitEvent(g_aux, g_fork, 0));
        if ((st = nonlocal<R>(p, scalar_proj(p), neig, psi, ldp, hpsi, ldh, 0, !kg, g_aux, w))) return st;
        HIPCHK(hipEventRecord(g_join, g_aux));
        return local_part<R>(p, neig, psi, ldp, hpsi, ldh, true, g_join, s, kg, w);
    }
    return local_part<R>(p, neig, psi, ldp, hpsi, ldh, false, nullptr, s, false, w);
}

int set_potential(cpwgpu_plan_s *p, const double *vscr, int mxdscr) {
    if (p->mwrk > mxdscr)
        return fail(CPWGPU_ERR_MXDSCR, "mxdwrk = %d is greater than mxdscr = %d", p->mwrk, mxdscr);
    if (p->ntot > (size_t)mxdscr)
        return fail(CPWGPU_ERR_MXDSCR, "mesh id*n2*n3 = %zu is greater than mxdscr = %d", p->ntot, mxdscr);
    int st = ensure_stream();
    if (st) return st;
    if ((st = upload(p->v64, vscr, p->ntot))) return st;
    {
        // V(G=0): mean over the n1*n2*n3 points (x beyond n1 is padding when id > n1)
        double s = 0.0;
        const int n1 = p->g.n[0], n2 = p->g.n[1], n3 = p->g.n[2], id = p->g.id;
        for (int k3 = 0; k3 < n3; k3++)
            for (int k2 = 0; k2 < n2; k2++) {
                const double *row = vscr + ((size_t)k3 * n2 + k2) * id;
                for (int k1 = 0; k1 < n1; k1++) s += row[k1];
            }
        p->v0 = s / ((double)n1 * n2 * n3);
    }
    if ((st = p->v32.reserve(sizeof(float) * p->ntot))) return st;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((p->ntot + 255) / 256)), dim3(256), 0, g_stream,
                       (const double *)p->v64.p, (float *)p->v32.p, (long)p->ntot);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// Replaces the plan's projectors with nanl columns: fill(anl, lda) writes
// them into the FP64 buffer (reserved for lda = max(1, mtxd) rows) and sets
// xnl, enqueued on g_stream.  nanl, lda and the FP32 copy change only after
// every step succeeded; a failure after the old buffer was released leaves
// the plan without projectors (nanl = 0), never with a partly written set.
// An FP32 copy that exists is converted again at once (same buffer when it
// is large enough), so a captured graph that reads it sees the new values.
template <typename F> int replace_projectors(cpwgpu_plan_s *p, int nanl, F fill) {
    int st = ensure_stream();
    if (st) return st;
    if (nanl == 0) {
        p->nanl = 0;
        p->anl32_valid = false;
        return CPWGPU_OK;
    }
    const long lda = std::max(1, p->mtxd);
    st = p->anl64.reserve(16 * (size_t)lda * nanl);
    if (!st) st = p->xnl.reserve(8 * (size_t)nanl);
    if (!st) st = fill((double2 *)p->anl64.p, lda);
    if (st) {
        p->nanl = 0;
        p->anl32_valid = false;
        return st;
    }
    p->nanl = nanl;
    p->lda = lda;
    p->anl32_valid = false;
    const long lda32 = (lda + 1) & ~1L;
    if (p->anl32.p && p->anl32.n >= sizeof(cplx<float>) * (size_t)lda32 * nanl) {
        p->lda32 = lda32;
        const long tot = (long)p->mtxd * nanl;
        if (tot > 0) {
            hipLaunchKernelGGL(k_to_c64, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g_stream,
                               (const double2 *)p->anl64.p, lda, (cplx<float> *)p->anl32.p, lda32, p->mtxd, nanl);
            HIPCHK(hipGetLastError());
        }
        p->anl32_valid = true;
    }
    return CPWGPU_OK;
}

int set_projectors_host(cpwgpu_plan_s *p, const double *anlga, const double *xnlkb, int nanl,
                        int mxddim) {
    if (nanl < 0) return fail(CPWGPU_ERR_ARG, "nanl = %d", nanl);
    if (nanl > 0 && (!anlga || !xnlkb)) return fail(CPWGPU_ERR_ARG, "null projector array");
    return replace_projectors(p, nanl, [&](double2 *anl, long lda) -> int {
        if (p->mtxd > 0)
            HIPCHK(hipMemcpy2DAsync(anl, 16 * (size_t)lda, anlga, 16 * (size_t)mxddim, 16 * (size_t)p->mtxd, nanl,
                                    hipMemcpyHostToDevice, g_stream));
        return upload(p->xnl, xnlkb, nanl);
    });
}

// ---- KB projectors on the device (Src/omp/proj_nl_kb_c16.f90:23-275) --------------
DevBuf g_kb_tab, g_kb_types, g_kb_atoms, g_kb_gw, g_kb_out;

struct KbSetup {
    KbCell cell{};
    std::vector<KbType> types;
    std::vector<KbAtom> atoms;   // atoms of the types that have projectors
    std::vector<double> tab;     // vkb(-2:mxdlqp, l, 0, type) lines, packed per (type, l)
    std::vector<double> xnl;     // nkb of each column
    long vstride = 0;
};

// Host half of proj_nl_kb_c16: checks (:143-157, and the cell of
// Src/crystal/adot_to_avec.f90 / adot_to_bdot.f90), column bookkeeping
// (:169-186), packed tables and xnlkb (:265-267).  Arrays in the reference's
// layouts: vkb(-2:mxdlqp, 0:3, -1:1, mxdtyp), nkb(0:3, -1:1, mxdtyp),
// rat(3, mxdatm, mxdtyp), adot(3, 3).
int kb_setup(const double *rkpt, int nanl, const int *nqnl, const double *delqnl, const double *vkb,
             const int *nkb, int ntype, const int *natom, const double *rat, const double *adot, int mxdtyp,
             int mxdatm, int mxdlqp, KbSetup &s) {
    if (!rkpt || !adot) return fail(CPWGPU_ERR_ARG, "null rkpt/adot");
    if (ntype < 0 || ntype > mxdtyp) return fail(CPWGPU_ERR_ARG, "ntype = %d, mxdtyp = %d", ntype, mxdtyp);
    if (mxdlqp < 0 || mxdatm < 0) return fail(CPWGPU_ERR_ARG, "mxdlqp = %d, mxdatm = %d", mxdlqp, mxdatm);
    if (ntype > 0 && (!nqnl || !delqnl || !vkb || !nkb || !natom || !rat))
        return fail(CPWGPU_ERR_ARG, "null pseudopotential array");
    // avec upper triangular with a1 along x and a2 in the xy plane: adot = avec^T avec
    const double eps = 1e-10;
    double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // a[r][c] = avec(r+1, c+1)
    const auto A = [&](int r, int c) { return adot[r + 3 * c]; };
    if (!(A(0, 0) >= eps)) return fail(CPWGPU_ERR_ARG, "adot_to_avec: avec(1,1) = %g", A(0, 0));
    a[0][0] = std::sqrt(A(0, 0));
    a[0][1] = A(0, 1) / a[0][0];
    double t = A(1, 1) - a[0][1] * a[0][1];
    if (!(t >= eps)) return fail(CPWGPU_ERR_ARG, "adot_to_avec: avec(2,2) = %g", t);
    a[1][1] = std::sqrt(t);
    a[0][2] = A(0, 2) / a[0][0];
    a[1][2] = (A(1, 2) - a[0][1] * a[0][2]) / a[1][1];
    t = A(2, 2) - a[1][2] * a[1][2] - a[0][2] * a[0][2];
    if (!(t >= eps)) return fail(CPWGPU_ERR_ARG, "adot_to_avec: avec(3,3) = %g", t);
    a[2][2] = std::sqrt(t);
    // bvec = 2 pi inv(avec)^T; inv of the upper triangular avec by back substitution
    double ia[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int c = 0; c < 3; c++) {
        ia[c][c] = 1.0 / a[c][c];
        for (int r = c - 1; r >= 0; r--) {
            double acc = 0.0;
            for (int k = r + 1; k <= c; k++) acc += a[r][k] * ia[k][c];
            ia[r][c] = -acc / a[r][r];
        }
    }
    const double twopi = 6.283185307179586477;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) s.cell.b[r + 3 * c] = twopi * ia[c][r];
    // vcell^2 = det adot
    const double det = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) -
                       A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                       A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    if (!(det >= eps)) return fail(CPWGPU_ERR_ARG, "adot_to_bdot: vcell squared = %g", det);
    s.cell.fac = 1.0 / std::sqrt(std::sqrt(det));
    for (int j = 0; j < 3; j++) s.cell.rk[j] = rkpt[j];
    // columns: type -> l -> m -> atom
    s.vstride = (long)mxdlqp + 3;
    s.types.assign(ntype, KbType{});
    s.tab.assign((size_t)std::max(ntype, 1) * 4 * s.vstride, 0.0);
    s.atoms.clear();
    s.xnl.clear();
    int ind = 0;
    for (int k = 0; k < ntype; k++) {
        KbType &ty = s.types[k];
        if (natom[k] < 0 || natom[k] > mxdatm)
            return fail(CPWGPU_ERR_ARG, "natom(%d) = %d, mxdatm = %d", k + 1, natom[k], mxdatm);
        ty.natom = natom[k];
        ty.nq = nqnl[k];
        ty.voff = (long)k * 4 * s.vstride;
        for (int l = 0; l < 4; l++) {
            ty.base[l] = ind;
            const int nk = nkb[l + 4 * (1 + 3 * k)];
            if (nk == 0) continue;
            ty.lmask |= 1 << l;
            ind += (2 * l + 1) * natom[k];
            for (int c = 0; c < (2 * l + 1) * natom[k]; c++) s.xnl.push_back((double)nk);
            std::memcpy(&s.tab[ty.voff + l * s.vstride], vkb + s.vstride * (l + 4 * (1 + 3L * k)),
                        sizeof(double) * s.vstride);
        }
        if (ty.lmask) {
            if (nqnl[k] > mxdlqp + 1)
                return fail(CPWGPU_ERR_ARG, "nqnl(%d) = %d > mxdlqp + 1 = %d", k + 1, nqnl[k], mxdlqp + 1);
            if (!(delqnl[k] > 0.0)) return fail(CPWGPU_ERR_ARG, "delqnl(%d) = %g", k + 1, delqnl[k]);
            ty.inv_delq = 1.0 / delqnl[k];
            for (int kk = 0; kk < natom[k]; kk++) {
                KbAtom at{};
                for (int j = 0; j < 3; j++) at.rat[j] = rat[j + 3 * ((long)kk + (long)mxdatm * k)];
                at.type = k;
                at.atom = kk;
                s.atoms.push_back(at);
            }
        }
    }
    if (ind != nanl) return fail(CPWGPU_ERR_ARG, "proj_nl_kb: inconsistent value of nanl %d should be %d", nanl, ind);
    if (s.atoms.size() > 65535) return fail(CPWGPU_ERR_SIZE, "%zu atoms with projectors", s.atoms.size());
    return CPWGPU_OK;
}

// anl(lda, nanl) complex128 on the device from the g_j of the plane waves
// (gw_dev[j * mtxd + i]); enqueued on g_stream.
int kb_build(const KbSetup &s, const double *gw_dev, int mtxd, double2 *anl, long lda) {
    if (mtxd == 0 || s.atoms.empty()) return CPWGPU_OK;
    int st;
    if ((st = upload(g_kb_tab, s.tab.data(), s.tab.size())) ||
        (st = upload(g_kb_types, s.types.data(), s.types.size())) ||
        (st = upload(g_kb_atoms, s.atoms.data(), s.atoms.size())))
        return st;
    const dim3 grid((unsigned)((mtxd + KB_THREADS - 1) / KB_THREADS), (unsigned)s.atoms.size());
    launch_timed(CPWGPU_K_PROJ_KB, k_proj_kb, grid, dim3(KB_THREADS), 0, g_stream, gw_dev, mtxd, s.cell,
                 (const KbAtom *)g_kb_atoms.p, (const KbType *)g_kb_types.p, (const double *)g_kb_tab.p,
                 s.vstride, anl, lda);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// ---- two-component H|psi> (Src/hk_psi/reference/hk_psi_spin_c16.f90) -------------
DevBuf g_spin_tmp;  // staging of vscr_sp

int set_potential_spin(cpwgpu_plan_s *p, const double *vscr_sp, int mxdscr, int nsp) {
    if (nsp != 1 && nsp != 2 && nsp != 4)
        return fail(CPWGPU_ERR_MODE, "nsp = %d is an incorrect value, should be 1,2,4", nsp);
    if (p->mwrk > mxdscr)
        return fail(CPWGPU_ERR_MXDSCR, "mxdwrk = %d is greater than mxdscr = %d", p->mwrk, mxdscr);
    if (p->ntot > (size_t)mxdscr)
        return fail(CPWGPU_ERR_MXDSCR, "mesh id*n2*n3 = %zu is greater than mxdscr = %d", p->ntot, mxdscr);
    int st = ensure_stream();
    if (st) return st;
    const long ntot = (long)p->ntot;
    const size_t ncomp = nsp == 1 ? 1 : nsp == 2 ? 2 : 4;
    if ((st = upload(g_spin_tmp, vscr_sp, (size_t)mxdscr * nsp)) || (st = p->vsp64.reserve(8 * ncomp * ntot)) ||
        (st = p->vsp32.reserve(4 * ncomp * ntot)))
        return st;
    const dim3 grid((unsigned)((ntot + SP_THREADS - 1) / SP_THREADS));
    hipLaunchKernelGGL(k_spin_potential<double>, grid, dim3(SP_THREADS), 0, g_stream, (const double *)g_spin_tmp.p,
                       (long)mxdscr, nsp, ntot, (double *)p->vsp64.p);
    hipLaunchKernelGGL(k_spin_potential<float>, grid, dim3(SP_THREADS), 0, g_stream, (const double *)g_spin_tmp.p,
                       (long)mxdscr, nsp, ntot, (float *)p->vsp32.p);
    HIPCHK(hipGetLastError());
    p->nsp = nsp;
    return CPWGPU_OK;
}

// anlsp(2*mxddim, nanlsp) complex128 host -> 2*mtxd rows on the device
int set_projectors_spin_host(cpwgpu_plan_s *p, const double *anlsp, const double *xnlkbsp, int nanlsp,
                             int mxddim) {
    if (nanlsp < 0) return fail(CPWGPU_ERR_ARG, "nanlsp = %d", nanlsp);
    if (nanlsp > 0 && (!anlsp || !xnlkbsp)) return fail(CPWGPU_ERR_ARG, "null projector array");
    int st = ensure_stream();
    if (st) return st;
    // as replace_projectors: the count changes only once the new set is in place
    p->sanl32_valid = false;
    if (nanlsp == 0) {
        p->snanl = 0;
        return CPWGPU_OK;
    }
    const long slda = std::max(1, 2 * p->mtxd);
    st = p->sanl64.reserve(16 * (size_t)slda * nanlsp);
    if (!st && p->mtxd > 0) {
        const hipError_t e = hipMemcpy2DAsync(p->sanl64.p, 16 * (size_t)slda, anlsp, 16 * (size_t)(2 * (size_t)mxddim),
                                              16 * (size_t)(2 * p->mtxd), nanlsp, hipMemcpyHostToDevice, g_stream);
        if (e != hipSuccess) st = fail(CPWGPU_ERR_HIP, "projector copy: %s", hipGetErrorString(e));
    }
    if (!st) st = upload(p->sxnl, xnlkbsp, nanlsp);
    p->snanl = st ? 0 : nanlsp;
    if (!st) p->slda = slda;
    return st;
}

// register shape of the coupled two-component plane kernel: (x lanes, y
// lanes, threads, waves per SIMD); both components' rows live in registers.
// FP64: the two sub-planes (about 86 KB at 72) allow one block per CU, so the
// block has 512 threads for 2 waves per SIMD (184 VGPRs fit that budget).
template <typename R> struct SpinShape;
template <> struct SpinShape<float> { static constexpr int GX = 8, GY = 8, T = 256, W = 3; };
template <> struct SpinShape<double> { static constexpr int GX = 8, GY = 8, T = 512, W = 2; };
template <typename R, int N> constexpr bool spin_reg_ok() {
    return reg_ok<R, N, SpinShape<R>::GX, SpinShape<R>::GY>();
}

// whether the register variant of the coupled plane kernel is built for the
// plan's mesh
template <typename R> bool spin_reg_built(const cpwgpu_plan_s *p) {
    const FftGeom &g = p->g;
    if (!fast_sizes(g)) return false;
    bool ok = false;
#define CPW_SR(N) if (g.n[0] == N) ok = spin_reg_ok<R, N>();
    CPW_FAST_SIZES(CPW_SR)
#undef CPW_SR
    return ok;
}

template <typename R> int spin_var_resolve(const cpwgpu_plan_s *p) {
    const int v = p->spin_var[sizeof(R) == 4 ? 1 : 0];
    if (!spin_reg_built<R>(p)) return 0;
    // measured on the 64-atom cell (64 spinor bands): register lines 11-12%
    // faster than the LDS two-pass kernel in both precisions
    return v < 0 ? 1 : v;
}

template <typename R, int N>
int launch_plane_spin_reg(cpwgpu_plan_s *p, cplx<R> *colbuf, const R *vr, long vstride, int npair,
                          hipStream_t s) {
    constexpr int GX = SpinShape<R>::GX, GY = SpinShape<R>::GY, T = SpinShape<R>::T, W = SpinShape<R>::W;
    if constexpr (!spin_reg_ok<R, N>()) {
        return fail(CPWGPU_ERR_MODE, "register two-component plane kernel is not built for N = %d", N);
    } else {
        const FftGeom &g = p->g;
        // a box inside N/4: the compile-time row stride
        const bool cs = prune_fits(g);
        const int S = cs ? pr_sub_stride<N>() : reg_stride(g.nx);
        const size_t dyn = sizeof(cplx<R>) * ((size_t)N + 2 * (size_t)N * S);
        if (dyn > 160 * 1024)
            return fail(CPWGPU_ERR_SIZE, "two-component plane of %d x %d does not fit in LDS", N, N);
        static bool attr = false;
        if (!attr) {
            HIPCHK(hipFuncSetAttribute((const void *)k_plane_spin_reg<R, N, GX, GY, T, W, false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            HIPCHK(hipFuncSetAttribute((const void *)k_plane_spin_reg<R, N, GX, GY, T, W, true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = true;
        }
        if (cs)
            launch_timed(CPWGPU_K_PLANE, (k_plane_spin_reg<R, N, GX, GY, T, W, true>), dim3(g.n[2] * npair), dim3(T), dyn, s,
                               colbuf, (const int *)p->col_uy.p, vr, vstride, (const double2 *)p->tw[0].p, g, npair, S);
        else
            launch_timed(CPWGPU_K_PLANE, (k_plane_spin_reg<R, N, GX, GY, T, W, false>), dim3(g.n[2] * npair), dim3(T), dyn, s,
                               colbuf, (const int *)p->col_uy.p, vr, vstride, (const double2 *)p->tw[0].p, g, npair, S);
        HIPCHK(hipGetLastError());
        return CPWGPU_OK;
    }
}

template <typename R, int N>
int launch_plane_spin(cpwgpu_plan_s *p, cplx<R> *colbuf, const R *vr, long vstride, int npair, hipStream_t s) {
    if (spin_var_resolve<R>(p) == 1) return launch_plane_spin_reg<R, N>(p, colbuf, vr, vstride, npair, s);
    const FftGeom &g = p->g;
    const size_t dyn = sizeof(cplx<R>) * (2 * (size_t)N * plane_sub_stride(g.nx) +
                                          (size_t)plane_rows_fast<N>() * row_stride<N>());
    const size_t stat = sizeof(cplx<R>) * 2 * N;
    if (dyn + stat > 160 * 1024)
        return fail(CPWGPU_ERR_SIZE, "two-component plane of %d x %d does not fit in LDS", N, N);
    static bool attr = false;
    if (!attr) {
        HIPCHK(hipFuncSetAttribute((const void *)k_plane_spin<R, N, N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(160 * 1024 - stat)));
        attr = true;
    }
    launch_timed(CPWGPU_K_PLANE, (k_plane_spin<R, N, N>), dim3(g.n[2] * npair), dim3(FT), dyn, s, colbuf,
                       (const int *)p->col_uy.p, vr, vstride, (const double2 *)p->tw[0].p,
                       (const double2 *)p->tw[1].p, g, npair);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// H psi_sp for neig two-component bands (rows 2i / 2i+1 = up / down of plane
// wave i): the components are split into scalar bands, the local part and
// the kinetic term go through the scalar column kernels (the potential per
// component, or the coupled plane kernel for nsp = 4), the non-local GEMMs run
// on the interleaved rows beside them, and the two are merged.
template <typename R>
int apply_spin(cpwgpu_plan_s *p, int neig, const double2 *psi, long ldp, double2 *hpsi, long ldh, hipStream_t s) {
    const FftGeom &g = p->g;
    const int mtxd = g.mtxd;
    if (!p->nsp) return fail(CPWGPU_ERR_ARG, "two-component potential not set");
    if (p->nsp == 4 && g.ncol > 0 && !fast_sizes(g))
        return fail(CPWGPU_ERR_SIZE, "the non-collinear potential needs one of the specialised meshes");
    if (neig == 0 || mtxd == 0) return CPWGPU_OK;
    int st;
    const size_t colb = 16 * (size_t)mtxd;
    if ((st = p->spsi.reserve(colb * 2 * neig)) || (st = p->shpsi.reserve(colb * 2 * neig))) return st;
    double2 *sp = (double2 *)p->spsi.p, *sh = (double2 *)p->shpsi.p;
    const dim3 eg((unsigned)((mtxd + SP_THREADS - 1) / SP_THREADS), (unsigned)neig);
    hipLaunchKernelGGL(k_spin_split, eg, dim3(SP_THREADS), 0, s, psi, ldp, mtxd, neig, sp);
    HIPCHK(hipGetLastError());
    const bool nl = p->snanl > 0;
    if (nl) {
        if ((st = ensure_aux())) return st;
        HIPCHK(hipEventRecord(g_fork, s));
        HIPCHK(hipStreamWaitEvent(g_aux, g_fork, 0));
        if ((st = nonlocal<R>(p, spin_proj(p), neig, (const double *)psi, ldp, (double *)hpsi, ldh, 0, false,
                              g_aux)))
            return st;
        HIPCHK(hipEventRecord(g_join, g_aux));
    }
    const R *vr = sizeof(R) == 8 ? (const R *)p->vsp64.p : (const R *)p->vsp32.p;
    const long ntot = (long)p->ntot;
    const size_t per_band = sizeof(cplx<R>) * (size_t)g.n[2] * g.ncolp;
    // scalar bands per pass: all of them if the column buffer stays under 4 GiB
    const int nbc = (int)std::max<size_t>(2, std::min<size_t>(2 * (size_t)neig, (4ull << 30) / per_band)) & ~1;
    if ((st = p->colbuf.reserve(per_band * nbc))) return st;
    cplx<R> *colbuf = (cplx<R> *)p->colbuf.p;
    auto columns = [&](int stage, int b0, int nb, cplx<R> *cbuf) -> int {
        const LocalArgs la{sp + (long)b0 * mtxd, mtxd, sh + (long)b0 * mtxd, mtxd, nb, COLF_KIN};
        return local_stage<R>(p, la, cbuf, vr, stage, s);
    };
    if (g.ncol > 0 && p->nsp != 4) {
        // potential diagonal in spin: scalar passes, up bands [0, neig) and down bands [neig, 2 neig)
        for (int b0 = 0; b0 < 2 * neig; b0 += nbc) {
            const int nb = std::min(nbc, 2 * neig - b0);
            if ((st = columns(ST_COL_BWD, b0, nb, colbuf))) return st;
            // bands of one component share a potential: split the plane pass at neig
            for (int c0 = b0; c0 < b0 + nb;) {
                const int comp = c0 < neig ? 0 : 1;
                const int c1 = comp == 0 ? std::min(b0 + nb, neig) : b0 + nb;
                const LocalArgs la{nullptr, 0, nullptr, 0, c1 - c0, COLF_KIN};
                const R *vc = vr + (p->nsp == 2 ? comp * ntot : 0);
                if ((st = local_stage<R>(p, la, colbuf + (size_t)(c0 - b0) * g.n[2] * g.ncolp, vc, ST_PLANE, s)))
                    return st;
                c0 = c1;
            }
            if ((st = columns(ST_COL_FWD, b0, nb, colbuf))) return st;
        }
    } else if (g.ncol > 0) {
        // coupled: pairs (up n, down n) in one plane kernel
        const int npc = nbc / 2;
        for (int n0 = 0; n0 < neig; n0 += npc) {
            const int np_ = std::min(npc, neig - n0);
            cplx<R> *cb_dn = colbuf + (size_t)np_ * g.n[2] * g.ncolp;
            if ((st = columns(ST_COL_BWD, n0, np_, colbuf)) || (st = columns(ST_COL_BWD, neig + n0, np_, cb_dn)))
                return st;
            st = CPWGPU_ERR_SIZE;
#define CPW_SPL(N) if (g.n[0] == N) st = launch_plane_spin<R, N>(p, colbuf, vr, ntot, np_, s);
            CPW_FAST_SIZES(CPW_SPL)
#undef CPW_SPL
            if (st) return st;
            if ((st = columns(ST_COL_FWD, n0, np_, colbuf)) || (st = columns(ST_COL_FWD, neig + n0, np_, cb_dn)))
                return st;
        }
    } else {
        // no z column (mtxd = 0 handled above): nothing local
    }
    if (nl) HIPCHK(hipStreamWaitEvent(s, g_join, 0));
    hipLaunchKernelGGL(k_spin_merge, eg, dim3(SP_THREADS), 0, s, (const double2 *)sh, mtxd, neig, hpsi, ldh,
                       nl ? 1 : 0);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// ---- drop-in state --------------------------------------------------------
// a plan kept between host-array calls while the plane-wave set is unchanged
struct PlanSlot {
    cpwgpu_plan_s *plan = nullptr;
    int key_ng = -1, key_mtxd = -1;
    int key_kmscr[7] = {0};
    void clear() {
        delete plan;
        plan = nullptr;
        key_mtxd = -1;
    }
};

struct DropIn {
    PlanSlot slot;
    DevBuf psi, hpsi, hs, hd, c8;
} g_drop;

// the slot's plan if it holds the same plane waves in the same order on the
// same mesh (ekpg refreshed when it changed), else a new plan in the slot
int slot_plan(PlanSlot &slot, int mtxd, int ng, const int *kgv, const double *ekpg, const int *isort,
              const int *kmscr, int mxddim, int mxdgve, cpwgpu_plan_s **out) {
    cpwgpu_plan_s *p = slot.plan;
    bool same = p && slot.key_mtxd == mtxd && slot.key_ng == ng &&
                std::equal(kmscr, kmscr + 7, slot.key_kmscr);
    if (same) {
        // same plane waves in the same order on the same mesh?
        std::vector<int> kx, ky, kz;
        int kd[3];
        int ns[3] = {kmscr[3], kmscr[4], kmscr[5]};
        if (ng > mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", ng, mxdgve);
        if (mtxd > mxddim) return fail(CPWGPU_ERR_MTXD, "mtxd = %d, mxddim = %d", mtxd, mxddim);
        int st = plane_wave_points(mtxd, ng, kgv, isort, ns, kx, ky, kz, kd);
        if (st) return st;
        for (int i = 0; i < mtxd && same; i++)
            same = p->key_pts[i] == ((long)kz[i] * ns[1] + ky[i]) * kmscr[6] + kx[i];
    }
    if (!same) {
        slot.clear();
        p = new cpwgpu_plan_s();
        int st = build_plan(p, mtxd, ng, kgv, ekpg, isort, kmscr, mxddim, mxdgve);
        if (st) {
            delete p;
            return st;
        }
        slot.plan = p;
        slot.key_mtxd = mtxd;
        slot.key_ng = ng;
        std::copy(kmscr, kmscr + 7, slot.key_kmscr);
    } else if (!std::equal(ekpg, ekpg + mtxd, p->h_ekpg.begin())) {
        p->h_ekpg.assign(ekpg, ekpg + mtxd);
        int st = upload(p->ekpg, ekpg, mtxd);
        if (st) return st;
    }
    *out = p;
    return CPWGPU_OK;
}

int drop_in_plan(int mtxd, int ng, const int *kgv, const double *ekpg, const int *isort,
                 const int *kmscr, int mxddim, int mxdgve, cpwgpu_plan_s **out) {
    return slot_plan(g_drop.slot, mtxd, ng, kgv, ekpg, isort, kmscr, mxddim, mxdgve, out);
}

// all-k host entry: one plan slot per k-point, device copies of psi / H psi
struct AllK {
    std::vector<PlanSlot> slots;
    DevBuf psi, hpsi;
} g_allk;

}  // namespace

namespace {

// ---- FP64 subspace workspace ------------------------------------------------
struct Subspace {
    DevBuf s, m, w, v, part, pos, rot, info, eig, res, tmp;
} g_ss;

// The subspace products use the direct-load FP64 GEMM shapes that measured
// fastest on the 64-atom cell (projection shape 2 at 0.72 and back-projection
// shape 2 at 0.67 of the FP64 matrix peak, against 0.35 / 0.49 for the
// LDS-staged kernels).
constexpr int SUB_PROJ_V = 2, SUB_BACK_V = 2;

// split-K partials of a^H W b (na x nbc over mtxd rows, W = diag(wrow) or
// the identity) into part; returns the number of slices in *nsl
int gram_partials(const double2 *a, long lda, int na, const double2 *b, long ldb, int nbc, int mtxd,
                  DevBuf &part, int *nsl_out, hipStream_t st, int lower = 0, const double *wrow = nullptr) {
    constexpr int BM = direct_bm<double, SUB_PROJ_V>(), BN = direct_bn<double, SUB_PROJ_V>();
    constexpr int BK = direct_bk<double, SUB_PROJ_V>();
    const int tm = (na + BM - 1) / BM, tn = (nbc + BN - 1) / BN;
    const int ksteps = (mtxd + BK - 1) / BK;
    int nsl = std::max(1, std::min(ksteps, (256 + tm * tn - 1) / (tm * tn)));
    // slices are whole K chunks of the kernel
    const int kslice = ((ksteps + nsl - 1) / nsl) * BK;
    nsl = (mtxd + kslice - 1) / kslice;
    int rc;
    if ((rc = part.reserve(16 * (size_t)nsl * na * nbc))) return rc;
    hipLaunchKernelGGL((k_proj_direct<double, SUB_PROJ_V>), dim3(tm, tn, nsl), dim3(GEMM_THREADS), 0, st,
                       (const cplx<double> *)a, lda, b, ldb, (cplx<double> *)part.p, mtxd, na, nbc, kslice,
                       wrow, lower);
    HIPCHK(hipGetLastError());
    *nsl_out = nsl;
    return CPWGPU_OK;
}

int rayleigh_ritz(int mtxd, int nb, double2 *psi, long ldp, double2 *hpsi, long ldh, int rotate,
                  double *eig_h, double *res_h, hipStream_t st) {
    int rc;
    const size_t mat = 16 * (size_t)nb * nb;
    if ((rc = g_ss.s.reserve(mat)) || (rc = g_ss.m.reserve(mat)) || (rc = g_ss.w.reserve(mat)) ||
        (rc = g_ss.v.reserve(mat)))
        return rc;
    const int np = (nb + 1) & ~1;
    if ((rc = g_ss.pos.reserve(4 * (size_t)np + 64)) || (rc = g_ss.rot.reserve(8 * (3 * (size_t)np + 8))) ||
        (rc = g_ss.info.reserve(64)) || (rc = g_ss.eig.reserve(8 * (size_t)nb)) ||
        (rc = g_ss.res.reserve(8 * (size_t)nb)))
        return rc;
    // Gram matrices on the matrix cores: split-K over the plane waves
    double2 *S = (double2 *)g_ss.s.p, *M = (double2 *)g_ss.m.p, *W = (double2 *)g_ss.w.p,
            *V = (double2 *)g_ss.v.p;
    for (int which = 0; which < 2; which++) {
        const double2 *y = which == 0 ? psi : hpsi;
        const long ldy = which == 0 ? ldp : ldh;
        int nsl;
        if ((rc = gram_partials(psi, ldp, nb, y, ldy, nb, mtxd, g_ss.part, &nsl, st))) return rc;
        launch_proj_reduce<double>((const cplx<double> *)g_ss.part.p, nullptr,
                                   (cplx<double> *)(which == 0 ? S : M), nb, nb, nsl, st);
    }
    int *dinfo = (int *)g_ss.info.p;
    hipLaunchKernelGGL(k_cholesky, dim3(1), dim3(SS_THREADS), 0, st, S, nb, dinfo);
    hipLaunchKernelGGL(k_to_standard, dim3(1), dim3(SS_THREADS), 0, st, (const double2 *)S, M, W, nb);
    // tol 1e-13: off-diagonal elements below 1e-13 |M| / nb are left, which moves
    // the Ritz values by at most ~1e-13 relative; a tighter tol sits at the
    // rounding-noise level the rotations themselves regenerate
    hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(SS_THREADS), 0, st, M, V, (double *)g_ss.eig.p, nb, 60,
                       1e-13, (int *)g_ss.pos.p, (double *)g_ss.rot.p, dinfo + 1);
    hipLaunchKernelGGL(k_back_substitute, dim3(1), dim3(SS_THREADS), 0, st, (const double2 *)S, V, nb);
    HIPCHK(hipGetLastError());
    int hinfo[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(hinfo, dinfo, sizeof hinfo, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hinfo[0]) return fail(CPWGPU_ERR_NOTPD, "overlap matrix not positive definite (pivot %d)", hinfo[0]);
    if (rotate) {
        if ((rc = g_ss.tmp.reserve(16 * (size_t)mtxd * nb))) return rc;
        double2 *T = (double2 *)g_ss.tmp.p;
        const long cnt = (long)mtxd * nb;
        for (int which = 0; which < 2; which++) {
            double2 *x = which == 0 ? psi : hpsi;
            const long ldx = which == 0 ? ldp : ldh;
            constexpr int BI = back_bi<double, SUB_BACK_V>(), BN = back_bn<double, SUB_BACK_V>();
            hipLaunchKernelGGL((k_back_direct<double, SUB_BACK_V>), dim3((mtxd + BI - 1) / BI, (nb + BN - 1) / BN),
                               dim3(GEMM_THREADS), 0, st, (const cplx<double> *)x, ldx,
                               (const cplx<double> *)V, T, (long)mtxd, mtxd, nb, nb, 0,
                               (const double2 *)nullptr, 0L, (const double *)nullptr);
            hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                               (const double2 *)T, (long)mtxd, x, ldx, mtxd, nb);
        }
    }
    if (res_h) {
        // residual of the (rotated) pairs; without rotation it uses psi, hpsi as given
        hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, (const double2 *)psi, ldp,
                           (const double2 *)hpsi, ldh, (const double *)g_ss.eig.p, (double *)g_ss.res.p,
                           mtxd);
        HIPCHK(hipMemcpyAsync(res_h, g_ss.res.p, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(eig_h, g_ss.eig.p, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CPWGPU_OK;
}

// ---- block iterative eigensolver (device-resident ditsp) ----------------------
// The lowest eigenpairs of the plan's H with everything on the device: the
// small-matrix start of Src/ditsp/ditsp_c16_start.f90 (zheevd of the leading
// block of H, its eigenvectors padded with zeros), then steps of a locally
// optimal block preconditioned iteration: the Jacobi-relaxed residuals of the
// vectors not yet converged (Src/omp/diag_rq_jac_c16.f90) and their previous
// search directions are orthonormalised against the vectors, H is applied to
// them, and a Rayleigh-Ritz step on [vectors, residuals, directions] keeps
// the lowest ones, until max_n |H psi_n - e_n psi_n|^2 < epspsi^2 over the
// wanted bands (the test of ditsp_c16.f90:375-377).  The reference relaxes
// each band with a fixed step between Rayleigh-Ritz steps instead; its
// convergence test, preconditioner and start are kept.  The
// orthonormalisation is Gram-Schmidt against the vectors plus Cholesky-QR of
// the new block (two passes each) instead of the reference's modified
// Gram-Schmidt; the small dense eigenproblems go to rocSOLVER's zheevd.
rocblas_handle g_blas = nullptr;

#define BLASCHK(x)                                                                         \
    do {                                                                                   \
        rocblas_status s_ = (x);                                                           \
        if (s_ != rocblas_status_success)                                                  \
            return fail(CPWGPU_ERR_HIP, "%s: rocBLAS/rocSOLVER status %d (%s:%d)", #x,     \
                        (int)s_, __FILE__, __LINE__);                                      \
    } while (0)

int ensure_blas(hipStream_t st) {
    if (!g_blas) {
        HIPCHK(hipSetDevice(g_device));
        BLASCHK(rocblas_create_handle(&g_blas));
    }
    BLASCHK(rocblas_set_stream(g_blas, st));
    return CPWGPU_OK;
}

struct DitspWs {
    DevBuf xa, hxa, xb, hxb, pb, cpk, hd, m, g, g0, pm, w, e, info, res, act, d, mones, eall;
} g_dt;

constexpr double DT_PIVOT = 1e-10;      // squared norm a unit column must keep outside the span
                                        // of the columns before it
constexpr double DT_SWITCH2 = 1e-10;    // FP32-mixed: residual^2 where FP64 H takes over
constexpr int DT_MAX_SMALL = 4096;      // largest small matrix of the start
constexpr unsigned long long DT_SEED = 0x243F6A8885A308D3ull;

// out(na, nbc) = a^H b over mtxd rows on the matrix cores (split-K over the
// plane waves, fixed-order reduction); row r of out is multiplied by scale[r]
// when scale is given.  lower != 0 (a Hermitian result, na == nbc): only the
// lower triangle is computed, the strict upper one is set to zero.
int gram(const double2 *a, long lda, int na, const double2 *b, long ldb, int nbc, int mtxd,
         const double *scale, double2 *out, hipStream_t st, int lower = 0) {
    int nsl, rc;
    if ((rc = gram_partials(a, lda, na, b, ldb, nbc, mtxd, g_ss.part, &nsl, st, lower))) return rc;
    launch_proj_reduce<double>((const cplx<double> *)g_ss.part.p, scale, (cplx<double> *)out, na,
                               nbc, nsl, st, lower);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// y(:, 0:nc) = beta*y + x(:, 0:nk) c; c column-major (nk, nc) with ld nk; beta 0 or 1
int rotate(const double2 *x, long ldx, int nk, const double2 *c, int nc, double2 *y, long ldy,
           int mtxd, int beta, hipStream_t st) {
    constexpr int BI = back_bi<double, SUB_BACK_V>(), BN = back_bn<double, SUB_BACK_V>();
    hipLaunchKernelGGL((k_back_direct<double, SUB_BACK_V>), dim3((mtxd + BI - 1) / BI, (nc + BN - 1) / BN),
                       dim3(GEMM_THREADS), 0, st, (const cplx<double> *)x, ldx, (const cplx<double> *)c, y,
                       ldy, mtxd, nk, nc, beta, (const double2 *)nullptr, 0L, (const double *)nullptr);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

inline dim3 dt_grid(int rows, int cols) {
    return dim3((unsigned)((rows + DT_THREADS - 1) / DT_THREADS), (unsigned)cols);
}

// Orthonormalises t(:, 0:m) (unit columns on entry) against the orthonormal
// x(:, 0:nx) and among themselves: Gram-Schmidt against x, twice, then
// Cholesky-QR of t, then once more of each.  In the first Cholesky pass the
// columns from the first one whose squared norm outside the span of x and of
// the columns before it is below DT_PIVOT are dropped; *kept is what is left.
int orthonormalize(int mtxd, const double2 *x, long ldx, int nx, double2 *t, long ldt, int m,
                   int *kept, hipStream_t st) {
    int rc;
    *kept = 0;
    if (m <= 0) return CPWGPU_OK;
    if ((rc = g_dt.pm.reserve(16 * (size_t)std::max(nx, 1) * m)) ||
        (rc = g_dt.g.reserve(16 * (size_t)m * m)) || (rc = g_dt.g0.reserve(16 * (size_t)m * m)) ||
        (rc = g_dt.info.reserve(64)) || (rc = g_dt.d.reserve(8 * (size_t)m)) ||
        (rc = g_dt.mones.reserve(8 * (size_t)std::max(nx, 1))))
        return rc;
    double2 *P = (double2 *)g_dt.pm.p, *G = (double2 *)g_dt.g.p, *G0 = (double2 *)g_dt.g0.p;
    int *dinfo = (int *)g_dt.info.p;
    if (nx > 0)
        hipLaunchKernelGGL(k_fill, dim3((nx + DT_THREADS - 1) / DT_THREADS), dim3(DT_THREADS), 0, st,
                           (double *)g_dt.mones.p, nx, -1.0);
    auto project = [&]() -> int {
        // P = -x^H t;  t += x P
        int r;
        if ((r = gram(x, ldx, nx, t, ldt, m, mtxd, (const double *)g_dt.mones.p, P, st))) return r;
        return rotate(x, ldx, nx, P, m, t, ldt, mtxd, 1, st);
    };
    const rocblas_double_complex one(1.0, 0.0);
    for (int pass = 0; pass < 2 && m > 0; pass++) {
        if (nx > 0) {
            if ((rc = project())) return rc;
            if (pass == 0 && (rc = project())) return rc;
        }
        const int ld = m;
        if ((rc = gram(t, ldt, m, t, ldt, m, mtxd, nullptr, G, st, 1))) return rc;
        HIPCHK(hipMemcpyAsync(G0, G, 16 * (size_t)ld * ld, hipMemcpyDeviceToDevice, st));
        BLASCHK(rocsolver_zpotrf(g_blas, rocblas_fill_lower, m, (rocblas_double_complex *)G, ld, dinfo));
        int info = 0;
        HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof info, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (info > 0) {
            // leading minor of order info not positive definite: keep the columns before it
            m = info - 1;
            if (m == 0) break;
            HIPCHK(hipMemcpyAsync(G, G0, 16 * (size_t)ld * ld, hipMemcpyDeviceToDevice, st));
            BLASCHK(rocsolver_zpotrf(g_blas, rocblas_fill_lower, m, (rocblas_double_complex *)G, ld, dinfo));
            HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof info, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (info != 0) return fail(CPWGPU_ERR_NOTPD, "Cholesky of a leading block failed (%d)", info);
        }
        if (pass == 0) {
            std::vector<double> d(m);
            hipLaunchKernelGGL(k_diag_sq, dim3((m + DT_THREADS - 1) / DT_THREADS), dim3(DT_THREADS), 0, st,
                               (const double2 *)G, (long)ld, m, (double *)g_dt.d.p);
            HIPCHK(hipMemcpyAsync(d.data(), g_dt.d.p, 8 * (size_t)m, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < m; k++)
                if (d[k] < DT_PIVOT) {
                    m = k;
                    break;
                }
            if (m == 0) break;
        }
        // t <- t L^-H
        BLASCHK(rocblas_ztrsm(g_blas, rocblas_side_right, rocblas_fill_lower,
                              rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, mtxd, m,
                              &one, (const rocblas_double_complex *)G, ld, (rocblas_double_complex *)t,
                              (int)ldt));
    }
    *kept = m;
    return CPWGPU_OK;
}

// Rayleigh-Ritz on the orthonormal basis xa(:, 0:nt) with hxa = H xa (ld
// mtxd): the lowest nk Ritz vectors go to xo and H times them to hxo (ld
// mtxd); all nt Ritz values to g_dt.w (ascending), zheevd's info to
// g_dt.info[1].  When po is given and nt > nk, po receives the part of the
// Ritz vectors from the columns beyond nk, xa(:, nk:nt) C(nk:nt, 0:nk): the
// search directions of the next step.
int ritz_step(int mtxd, const double2 *xa, const double2 *hxa, int nt, int nk, double2 *xo,
              double2 *hxo, double2 *po, hipStream_t st) {
    int rc;
    double2 *M = (double2 *)g_dt.m.p;
    if ((rc = gram(xa, mtxd, nt, hxa, mtxd, nt, mtxd, nullptr, M, st, 1))) return rc;
    BLASCHK(rocsolver_zheevd(g_blas, rocblas_evect_original, rocblas_fill_lower, nt,
                             (rocblas_double_complex *)M, nt, (double *)g_dt.w.p, (double *)g_dt.e.p,
                             (int *)g_dt.info.p + 1));
    if ((rc = rotate(xa, mtxd, nt, M, nk, xo, mtxd, mtxd, 0, st))) return rc;
    if ((rc = rotate(hxa, mtxd, nt, M, nk, hxo, mtxd, mtxd, 0, st))) return rc;
    if (po && nt > nk) {
        double2 *cp = (double2 *)g_dt.cpk.p;
        hipLaunchKernelGGL(k_pack_rows, dt_grid(nt - nk, nk), dim3(DT_THREADS), 0, st, (const double2 *)M,
                           (long)nt, nk, nt - nk, nk, cp);
        if ((rc = rotate(xa + (size_t)nk * mtxd, mtxd, nt - nk, cp, nk, po, mtxd, mtxd, 0, st))) return rc;
    }
    return CPWGPU_OK;
}

// mtxds from the diagonal of H, the rule of Src/ALFA/size_mtxds.f90:60-85
int size_mtxds(const std::vector<double> &hd, int neig) {
    const int mtxd = (int)hd.size();
    constexpr int NFULL = 250;
    constexpr double EPS = 1e-6;
    if (mtxd < NFULL || 3 * neig + 20 >= mtxd) return mtxd;
    const double href = std::fabs(2 * hd[neig - 1] - hd[0]) + EPS;
    for (int i = 3 * neig + 20; i <= mtxd - 1; i++)  // 1-based i as in the reference
        if (hd[i - 1] > href && hd[i] - hd[i - 1] > EPS) return i;
    return mtxd;
}

// Largest eigenvalue of the symmetric tridiagonal (a, b) (b[j] couples j, j+1):
// bisection on the Sturm count in the Gershgorin interval
double tridiag_max(const std::vector<double> &a, const std::vector<double> &b) {
    const int k = (int)a.size();
    double lo = HUGE_VAL, hi = -HUGE_VAL;
    for (int j = 0; j < k; j++) {
        const double r = (j > 0 ? std::fabs(b[j - 1]) : 0.0) + (j + 1 < k ? std::fabs(b[j]) : 0.0);
        lo = std::min(lo, a[j] - r);
        hi = std::max(hi, a[j] + r);
    }
    // number of eigenvalues below x
    auto below = [&](double x) {
        int n = 0;
        double d = 1.0;
        for (int j = 0; j < k; j++) {
            d = a[j] - x - (j > 0 ? b[j - 1] * b[j - 1] / d : 0.0);
            if (d == 0.0) d = -1e-300;
            if (d < 0) n++;
        }
        return n;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); it++) {
        const double mid = 0.5 * (lo + hi);
        if (below(mid) == k) hi = mid;
        else lo = mid;
    }
    return hi;
}

// Upper bound of the spectrum of H: DT_LANCZOS Lanczos steps (FP64 H) from a
// random vector; the largest Ritz value plus the norm of the last residual.
// buf: three scratch columns (ld mtxd).
constexpr int DT_LANCZOS = 12;
template <typename HA>
int lanczos_upper(HA &happly, cpwgpu_plan_s *p, int mtxd, double2 *buf, hipStream_t st, double *bup) {
    int rc;
    double2 *v = buf, *w = buf + mtxd, *vp = buf + 2 * (size_t)mtxd;
    double2 *dot = (double2 *)g_dt.m.p;
    hipLaunchKernelGGL(k_random_cols, dt_grid(mtxd, 1), dim3(DT_THREADS), 0, st, v, (long)mtxd, mtxd, 0, 1,
                       (const double *)p->ekpg.p, DT_SEED ^ 0x9E3779B97F4A7C15ull);
    hipLaunchKernelGGL(k_normalize_cols, dim3(1), dim3(DT_THREADS), 0, st, v, (long)mtxd, mtxd);
    std::vector<double> al, be;
    double bprev = 0.0;
    for (int j = 0; j < std::min(DT_LANCZOS, mtxd); j++) {
        if ((rc = happly(CPWGPU_PREC_FP64, v, w, 1))) return rc;
        double2 h;
        if ((rc = gram(v, mtxd, 1, w, mtxd, 1, mtxd, nullptr, dot, st))) return rc;
        HIPCHK(hipMemcpyAsync(&h, dot, sizeof h, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const double a = h.x;
        // w -= a v + bprev vp
        hipLaunchKernelGGL(k_cheb, dt_grid(mtxd, 1), dim3(DT_THREADS), 0, st, w, (const double2 *)w,
                           (const double2 *)v, (const double2 *)vp, (long)mtxd, mtxd, a, 1.0, -bprev);
        if ((rc = gram(w, mtxd, 1, w, mtxd, 1, mtxd, nullptr, dot, st))) return rc;
        HIPCHK(hipMemcpyAsync(&h, dot, sizeof h, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const double b = std::sqrt(std::max(h.x, 0.0));
        al.push_back(a);
        if (b <= 1e-12 * std::fabs(a) || j + 1 == std::min(DT_LANCZOS, mtxd)) {
            be.push_back(b);
            break;
        }
        be.push_back(b);
        // vp <- w / b, then (v, vp) swap roles
        hipLaunchKernelGGL(k_cheb, dt_grid(mtxd, 1), dim3(DT_THREADS), 0, st, vp, (const double2 *)w,
                           (const double2 *)w, (const double2 *)w, (long)mtxd, mtxd, 0.0, 1.0 / b, 0.0);
        std::swap(v, vp);
        bprev = b;
    }
    HIPCHK(hipGetLastError());
    std::vector<double> off(be.begin(), be.end() - 1);
    *bup = tridiag_max(al, off) + be.back();
    return CPWGPU_OK;
}

// Degree of the Chebyshev filter.  The filter grows the lowest Ritz value's
// direction by T_m(|l0|) (l0 = its image in the mapped interval) against at
// most 1 for the damped part; the degree keeps that ratio near DT_CHEB_AMP, so
// the Gram matrix of the filtered block stays well conditioned for the
// Cholesky-QR.  CPWGPU_DITSP_DEGREE (4..40) fixes it instead.
constexpr double DT_CHEB_AMP = 1e4;
int cheb_degree(double l0) {
    const char *env = getenv("CPWGPU_DITSP_DEGREE");
    const int fixed = env && *env ? atoi(env) : 0;
    if (fixed >= 4 && fixed <= 40) return fixed;
    const double x = std::max(std::fabs(l0), 1.0 + 1e-9);
    const int m = (int)(std::acosh(DT_CHEB_AMP) / std::acosh(x));
    return std::min(24, std::max(4, m));
}

// One step of Chebyshev-filtered subspace iteration (Zhou and Saad, the
// scaled three-term recurrence): the nb vectors X (HX = H X, Ritz values E,
// host copy wv) are filtered by the degree-m Chebyshev polynomial that damps
// [wv[nb-1], bup] and grows below it, orthonormalised (Cholesky-QR, columns
// lost to dependence replaced by random ones) and rotated to the Ritz vectors
// of H in their span: X, HX, E on return.  Work space: columns nb..3nb of
// X and HX, and XB, HXB.
template <typename HA>
int cheb_filter_step(HA &happly, cpwgpu_plan_s *p, int pc, int mtxd, int nb, int neig, double2 *X, double2 *HX,
                     double2 *XB, double2 *HXB, double *E, const std::vector<double> &wv, double &bup,
                     hipStream_t st) {
    int rc;
    const size_t col = 16 * (size_t)mtxd;
    const double a0 = wv[0], a = wv[nb - 1];
    // a spectrum bound below the block's top Ritz value cannot be right: widen it
    if (!(bup > a + 1e-6 * std::max(1.0, std::fabs(a)))) bup = a + std::max(1.0, a - a0);
    const double e = 0.5 * (bup - a), c = 0.5 * (bup + a);
    double sigma = e / (a0 - c);
    const double tau = 2.0 / sigma;
    const int deg = cheb_degree((a0 - c) / e);
    double2 *B0 = X, *B1 = X + (size_t)nb * mtxd, *B2 = X + 2 * (size_t)nb * mtxd;
    double2 *HB = HX + (size_t)nb * mtxd;
    const dim3 grid = dt_grid(mtxd, nb);
    // y1 = (H x - c x) sigma / e
    hipLaunchKernelGGL(k_cheb, grid, dim3(DT_THREADS), 0, st, B1, (const double2 *)HX, (const double2 *)B0,
                       (const double2 *)B0, (long)mtxd, mtxd, c, sigma / e, 0.0);
    for (int i = 2; i <= deg; i++) {
        if ((rc = happly(pc, B1, HB, nb))) return rc;
        const double s1 = 1.0 / (tau - sigma);
        hipLaunchKernelGGL(k_cheb, grid, dim3(DT_THREADS), 0, st, B2, (const double2 *)HB, (const double2 *)B1,
                           (const double2 *)B0, (long)mtxd, mtxd, c, 2.0 * s1 / e, -sigma * s1);
        double2 *t = B0;
        B0 = B1;
        B1 = B2;
        B2 = t;
        sigma = s1;
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_normalize_cols, dim3(nb), dim3(DT_THREADS), 0, st, B1, (long)mtxd, mtxd);
    int kept = 0;
    if ((rc = orthonormalize(mtxd, nullptr, mtxd, 0, B1, mtxd, nb, &kept, st))) return rc;
    if (kept < nb) {
        hipLaunchKernelGGL(k_random_cols, dt_grid(mtxd, nb - kept), dim3(DT_THREADS), 0, st, B1, (long)mtxd,
                           mtxd, kept, nb - kept, (const double *)p->ekpg.p, DT_SEED + (unsigned long long)kept);
        hipLaunchKernelGGL(k_normalize_cols, dim3(nb - kept), dim3(DT_THREADS), 0, st, B1 + (size_t)kept * mtxd,
                           (long)mtxd, mtxd);
        int k2 = 0;
        if ((rc = orthonormalize(mtxd, B1, mtxd, kept, B1 + (size_t)kept * mtxd, mtxd, nb - kept, &k2, st)))
            return rc;
        if (k2 < nb - kept) return fail(CPWGPU_ERR_NOTPD, "filtered vectors are linearly dependent");
    }
    if ((rc = happly(pc, B1, HB, nb))) return rc;
    if ((rc = ritz_step(mtxd, B1, HB, nb, nb, XB, HXB, nullptr, st))) return rc;
    HIPCHK(hipMemcpyAsync(X, XB, col * nb, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(HX, HXB, col * nb, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(E, g_dt.w.p, 8 * (size_t)nb, hipMemcpyDeviceToDevice, st));
    (void)neig;
    return CPWGPU_OK;
}

int plan_ditsp(cpwgpu_plan_s *p, int prec, int icmax, int iguess, double epspsi, int neig, int mtxds,
               const double2 *hamsm, long ldhs, const double *hdiag_in, double2 *psi, long ldp,
               double2 *hpsi, long ldh, double *ei, int *ifail, int *stats, hipStream_t st) {
    const int mtxd = p->mtxd;
    int rc;
    if (!p->v64.p) return fail(CPWGPU_ERR_ARG, "the plan has no potential");
    if ((rc = ensure_blas(st))) return rc;
    long ihkp = 0;
    auto happly = [&](int pr, const double2 *x, double2 *hx, int n) -> int {
        ihkp += n;
        return pr == CPWGPU_PREC_FP32
                   ? apply_t<float>(p, n, (const double *)x, mtxd, (double *)hx, mtxd, st)
                   : apply_t<double>(p, n, (const double *)x, mtxd, (double *)hx, mtxd, st);
    };
    const int nbuf = std::min(mtxd - neig, std::max(8, neig / 8));
    const int nb = neig + nbuf;
    // diagonal of H (preconditioner and the choice of mtxds)
    if ((rc = g_dt.hd.reserve(8 * (size_t)mtxd))) return rc;
    double *hd = (double *)g_dt.hd.p;
    if (hdiag_in) {
        HIPCHK(hipMemcpyAsync(hd, hdiag_in, 8 * (size_t)mtxd, hipMemcpyDeviceToDevice, st));
    } else {
        hipLaunchKernelGGL(k_hdiag, dim3((mtxd + DT_THREADS - 1) / DT_THREADS), dim3(DT_THREADS), 0, st,
                           (const double *)p->ekpg.p, (const double2 *)p->anl64.p, p->lda,
                           (const double *)p->xnl.p, p->nanl, p->v0, hd, mtxd);
        HIPCHK(hipGetLastError());
    }
    if (mtxds <= 0) {
        std::vector<double> h(mtxd);
        HIPCHK(hipMemcpyAsync(h.data(), hd, 8 * (size_t)mtxd, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        mtxds = size_mtxds(h, neig);
    }
    mtxds = std::min(mtxds, mtxd);
    if (neig > mtxds) return fail(CPWGPU_ERR_ARG, "neig = %d is greater than mtxds = %d", neig, mtxds);
    const bool full = mtxds >= mtxd;
    const bool small_start = full || iguess != 1;
    if (small_start && mtxds > DT_MAX_SMALL)
        return fail(CPWGPU_ERR_SIZE, "small matrix of %d > %d rows", mtxds, DT_MAX_SMALL);
    const int ms = small_start ? mtxds : 0;
    const int cap = std::max(3 * nb, small_start && !hamsm ? ms : 0);
    const int mdim = std::max(3 * nb, ms);
    const size_t col = 16 * (size_t)mtxd;
    if ((rc = g_dt.xa.reserve(col * cap)) || (rc = g_dt.hxa.reserve(col * cap)) ||
        (rc = g_dt.xb.reserve(col * nb)) || (rc = g_dt.hxb.reserve(col * nb)) ||
        (rc = g_dt.pb.reserve(col * nb)) || (rc = g_dt.cpk.reserve(16 * (size_t)2 * nb * nb)) ||
        (rc = g_dt.m.reserve(16 * (size_t)mdim * mdim)) || (rc = g_dt.w.reserve(8 * (size_t)mdim)) ||
        (rc = g_dt.e.reserve(8 * (size_t)mdim)) || (rc = g_dt.info.reserve(64)) ||
        (rc = g_dt.res.reserve(8 * (size_t)nb)) || (rc = g_dt.act.reserve(8 * (size_t)nb)) ||
        (rc = g_dt.eall.reserve(8 * (size_t)nb)))
        return rc;
    double2 *X = (double2 *)g_dt.xa.p, *HX = (double2 *)g_dt.hxa.p;
    double2 *XB = (double2 *)g_dt.xb.p, *HXB = (double2 *)g_dt.hxb.p, *PB = (double2 *)g_dt.pb.p;
    double2 *M = (double2 *)g_dt.m.p;
    double *W = (double *)g_dt.w.p;
    int *dinfo = (int *)g_dt.info.p;
    HIPCHK(hipMemsetAsync(dinfo, 0, 64, st));
    int refine_at = -1, iters = 0;
    auto finish = [&](int conv, double xmax) {
        if (ifail) {
            int f = 0;
            if (!conv) {
                f = 1 + (int)(-std::log10(xmax));  // Src/ditsp/ditsp_c16.f90:408-409
                if (f == 0) f = 1;
            }
            *ifail = f;
        }
        if (stats) {
            stats[0] = iters;
            stats[1] = (int)ihkp;
            stats[2] = nb;
            stats[3] = mtxds;
            stats[4] = refine_at;
        }
    };
    if (small_start) {
        // small matrix: given, or the leading block of H from its action on unit vectors
        if (hamsm) {
            HIPCHK(hipMemcpy2DAsync(M, 16 * (size_t)ms, hamsm, 16 * (size_t)ldhs, 16 * (size_t)ms, ms,
                                    hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL(k_unit_cols, dt_grid(mtxd, ms), dim3(DT_THREADS), 0, st, X, (long)mtxd, mtxd, ms);
            if ((rc = happly(CPWGPU_PREC_FP64, X, HX, ms))) return rc;
            hipLaunchKernelGGL(k_leading_block, dt_grid(ms, ms), dim3(DT_THREADS), 0, st,
                               (const double2 *)HX, (long)mtxd, ms, M);
        }
        BLASCHK(rocsolver_zheevd(g_blas, rocblas_evect_original, rocblas_fill_lower, ms,
                                 (rocblas_double_complex *)M, ms, W, (double *)g_dt.e.p, dinfo + 1));
        int info = 0;
        HIPCHK(hipMemcpyAsync(&info, dinfo + 1, sizeof info, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (info) return fail(CPWGPU_ERR_HIP, "zheevd of the small matrix failed (info %d)", info);
        if (full) {
            // the whole H was diagonalised (Src/ditsp/ditsp_c16.f90:218-245)
            hipLaunchKernelGGL(k_pad_rows, dt_grid(mtxd, neig), dim3(DT_THREADS), 0, st, (const double2 *)M,
                               (long)ms, ms, X, (long)mtxd, mtxd, neig);
            if ((rc = happly(CPWGPU_PREC_FP64, X, HX, neig))) return rc;
            HIPCHK(hipMemcpy2DAsync(psi, 16 * (size_t)ldp, X, col, col, neig, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpy2DAsync(hpsi, 16 * (size_t)ldh, HX, col, col, neig, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(ei, W, 8 * (size_t)neig, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            finish(1, 0.0);
            return CPWGPU_OK;
        }
        const int nu = std::min(nb, ms);
        hipLaunchKernelGGL(k_pad_rows, dt_grid(mtxd, nu), dim3(DT_THREADS), 0, st, (const double2 *)M,
                           (long)ms, ms, X, (long)mtxd, mtxd, nu);
        if (nu < nb) {
            hipLaunchKernelGGL(k_random_cols, dt_grid(mtxd, nb - nu), dim3(DT_THREADS), 0, st, X, (long)mtxd,
                               mtxd, nu, nb - nu, (const double *)p->ekpg.p, DT_SEED);
            hipLaunchKernelGGL(k_normalize_cols, dim3(nb - nu), dim3(DT_THREADS), 0, st, X + (size_t)nu * mtxd,
                               (long)mtxd, mtxd);
            int kept = 0;
            if ((rc = orthonormalize(mtxd, X, mtxd, nu, X + (size_t)nu * mtxd, mtxd, nb - nu, &kept, st)))
                return rc;
            if (kept < nb - nu) return fail(CPWGPU_ERR_NOTPD, "start vectors are linearly dependent");
        }
    } else {
        // neig guesses from the caller, buffer vectors random
        HIPCHK(hipMemcpy2DAsync(X, col, psi, 16 * (size_t)ldp, col, neig, hipMemcpyDeviceToDevice, st));
        if (nb > neig)
            hipLaunchKernelGGL(k_random_cols, dt_grid(mtxd, nb - neig), dim3(DT_THREADS), 0, st, X,
                               (long)mtxd, mtxd, neig, nb - neig, (const double *)p->ekpg.p, DT_SEED);
        hipLaunchKernelGGL(k_normalize_cols, dim3(nb), dim3(DT_THREADS), 0, st, X, (long)mtxd, mtxd);
        int kept = 0;
        if ((rc = orthonormalize(mtxd, X, mtxd, 0, X, mtxd, nb, &kept, st))) return rc;
        if (kept < nb) return fail(CPWGPU_ERR_NOTPD, "guess vectors are linearly dependent");
    }
    int pc = prec;
    bool refined = prec == CPWGPU_PREC_FP64;
    // Locked vectors: X(:, 0:nl), the lowest converged ones (their residual^2
    // below 0.01 epspsi^2); they stay fixed, new directions are still made
    // orthogonal to them, and each Rayleigh-Ritz step only diagonalises
    // [X(:, nl:nb), new block].  E holds the Ritz value of every column of X.
    int nl = 0;
    bool havep = false;
    int pbase = 0;  // column j of PB is the direction of vector pbase + j
    double *E = (double *)g_dt.eall.p;
    int *dact = (int *)g_dt.act.p, *dactp = dact + nb;
    auto ritz_free = [&](int kept) -> int {
        const int nu = nb - nl;
        double2 *Xu = X + (size_t)nl * mtxd, *HXu = HX + (size_t)nl * mtxd;
        int r = ritz_step(mtxd, Xu, HXu, nu + kept, nu, XB, HXB, kept > 0 ? PB : nullptr, st);
        if (r) return r;
        havep = kept > 0;
        pbase = nl;
        HIPCHK(hipMemcpyAsync(Xu, XB, col * nu, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(HXu, HXB, col * nu, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(E + nl, W, 8 * (size_t)nu, hipMemcpyDeviceToDevice, st));
        return CPWGPU_OK;
    };
    if ((rc = happly(pc, X, HX, nb)) || (rc = ritz_free(0))) return rc;
    const double eps2 = epspsi * epspsi;
    const double sw2 = std::max(eps2, DT_SWITCH2);
    // CPWGPU_DITSP_TRACE=1: one line per iteration on stderr (the ipr > 3
    // diagnostics of Src/ditsp/ditsp_c16.f90:306-336)
    static const bool trace = [] {
        const char *e = getenv("CPWGPU_DITSP_TRACE");
        return e && *e && *e != '0';
    }();
    std::vector<double> res(nb, HUGE_VAL), wv(nb);
    std::vector<int> act, actp;
    act.reserve(nb);
    actp.reserve(nb);
    double xmax = 0.0, xprev = HUGE_VAL;
    bool conv = false;
    // Chebyshev-filtered subspace iteration: an upper bound of the spectrum
    // from a few Lanczos steps (FP64 H on one vector)
    const bool cheb = g_ditsp_method == CPWGPU_DITSP_CHEB;
    double bup = 0.0;
    if (cheb && (rc = lanczos_upper(happly, p, mtxd, XB, st, &bup))) return rc;
    for (int ic = 0;; ic++) {
        const int nu = nb - nl;
        // residuals of the unlocked vectors; the locked ones keep theirs
        hipLaunchKernelGGL(k_residual, dim3(nu), dim3(256), 0, st, (const double2 *)(X + (size_t)nl * mtxd),
                           (long)mtxd, (const double2 *)(HX + (size_t)nl * mtxd), (long)mtxd,
                           (const double *)(E + nl), (double *)g_dt.res.p + nl, mtxd);
        HIPCHK(hipGetLastError());
        int info = 0;
        HIPCHK(hipMemcpyAsync(res.data() + nl, (double *)g_dt.res.p + nl, 8 * (size_t)nu, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(wv.data(), E, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&info, dinfo + 1, sizeof info, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (info) return fail(CPWGPU_ERR_HIP, "zheevd of the subspace failed (info %d)", info);
        iters = ic;
        xmax = *std::max_element(res.begin(), res.begin() + neig);
        if (trace)
            fprintf(stderr, "ditsp: ic %4d  max res^2 %.3e  locked %d  %s\n", ic, xmax, nl, refined ? "fp64" : "fp32");
        if (!refined) {
            // the FP32 products stop helping near their rounding level: switch
            // there, or once the residual stalls below 1e-6
            if (xmax < sw2 || (xmax < 1e-6 && xmax > 0.9 * xprev)) {
                // FP32-mixed: unlock everything, FP64 H on the current vectors, then FP64 iterations
                refined = true;
                refine_at = ic;
                pc = CPWGPU_PREC_FP64;
                nl = 0;
                std::fill(res.begin(), res.end(), HUGE_VAL);
                if ((rc = happly(pc, X, HX, nb)) || (rc = ritz_free(0))) return rc;
                xprev = HUGE_VAL;
                continue;
            }
        } else if (xmax < eps2) {
            conv = true;
            break;
        }
        if (ic >= icmax) break;
        if (cheb) {
            if ((rc = cheb_filter_step(happly, p, pc, mtxd, nb, neig, X, HX, XB, HXB, E, wv, bup, st))) return rc;
            xprev = xmax;
            continue;
        }
        const double thr = 0.01 * (refined ? eps2 : sw2);
        // lock the lowest unlocked vectors that have converged
        while (nl < nb - 1 && res[nl] <= thr) nl++;
        // active vectors (unlocked, not converged), largest residual first
        act.clear();
        actp.clear();
        for (int n = nl; n < nb; n++)
            if (res[n] > thr) act.push_back(n);
        std::stable_sort(act.begin(), act.end(), [&](int a, int b) { return res[a] > res[b]; });
        const int m = (int)act.size();
        if (m == 0) break;
        for (int n : act) actp.push_back(n - pbase);
        HIPCHK(hipMemcpyAsync(dact, act.data(), 4 * (size_t)m, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dactp, actp.data(), 4 * (size_t)m, hipMemcpyHostToDevice, st));
        double2 *T = X + (size_t)nb * mtxd, *HT = HX + (size_t)nb * mtxd;
        hipLaunchKernelGGL(k_precondition, dt_grid(mtxd, m), dim3(DT_THREADS), 0, st, (const double2 *)X,
                           (long)mtxd, (const double2 *)HX, (long)mtxd, (const double *)E, (const double *)hd,
                           (const int *)dact, T, (long)mtxd, mtxd, m);
        // previous search directions of the active vectors (the third block
        // of a locally optimal block preconditioned iteration)
        int mt = m;
        if (havep) {
            hipLaunchKernelGGL(k_gather_cols, dt_grid(mtxd, m), dim3(DT_THREADS), 0, st, (const double2 *)PB,
                               (long)mtxd, (const int *)dactp, T + (size_t)m * mtxd, (long)mtxd, mtxd, m);
            mt = 2 * m;
        }
        hipLaunchKernelGGL(k_normalize_cols, dim3(mt), dim3(DT_THREADS), 0, st, T, (long)mtxd, mtxd);
        int kept = 0;
        if ((rc = orthonormalize(mtxd, X, mtxd, nb, T, mtxd, mt, &kept, st))) return rc;
        if (trace) fprintf(stderr, "ditsp:        active %d  new directions %d of %d\n", m, kept, mt);
        if (kept == 0) break;  // no direction left outside the current vectors
        if ((rc = happly(pc, T, HT, kept)) || (rc = ritz_free(kept))) return rc;
        xprev = xmax;
    }
    // the neig lowest pairs in ascending order (a vector found late can lie below a locked one)
    std::vector<int> order(nb);
    for (int n = 0; n < nb; n++) order[n] = n;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return wv[a] < wv[b]; });
    bool ident = true;
    for (int n = 0; n < neig; n++) ident = ident && order[n] == n;
    if (ident) {
        HIPCHK(hipMemcpy2DAsync(psi, 16 * (size_t)ldp, X, col, col, neig, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpy2DAsync(hpsi, 16 * (size_t)ldh, HX, col, col, neig, hipMemcpyDeviceToDevice, st));
    } else {
        HIPCHK(hipMemcpyAsync(dact, order.data(), 4 * (size_t)neig, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_gather_cols, dt_grid(mtxd, neig), dim3(DT_THREADS), 0, st, (const double2 *)X,
                           (long)mtxd, (const int *)dact, psi, ldp, mtxd, neig);
        hipLaunchKernelGGL(k_gather_cols, dt_grid(mtxd, neig), dim3(DT_THREADS), 0, st, (const double2 *)HX,
                           (long)mtxd, (const int *)dact, hpsi, ldh, mtxd, neig);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    for (int n = 0; n < neig; n++) ei[n] = wv[order[n]];
    finish(conv, xmax);
    return CPWGPU_OK;
}

// ---- atomic forces ------------------------------------------------------------
struct ForceWs {
    DevBuf t2, t2s, rho, qv, fq, rat, species, patom, occ, fl, fnl, enl, p, rhor;
} g_fw;

// register shape of the density pass: x lanes, y lanes, waves per SIMD
constexpr int DEN_GX = 8, DEN_GY = 8, DEN_W = 3;
// threads of the register density pass for line length N: the smallest of 256,
// 320 and 384 that keeps a lane's row accumulators (rows per x task x N / GX
// doubles) within 27, for lines of at most 10 values per lane: spill-free at 3
// waves per SIMD (24-72: 256, 80: 320); 0 = not built (96 and 128 spill, 108 has
// no 8-lane split: k_density_fast)
template <int N> constexpr int den_threads() {
    if (!reg_ok<double, N, DEN_GX, DEN_GY>() || N / DEN_GX > 10) return 0;
    for (int t : {256, 320, 384})
        if (((N + t / DEN_GX - 1) / (t / DEN_GX)) * (N / DEN_GX) <= 27) return t;
    return 0;
}
template <int N> constexpr bool den_reg_ok() { return den_threads<N>() > 0; }

template <int N>
int launch_density(cpwgpu_plan_s *p, const cplx<double> *colbuf, const double *occ, int nb, int bpg,
                   int ngrp, int Q1, int Q2, double2 *t2, hipStream_t s, double *rhor = nullptr) {
    const FftGeom &g = p->g;
    if constexpr (den_reg_ok<N>()) {
        if (rhor) {
            // the density planes alone: register lines, accumulation in registers
            constexpr int DT = den_threads<N>();
            // a box inside N/4: the compile-time row stride
            const bool cs = prune_fits(g);
            const int S = cs ? pr_sub_stride<N>() : reg_stride(g.nx);
            const size_t dyn = 16 * ((size_t)N + (size_t)N * S);
            if (dyn > 160 * 1024) return fail(CPWGPU_ERR_SIZE, "density plane %d does not fit in LDS", N);
            static bool attr_r = false;
            if (!attr_r) {
                HIPCHK(hipFuncSetAttribute((const void *)k_density_reg<N, DEN_GX, DEN_GY, DT, DEN_W, false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                HIPCHK(hipFuncSetAttribute((const void *)k_density_reg<N, DEN_GX, DEN_GY, DT, DEN_W, true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                attr_r = true;
            }
            if (cs)
                hipLaunchKernelGGL((k_density_reg<N, DEN_GX, DEN_GY, DT, DEN_W, true>), dim3(g.n[2] * ngrp),
                                   dim3(DT), dyn, s, colbuf, (const int *)p->col_uy.p, occ,
                                   (const double2 *)p->tw[0].p, g, nb, bpg, S, rhor);
            else
                hipLaunchKernelGGL((k_density_reg<N, DEN_GX, DEN_GY, DT, DEN_W, false>), dim3(g.n[2] * ngrp),
                                   dim3(DT), dyn, s, colbuf, (const int *)p->col_uy.p, occ,
                                   (const double2 *)p->tw[0].p, g, nb, bpg, S, rhor);
            HIPCHK(hipGetLastError());
            return CPWGPU_OK;
        }
    }
    constexpr int RB = plane_rows_fast<N>();
    const size_t dyn = 16 * ((size_t)N * (g.nx | 1) + (size_t)RB * row_stride<N>()) + 8 * (size_t)N * N;
    const size_t stat = 16 * 2 * (size_t)N;
    if (dyn + stat > 160 * 1024) return fail(CPWGPU_ERR_SIZE, "density plane %d does not fit in LDS", N);
    static bool attr = false;
    if (!attr) {
        HIPCHK(hipFuncSetAttribute((const void *)k_density_fast<N, N>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(160 * 1024 - stat)));
        attr = true;
    }
    hipLaunchKernelGGL((k_density_fast<N, N>), dim3(g.n[2] * ngrp), dim3(FT), dyn, s, colbuf,
                       (const int *)p->col_uy.p, occ, (const double2 *)p->tw[0].p,
                       (const double2 *)p->tw[1].p, g, nb, bpg, Q1, Q2, t2, rhor);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

// Electron density on the mesh, the band loop of Src/omp/charge_by_fft.f90:170-197:
// rho(r) = sum_n occ_n |psi_n(r)|^2, psi_n(r) the unnormalised G -> r transform.
// The inverse transforms are those of H|psi> (k_col_bwd, then y and x per plane).
int plan_density(cpwgpu_plan_s *p, int neig, const double2 *psi, long ldp, const double *occ_h, double *rho,
                 hipStream_t s) {
    const FftGeom &g = p->g;
    if (!fast_sizes(g)) return fail(CPWGPU_ERR_SIZE, "the density needs a mesh with a two-pass split");
    int rc;
    const int bpg = 8, ngrp = (neig + bpg - 1) / bpg;
    const long nmesh = (long)g.n[0] * g.n[1] * g.n[2];
    if ((rc = p->colbuf.reserve(16 * (size_t)g.n[2] * g.ncolp * neig)) ||
        (rc = g_fw.rhor.reserve(8 * (size_t)ngrp * nmesh)))
        return rc;
    const double *occ = nullptr;
    if (occ_h) {
        if ((rc = upload(g_fw.occ, occ_h, (size_t)neig))) return rc;
        occ = (const double *)g_fw.occ.p;
        HIPCHK(hipStreamSynchronize(g_stream));  // the upload went through the library stream
    }
    const LocalArgs la{psi, ldp, nullptr, 0, neig, 0};
    if ((rc = local_stage<double>(p, la, (cplx<double> *)p->colbuf.p, (const double *)p->v64.p, ST_COL_BWD, s)))
        return rc;
    rc = CPWGPU_ERR_SIZE;
#define CPW_DEN(N) if (g.n[0] == N) rc = launch_density<N>(p, (const cplx<double> *)p->colbuf.p, occ, neig, bpg, ngrp, 0, 0, nullptr, s, (double *)g_fw.rhor.p);
    CPW_FAST_SIZES(CPW_DEN)
#undef CPW_DEN
    if (rc) return rc;
    const long tot = (long)g.id * g.n[1] * g.n[2];
    hipLaunchKernelGGL(k_rho_r_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       (const double *)g_fw.rhor.p, ngrp, g.n[0], g.n[1], g.n[2], g.id, rho);
    HIPCHK(hipGetLastError());
    return CPWGPU_OK;
}

int plan_forces(cpwgpu_plan_s *p, int neig, const double2 *psi, long ldp, const double *occ_h,
                int natom, const double *rat, const double *avec, const int *proj_atom, int nq,
                const int *qv, int nspec, const int *species, const double *fq, double *force,
                double *enl, hipStream_t s) {
    const FftGeom &g = p->g;
    if (!fast_sizes(g)) return fail(CPWGPU_ERR_SIZE, "forces need a mesh with a two-pass split");
    int rc;
    std::vector<double> fcov(3 * (size_t)natom, 0.0);
    // local part ------------------------------------------------------------
    if (nq > 0) {
        int Q1 = 0, Q2 = 0;
        for (int k = 0; k < nq; k++) {
            Q1 = std::max(Q1, std::abs(qv[3 * k]));
            Q2 = std::max(Q2, std::abs(qv[3 * k + 1]));
            if (2 * std::abs(qv[3 * k]) >= g.n[0] || 2 * std::abs(qv[3 * k + 1]) >= g.n[1] ||
                2 * std::abs(qv[3 * k + 2]) >= g.n[2])
                return fail(CPWGPU_ERR_ARG, "q vector %d outside the mesh", k);
        }
        if (2 * Q1 + 1 > (g.nx | 1) || 2 * Q2 + 1 > g.n[1])
            return fail(CPWGPU_ERR_SIZE, "local potential q box wider than the wave-function box");
        const int bpg = 8, ngrp = (neig + bpg - 1) / bpg;
        const size_t per_band = 16 * (size_t)g.n[2] * g.ncolp;
        if ((rc = p->colbuf.reserve(per_band * neig))) return rc;
        const size_t t2n = (size_t)ngrp * g.n[2] * (2 * Q2 + 1) * (2 * Q1 + 1);
        if ((rc = g_fw.t2.reserve(16 * t2n)) || (rc = g_fw.rho.reserve(16 * (size_t)nq)) ||
            (rc = upload(g_fw.qv, qv, 3 * (size_t)nq)) ||
            (rc = upload(g_fw.fq, fq, (size_t)nq * nspec)) || (rc = upload(g_fw.rat, rat, 3 * (size_t)natom)) ||
            (rc = upload(g_fw.species, species, (size_t)natom)) || (rc = g_fw.fl.reserve(24 * (size_t)natom)))
            return rc;
        const double *occ = nullptr;
        if (occ_h) {
            if ((rc = upload(g_fw.occ, occ_h, (size_t)neig))) return rc;
            occ = (const double *)g_fw.occ.p;
        }
        HIPCHK(hipStreamSynchronize(g_stream));  // uploads went through the library stream
        const LocalArgs la{psi, ldp, nullptr, 0, neig, 0};
        if ((rc = local_stage<double>(p, la, (cplx<double> *)p->colbuf.p, (const double *)p->v64.p,
                                      ST_COL_BWD, s)))
            return rc;
        bool reg = false;
#define CPW_DR(N) if (g.n[0] == N) reg = den_reg_ok<N>();
        CPW_FAST_SIZES(CPW_DR)
#undef CPW_DR
        const int W1 = 2 * Q1 + 1;
        const long nplane = (long)g.n[2] * (2 * Q2 + 1) * W1;
        if ((rc = g_fw.t2s.reserve(16 * (size_t)nplane))) return rc;
        rc = CPWGPU_ERR_SIZE;
        if (reg) {
            // density planes per band group (register lines), then per z-plane
            // the group sum and the partial transforms along x and y
            const long nmesh = (long)g.n[0] * g.n[1] * g.n[2];
            const size_t lds = 16 * ((size_t)g.n[0] + g.n[1] + (size_t)g.n[1] * W1) + 8 * (size_t)g.n[0] * g.n[1];
            if (lds > 160 * 1024) return fail(CPWGPU_ERR_SIZE, "density plane of %d x %d does not fit in LDS", g.n[0], g.n[1]);
            if ((rc = g_fw.rhor.reserve(8 * (size_t)ngrp * nmesh))) return rc;
            rc = CPWGPU_ERR_SIZE;
#define CPW_DEN(N) if (g.n[0] == N) rc = launch_density<N>(p, (const cplx<double> *)p->colbuf.p, occ, neig, bpg, ngrp, 0, 0, nullptr, s, (double *)g_fw.rhor.p);
            CPW_FAST_SIZES(CPW_DEN)
#undef CPW_DEN
            if (rc) return rc;
            static bool attr = false;
            if (!attr) {
                HIPCHK(hipFuncSetAttribute((const void *)k_rho_partial, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024));
                attr = true;
            }
            hipLaunchKernelGGL(k_rho_partial, dim3(g.n[2]), dim3(256), lds, s, (const double *)g_fw.rhor.p, ngrp,
                               g.n[0], g.n[1], g.n[2], Q1, Q2, (const double2 *)p->tw[0].p,
                               (const double2 *)p->tw[1].p, (double2 *)g_fw.t2s.p);
        } else {
            // partial transforms per band group inside the density pass, then the
            // band groups summed (fixed order, all of the GPU)
#define CPW_DEN(N) if (g.n[0] == N) rc = launch_density<N>(p, (const cplx<double> *)p->colbuf.p, occ, neig, bpg, ngrp, Q1, Q2, (double2 *)g_fw.t2.p, s);
            CPW_FAST_SIZES(CPW_DEN)
#undef CPW_DEN
            if (rc) return rc;
            hipLaunchKernelGGL(k_t2_sum, dim3((unsigned)((nplane + 255) / 256)), dim3(256), 0, s,
                               (const double2 *)g_fw.t2.p, nplane, ngrp, (double2 *)g_fw.t2s.p);
        }
        // per q the z sum
        hipLaunchKernelGGL(k_rho_q, dim3((nq + 255) / 256), dim3(256), 0, s, (const double2 *)g_fw.t2s.p,
                           (const int *)g_fw.qv.p, nq, (const double2 *)p->tw[2].p, g.n[2], Q1, Q2, p->inv_n,
                           (double2 *)g_fw.rho.p);
        hipLaunchKernelGGL(k_force_local, dim3(natom), dim3(256), 0, s, (const double2 *)g_fw.rho.p,
                           (const int *)g_fw.qv.p, (const double *)g_fw.fq.p, nq,
                           (const double *)g_fw.rat.p, (const int *)g_fw.species.p, (double *)g_fw.fl.p);
        HIPCHK(hipGetLastError());
        std::vector<double> fl(3 * (size_t)natom);
        HIPCHK(hipMemcpyAsync(fl.data(), g_fw.fl.p, 24 * (size_t)natom, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (size_t k = 0; k < fl.size(); k++) fcov[k] += fl[k];
    }
    // non-local part --------------------------------------------------------
    double e_nl = 0.0;
    const int nanl = p->nanl;
    if (nanl > 0) {
        if (!proj_atom) return fail(CPWGPU_ERR_ARG, "proj_atom missing");
        for (int a = 0; a < nanl; a++)
            if (proj_atom[a] < 0 || proj_atom[a] >= natom) return fail(CPWGPU_ERR_ARG, "proj_atom[%d] = %d", a, proj_atom[a]);
        const long mat = (long)nanl * neig;
        if ((rc = g_fw.p.reserve(16 * 4 * (size_t)mat)) || (rc = upload(g_fw.patom, proj_atom, (size_t)nanl)) ||
            (rc = g_fw.fnl.reserve(24 * (size_t)natom)) || (rc = g_fw.enl.reserve(8 * (size_t)natom)))
            return rc;
        const double *occ = nullptr;
        if (occ_h) {
            if ((rc = upload(g_fw.occ, occ_h, (size_t)neig))) return rc;
            occ = (const double *)g_fw.occ.p;
        }
        HIPCHK(hipStreamSynchronize(g_stream));
        for (int j = 0; j < 4; j++) {
            // <beta|psi>, then <beta|g_j psi> for the three Cartesian components
            const double *w = j == 0 ? nullptr : (const double *)p->gw.p + (size_t)(j - 1) * g.mtxd;
            int nsl;
            if ((rc = gram_partials((const double2 *)p->anl64.p, p->lda, nanl, psi, ldp, neig, g.mtxd, p->part,
                                    &nsl, s, 0, w)))
                return rc;
            launch_proj_reduce<double>((const cplx<double> *)p->part.p, nullptr,
                                       (cplx<double> *)g_fw.p.p + (size_t)j * mat, nanl, neig, nsl, s);
        }
        hipLaunchKernelGGL(k_force_nl, dim3(natom), dim3(256), 0, s, (const double2 *)g_fw.p.p, nanl,
                           neig, (const double *)p->xnl.p, occ, (const int *)g_fw.patom.p,
                           (double *)g_fw.fnl.p, (double *)g_fw.enl.p);
        HIPCHK(hipGetLastError());
        std::vector<double> fn(3 * (size_t)natom), en(natom);
        HIPCHK(hipMemcpyAsync(fn.data(), g_fw.fnl.p, 24 * (size_t)natom, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(en.data(), g_fw.enl.p, 8 * (size_t)natom, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (size_t k = 0; k < fn.size(); k++) fcov[k] += fn[k];
        for (int a = 0; a < natom; a++) e_nl += en[a];
    }
    // Cartesian: F = inv(avec)^T f_cov  (avec columns are the lattice vectors)
    const double *A = avec;  // A[i + 3 j] = component i of a_j
    const double det = A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
                       A[6] * (A[1] * A[5] - A[4] * A[2]);
    if (std::fabs(det) < 1e-300) return fail(CPWGPU_ERR_ARG, "singular lattice vectors");
    double inv[9];  // inv = A^-1, inv[i + 3 j]
    inv[0] = (A[4] * A[8] - A[7] * A[5]) / det;
    inv[3] = -(A[3] * A[8] - A[6] * A[5]) / det;
    inv[6] = (A[3] * A[7] - A[6] * A[4]) / det;
    inv[1] = -(A[1] * A[8] - A[7] * A[2]) / det;
    inv[4] = (A[0] * A[8] - A[6] * A[2]) / det;
    inv[7] = -(A[0] * A[7] - A[6] * A[1]) / det;
    inv[2] = (A[1] * A[5] - A[4] * A[2]) / det;
    inv[5] = -(A[0] * A[5] - A[3] * A[2]) / det;
    inv[8] = (A[0] * A[4] - A[3] * A[1]) / det;
    for (int a = 0; a < natom; a++)
        for (int i = 0; i < 3; i++) {
            // (inv^T f)_i = sum_j inv[j][i] f_j = sum_j inv[j + 3 i] f_j
            double v = 0.0;
            for (int j = 0; j < 3; j++) v += inv[j + 3 * i] * fcov[3 * a + j];
            force[3 * a + i] = v;
        }
    if (enl) *enl = e_nl;
    return CPWGPU_OK;
}

}  // namespace

namespace {

// ---- autotuner ---------------------------------------------------------------
// Key of a tuned shape: the mesh, the x-box width, the number of z columns, the
// projector count and the band count decide which kernels win.
struct TuneKey {
    int prec, n1, n2, n3, nx, ncol, nanl, neig;
    bool operator==(const TuneKey &o) const {
        return prec == o.prec && n1 == o.n1 && n2 == o.n2 && n3 == o.n3 && nx == o.nx &&
               ncol == o.ncol && nanl == o.nanl && neig == o.neig;
    }
};

// one line per key: prec n1 n2 n3 nx ncol nanl neig | plane proj back splitk kin col sched
// (lines without the last fields, kin, col and sched, read them as -1)
bool tune_cache_lookup(const TuneKey &k, int *plane, int *proj, int *back, int *splitk, int *kin, int *col,
                       int *sched) {
    const char *path = std::getenv("CPWGPU_TUNE_CACHE");
    if (!path || !*path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        TuneKey r;
        int pv, jv, bv, sk, kv = -1, cv = -1, sv = -1;
        const int n = std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d %d %d %d", &r.prec, &r.n1, &r.n2,
                                  &r.n3, &r.nx, &r.ncol, &r.nanl, &r.neig, &pv, &jv, &bv, &sk, &kv, &cv, &sv);
        if (n < 12) continue;
        if (r == k) {  // the last entry for a key wins
            *plane = pv;
            *proj = jv;
            *back = bv;
            *splitk = sk;
            *kin = n >= 13 ? kv : -1;
            *col = n >= 14 ? cv : -1;
            *sched = n >= 15 ? sv : -1;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void tune_cache_store(const TuneKey &k, int plane, int proj, int back, int splitk, int kin, int col,
                      int sched) {
    const char *path = std::getenv("CPWGPU_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d %d %d %d\n", k.prec, k.n1, k.n2, k.n3, k.nx, k.ncol,
                 k.nanl, k.neig, plane, proj, back, splitk, kin, col, sched);
    std::fclose(f);
}

// median-free timing: the best of `reps` applies (two-component ones with
// spin) after one warm-up
int time_apply(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh,
               int prec, hipStream_t s, int reps, float *best, bool spin = false) {
    auto apply = [&]() {
        return spin ? cpwgpu_plan_apply_spin(p, neig, psi, ldp, hpsi, ldh, prec, s)
                    : cpwgpu_plan_apply(p, neig, psi, ldp, hpsi, ldh, prec, s);
    };
    // batches of back-to-back applies, as a caller's loop runs them (the
    // next apply's first kernels overlap the previous one's last): the time
    // per apply of the fastest batch; applies of more than 2 ms are timed
    // one by one (their ends are a small part of them)
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, s));
    int st = apply();
    HIPCHK(hipEventRecord(b, s));
    HIPCHK(hipEventSynchronize(b));
    float first = 0.f;
    HIPCHK(hipEventElapsedTime(&first, a, b));
    const int BATCH = first > 2.f ? 1 : 3;
    *best = 3.4e38f;
    for (int r = 0; r < reps && !st; r++) {
        HIPCHK(hipEventRecord(a, s));
        for (int i = 0; i < BATCH && !st; i++) st = apply();
        HIPCHK(hipEventRecord(b, s));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        *best = std::min(*best, ms / BATCH);
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    return st;
}

int tune_plan(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh,
              int prec, hipStream_t s) {
    const int ix = prec == CPWGPU_PREC_FP32 ? 1 : 0;
    const FftGeom &g = p->g;
    const TuneKey key{prec, g.n[0], g.n[1], g.n[2], g.nx, g.ncol, p->nanl, neig};
    int pv = -1, jv = -1, bv = -1, sk = -1, kv = -1, cv = -1, sv = -1;
    if (tune_cache_lookup(key, &pv, &jv, &bv, &sk, &kv, &cv, &sv)) {
        p->sched_var[ix] = sv >= 0 && sv <= 2 ? sv : -1;
        bool ok;
        if (prec == CPWGPU_PREC_FP32) plane_var_resolve<float>(p, pv, &ok);
        else plane_var_resolve<double>(p, pv, &ok);
        p->plane_var[ix] = ok ? pv : -1;
        p->proj_var[ix] = jv >= 0 && jv < NPROJ_VAR ? jv : -1;
        p->back_var[ix] = bv >= 0 && bv < NBACK_VAR ? bv : -1;
        p->splitk[ix] = sk >= 64 && sk <= 4096 ? sk : -1;
        p->kin_var[ix] = kv == 0 || kv == 1 ? kv : -1;
        bool cok;
        if (prec == CPWGPU_PREC_FP32) col_var_resolve<float>(p, cv, &cok);
        else col_var_resolve<double>(p, cv, &cok);
        p->col_var[ix] = cok && cv >= 0 ? cv : -1;
        return cpwgpu_plan_apply(p, neig, psi, ldp, hpsi, ldh, prec, s);
    }
    int st;
    float best = 3.4e38f, t;
    // column kernels (with the default plane kernel), then the plane kernels
    int best_cv = 0;
    for (int cv2 = 0; cv2 <= 1; cv2++) {
        bool ok;
        if (prec == CPWGPU_PREC_FP32) col_var_resolve<float>(p, cv2, &ok);
        else col_var_resolve<double>(p, cv2, &ok);
        if (!ok) continue;
        p->col_var[ix] = cv2;
        if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t))) return st;
        if (t < best) { best = t; best_cv = cv2; }
    }
    p->col_var[ix] = best_cv;
    best = 3.4e38f;
    int best_pv = 0;
    for (int v = 0; v < NPLANE_VAR; v++) {
        bool ok;
        if (prec == CPWGPU_PREC_FP32) plane_var_resolve<float>(p, v, &ok);
        else plane_var_resolve<double>(p, v, &ok);
        if (!ok) continue;
        p->plane_var[ix] = v;
        if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t))) return st;
        if (t < best) { best = t; best_pv = v; }
    }
    p->plane_var[ix] = best_pv;
    int best_sk = -1, best_jv = 0;
    if (p->nanl > 0) {
        // projection kernel x split-K target
        const int cand[4] = {128, 256, 512, 1024};
        best = 3.4e38f;
        for (int jv = 0; jv < NPROJ_VAR; jv++)
            for (int c : cand) {
                p->proj_var[ix] = jv;
                p->splitk[ix] = c;
                if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t))) return st;
                if (t < best) { best = t; best_sk = c; best_jv = jv; }
            }
    }
    p->proj_var[ix] = best_jv;
    p->splitk[ix] = best_sk;
    int best_bv = 0;
    if (p->nanl > 0) {
        best = 3.4e38f;
        for (int bv = 0; bv < NBACK_VAR; bv++) {
            p->back_var[ix] = bv;
            if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t))) return st;
            if (t < best) { best = t; best_bv = bv; }
        }
    }
    p->back_var[ix] = best_bv;
    // where the kinetic term goes: it moves work between the two streams
    int best_kv = 0;
    if (p->nanl > 0) {
        best = 3.4e38f;
        for (int kv2 = 0; kv2 <= 1; kv2++) {
            p->kin_var[ix] = kv2;
            if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t))) return st;
            if (t < best) { best = t; best_kv = kv2; }
        }
    }
    p->kin_var[ix] = best_kv;
    // stream schedule (two-stream applies only); schedules 1 and 2 change
    // what the back-projection runs beside: choose its kernel again
    int best_sv = 0;
    if (p->nanl > 0 && g_overlap) {
        p->sched_var[ix] = 0;
        if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &best))) return st;
        for (int sv = 1; sv <= 2; sv++) {
            p->sched_var[ix] = sv;
            for (int bv = 0; bv < NBACK_VAR; bv++) {
                p->back_var[ix] = bv;
                if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t))) return st;
                if (t < best) { best = t; best_sv = sv; best_bv = bv; }
            }
        }
        p->back_var[ix] = best_bv;
    }
    p->sched_var[ix] = best_sv;
    // the plane kernels once more, beside the chosen non-local kernels (how
    // they share the GPU with the GEMM stream decides as much as their speed)
    best = 3.4e38f;
    for (int v = 0; v < NPLANE_VAR; v++) {
        bool ok;
        if (prec == CPWGPU_PREC_FP32) plane_var_resolve<float>(p, v, &ok);
        else plane_var_resolve<double>(p, v, &ok);
        if (!ok) continue;
        p->plane_var[ix] = v;
        if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 8, &t))) return st;
        if (t < best) { best = t; best_pv = v; }
    }
    p->plane_var[ix] = best_pv;
    // leave hpsi holding H psi of the chosen configuration
    if ((st = cpwgpu_plan_apply(p, neig, psi, ldp, hpsi, ldh, prec, s))) return st;
    tune_cache_store(key, best_pv, best_jv, best_bv, best_sk, best_kv, best_cv, best_sv);
    return CPWGPU_OK;
}

// Tuner of the two-component path: times cpwgpu_plan_apply_spin for the coupled
// plane kernels (nsp = 4), then the projection kernel x split-K target, then
// the back-projection kernel, on the spinor GEMM shapes (2 mtxd rows, nanlsp
// projectors).  The GEMM choices are the plan's (shared with the scalar path).
int tune_spin(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh, int prec,
              hipStream_t s) {
    const int ix = prec == CPWGPU_PREC_FP32 ? 1 : 0;
    int st;
    float best, t;
    if (p->nsp == 4) {
        const bool built = prec == CPWGPU_PREC_FP32 ? spin_reg_built<float>(p) : spin_reg_built<double>(p);
        int best_sv = 0;
        best = 3.4e38f;
        for (int sv = 0; sv <= (built ? 1 : 0); sv++) {
            p->spin_var[ix] = sv;
            if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t, true))) return st;
            if (t < best) { best = t; best_sv = sv; }
        }
        p->spin_var[ix] = best_sv;
    }
    if (p->snanl > 0) {
        const int cand[4] = {128, 256, 512, 1024};
        int best_jv = 0, best_sk = -1;
        best = 3.4e38f;
        for (int jv = 0; jv < NPROJ_VAR; jv++)
            for (int c : cand) {
                p->proj_var[ix] = jv;
                p->splitk[ix] = c;
                if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t, true))) return st;
                if (t < best) { best = t; best_sk = c; best_jv = jv; }
            }
        p->proj_var[ix] = best_jv;
        p->splitk[ix] = best_sk;
        int best_bv = 0;
        best = 3.4e38f;
        for (int bv = 0; bv < NBACK_VAR; bv++) {
            p->back_var[ix] = bv;
            if ((st = time_apply(p, neig, psi, ldp, hpsi, ldh, prec, s, 5, &t, true))) return st;
            if (t < best) { best = t; best_bv = bv; }
        }
        p->back_var[ix] = best_bv;
    }
    // leave hpsi holding H psi of the chosen configuration
    return cpwgpu_plan_apply_spin(p, neig, psi, ldp, hpsi, ldh, prec, s);
}

}  // namespace

// ---- allocation-free apply ----------------------------------------------------
// Workspace of cpwgpu_plan_apply_noalloc: a fixed device array, one copy per
// device, held until the process ends.  Host code never passes the array's
// name to a kernel: it is a host-side stand-in.  The device address of the
// current device's copy comes from hipGetSymbolAddress (scratch_addr).
constexpr size_t NOALLOC_SCRATCH = (size_t)CPWGPU_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH];

namespace {
constexpr int MAXDEV = 64;
void *g_scratch_ptr[MAXDEV] = {};

// device address of the current device's scratch array; resolved once per device
int scratch_addr(void **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAXDEV) return fail(CPWGPU_ERR_ARG, "device %d out of range", dev);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(CPWGPU_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return CPWGPU_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// bytes of scratch an apply of nb bands takes: z-column buffer, split-K
// partials and D<beta|psi>, each 256-byte aligned
template <typename R> size_t noalloc_need(const cpwgpu_plan_s *p, int nb) {
    const size_t cs = sizeof(cplx<R>);
    size_t n = align256(cs * (size_t)p->g.n[2] * p->g.ncolp * nb);
    if (p->nanl > 0 && p->g.mtxd > 0) {
        int nsl, kslice;
        splitk_shape<R>(p, p->g.mtxd, p->nanl, nb, &nsl, &kslice);
        n += align256(cs * (size_t)nsl * nb * p->nanl) + align256(cs * (size_t)nb * p->nanl);
    }
    return n;
}

// H|psi> in band chunks whose workspace fits in the scratch array at scr
template <typename R>
int apply_noalloc_t(cpwgpu_plan_s *p, int neig, const double *psi, long ldp, double *hpsi, long ldh,
                    hipStream_t s, unsigned char *scr) {
    const bool nl = p->nanl > 0 && p->g.mtxd > 0;
    if (sizeof(R) == 4 && nl && !p->anl32_valid)
        return fail(CPWGPU_ERR_ARG, "no FP32 copy of the projectors: it is made by the first FP32 "
                                    "cpwgpu_plan_apply after the projectors are set");
    if (noalloc_need<R>(p, 1) > NOALLOC_SCRATCH)
        return fail(CPWGPU_ERR_SIZE, "one band needs %zu bytes of workspace, more than the %zu of the "
                                     "allocation-free entry point",
                    noalloc_need<R>(p, 1), NOALLOC_SCRATCH);
    int nbc = neig;
    while (nbc > 1 && noalloc_need<R>(p, nbc) > NOALLOC_SCRATCH) nbc--;
    for (int b0 = 0; b0 < neig; b0 += nbc) {
        const int nb = std::min(nbc, neig - b0);
        const size_t cs = sizeof(cplx<R>);
        Work w;
        w.colbuf = scr;
        w.colbuf_n = cs * (size_t)p->g.n[2] * p->g.ncolp * nb;
        size_t off = align256(w.colbuf_n);
        if (nl) {
            int nsl, kslice;
            splitk_shape<R>(p, p->g.mtxd, p->nanl, nb, &nsl, &kslice);
            w.part = scr + off;
            w.part_n = cs * (size_t)nsl * nb * p->nanl;
            off += align256(w.part_n);
            w.dhd = scr + off;
            w.dhd_n = cs * (size_t)nb * p->nanl;
        }
        const int st = apply_t<R>(p, nb, psi + 2 * (long)b0 * ldp, ldp, hpsi + 2 * (long)b0 * ldh, ldh, s, &w);
        if (st) return st;
    }
    return CPWGPU_OK;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int cpwgpu_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(CPWGPU_ERR_ARG, "device %d of %d", device, n);
    if (g_stream && device != g_device) {
        HIPCHK(hipStreamDestroy(g_stream));
        g_stream = nullptr;
    }
    g_device = device;
    HIPCHK(hipSetDevice(device));
    void *scr = nullptr;
    int st = scratch_addr(&scr);  // the allocation-free entry point's workspace on this device
    if (st) return st;
    return ensure_stream();
}

int cpwgpu_finalize(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_drop.slot.clear();
    for (PlanSlot &sl : g_allk.slots) sl.clear();
    g_allk.slots.clear();
    g_allk.psi.release();
    g_allk.hpsi.release();
    g_drop.psi.release();
    g_drop.hpsi.release();
    g_drop.hs.release();
    g_drop.hd.release();
    g_drop.c8.release();
    g_spin_tmp.release();
    g_kb_tab.release(); g_kb_types.release(); g_kb_atoms.release(); g_kb_gw.release(); g_kb_out.release();
    if (g_blas) {
        if (g_stream) HIPCHK(hipStreamSynchronize(g_stream));
        rocblas_destroy_handle(g_blas);
        g_blas = nullptr;
    }
    for (DevBuf *b : {&g_dt.xa, &g_dt.hxa, &g_dt.xb, &g_dt.hxb, &g_dt.pb, &g_dt.cpk, &g_dt.hd, &g_dt.m,
                      &g_dt.g, &g_dt.g0,
                      &g_dt.pm, &g_dt.w, &g_dt.e, &g_dt.info, &g_dt.res, &g_dt.act, &g_dt.d, &g_dt.mones,
                      &g_dt.eall, &g_fw.rhor, &g_fw.t2, &g_fw.t2s})
        b->release();
    if (g_aux) {
        HIPCHK(hipStreamSynchronize(g_aux));
        HIPCHK(hipStreamDestroy(g_aux));
        HIPCHK(hipEventDestroy(g_fork));
        HIPCHK(hipEventDestroy(g_join));
        g_aux = nullptr;
    }
    if (g_stream) {
        HIPCHK(hipStreamSynchronize(g_stream));
        HIPCHK(hipStreamDestroy(g_stream));
        g_stream = nullptr;
    }
    return CPWGPU_OK;
}

const char *cpwgpu_last_error(void) { return g_err.c_str(); }

int cpwgpu_set_ditsp_method(int method) {
    if (method != CPWGPU_DITSP_BLOCK && method != CPWGPU_DITSP_CHEB)
        return fail(CPWGPU_ERR_MODE, "diagonalisation method %d", method);
    g_ditsp_method = method;
    return CPWGPU_OK;
}

int cpwgpu_set_overlap(int on) {
    g_overlap = on ? 1 : 0;
    return CPWGPU_OK;
}

int cpwgpu_set_precision(int prec) {
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32)
        return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    g_prec = prec;
    return CPWGPU_OK;
}

int cpwgpu_plan_create(cpwgpu_plan_t *plan, int mtxd, int ng, const int *kgv, const double *ekpg,
                       const int *isort, const int *kmscr, int mxddim, int mxdgve) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || !kmscr || (mtxd > 0 && (!kgv || !ekpg || !isort)))
        return fail(CPWGPU_ERR_ARG, "null argument");
    *plan = nullptr;
    cpwgpu_plan_s *p = new cpwgpu_plan_s();
    int st = build_plan(p, mtxd, ng, kgv, ekpg, isort, kmscr, mxddim, mxdgve);
    if (st) {
        delete p;
        return st;
    }
    *plan = p;
    return CPWGPU_OK;
}

int cpwgpu_plan_destroy(cpwgpu_plan_t plan) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_stream) hipStreamSynchronize(g_stream);
    hipDeviceSynchronize();
    delete plan;
    return CPWGPU_OK;
}

int cpwgpu_plan_set_potential(cpwgpu_plan_t plan, const double *vscr, int mxdscr) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || !vscr) return fail(CPWGPU_ERR_ARG, "null argument");
    int st = set_potential(plan, vscr, mxdscr);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_set_projectors(cpwgpu_plan_t plan, const double *anlga, const double *xnlkb,
                               int nanl, int mxddim) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (mxddim < plan->mtxd) return fail(CPWGPU_ERR_MTXD, "mxddim %d < mtxd %d", mxddim, plan->mtxd);
    int st = set_projectors_host(plan, anlga, xnlkb, nanl, mxddim);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_set_projectors_device(cpwgpu_plan_t plan, const double *anlga_dev, long lda,
                                      const double *xnlkb_dev, int nanl) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || nanl < 0) return fail(CPWGPU_ERR_ARG, "bad argument");
    if (nanl > 0 && (!anlga_dev || !xnlkb_dev || lda < plan->mtxd)) return fail(CPWGPU_ERR_ARG, "bad projector arrays");
    int st = replace_projectors(plan, nanl, [&](double2 *anl, long ld) -> int {
        if (plan->mtxd > 0)
            HIPCHK(hipMemcpy2DAsync(anl, 16 * (size_t)ld, anlga_dev, 16 * (size_t)lda, 16 * (size_t)plan->mtxd, nanl,
                                    hipMemcpyDeviceToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(plan->xnl.p, xnlkb_dev, 8 * (size_t)nanl, hipMemcpyDeviceToDevice, g_stream));
        return CPWGPU_OK;
    });
    if (st) return st;
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_build_projectors_kb(cpwgpu_plan_t plan, const double *rkpt, int nanl, const int *nqnl,
                                    const double *delqnl, const double *vkb, const int *nkb, int ntype,
                                    const int *natom, const double *rat, const double *adot, int mxdtyp,
                                    int mxdatm, int mxdlqp) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || nanl < 0) return fail(CPWGPU_ERR_ARG, "bad argument");
    KbSetup s;
    int st = kb_setup(rkpt, nanl, nqnl, delqnl, vkb, nkb, ntype, natom, rat, adot, mxdtyp, mxdatm, mxdlqp, s);
    if (st) return st;
    st = replace_projectors(plan, nanl, [&](double2 *anl, long lda) -> int {
        int rc = upload(plan->xnl, s.xnl.data(), nanl);
        return rc ? rc : kb_build(s, (const double *)plan->gw.p, plan->mtxd, anl, lda);
    });
    if (st) return st;
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_proj_nl_kb_c16(const double *rkpt, const int *mtxd, const int *isort, const int *nanl, const int *ng,
                          const int *kgv, const int *nqnl, const double *delqnl, const double *vkb, const int *nkb,
                          const int *ntype, const int *natom, const double *rat, const double *adot, double *anlga,
                          double *xnlkb, const int *mxdtyp, const int *mxdatm, const int *mxdlqp, const int *mxddim,
                          const int *mxdanl, const int *mxdgve) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!mtxd || !nanl || !ng || !ntype || !mxdtyp || !mxdatm || !mxdlqp || !mxddim || !mxdanl || !mxdgve)
        return fail(CPWGPU_ERR_ARG, "null dimension argument");
    if (*ng > *mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", *ng, *mxdgve);
    if (*mtxd < 0 || *mtxd > *mxddim) return fail(CPWGPU_ERR_MTXD, "mtxd = %d, mxddim = %d", *mtxd, *mxddim);
    KbSetup s;
    int st = kb_setup(rkpt, *nanl, nqnl, delqnl, vkb, nkb, *ntype, natom, rat, adot, *mxdtyp, *mxdatm, *mxdlqp, s);
    if (st) return st;
    if (*nanl > *mxdanl) return fail(CPWGPU_ERR_ARG, "proj_nl_kb: increase mxdanl from %d to %d", *mxdanl, *nanl);
    const int m = *mtxd, na = *nanl;
    if (na == 0) return CPWGPU_OK;
    if (!xnlkb || !anlga) return fail(CPWGPU_ERR_ARG, "null anlga/xnlkb");
    std::memcpy(xnlkb, s.xnl.data(), sizeof(double) * na);
    if (m == 0) return CPWGPU_OK;
    if (!isort || !kgv) return fail(CPWGPU_ERR_ARG, "null isort/kgv");
    std::vector<double> gwh(3 * (size_t)m);
    for (int i = 0; i < m; i++) {
        if (isort[i] < 1 || isort[i] > *ng) return fail(CPWGPU_ERR_ARG, "isort(%d) = %d outside 1..ng", i + 1, isort[i]);
        for (int j = 0; j < 3; j++) gwh[(size_t)j * m + i] = (double)kgv[3L * (isort[i] - 1) + j];
    }
    if ((st = ensure_stream()) || (st = upload(g_kb_gw, gwh.data(), gwh.size())) ||
        (st = g_kb_out.reserve(16 * (size_t)m * na)) ||
        (st = kb_build(s, (const double *)g_kb_gw.p, m, (double2 *)g_kb_out.p, m)))
        return st;
    HIPCHK(hipMemcpy2DAsync(anlga, 16 * (size_t)*mxddim, g_kb_out.p, 16 * (size_t)m, 16 * (size_t)m, na,
                            hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_apply_noalloc(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp,
                              double *hpsi_dev, long ldh, int prec, void *stream) {
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (neig < 0) return fail(CPWGPU_ERR_ARG, "neig = %d", neig);
    if (neig == 0) return CPWGPU_OK;
    if (!psi_dev || !hpsi_dev) return fail(CPWGPU_ERR_ARG, "null psi/hpsi");
    if (ldp < plan->mtxd || ldh < plan->mtxd) return fail(CPWGPU_ERR_ARG, "leading dimension < mtxd");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    void *scr = nullptr;
    int st = scratch_addr(&scr);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    if (prec == CPWGPU_PREC_FP64)
        return apply_noalloc_t<double>(plan, neig, psi_dev, ldp, hpsi_dev, ldh, s, (unsigned char *)scr);
    return apply_noalloc_t<float>(plan, neig, psi_dev, ldp, hpsi_dev, ldh, s, (unsigned char *)scr);
}

int cpwgpu_plan_apply(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp,
                      double *hpsi_dev, long ldh, int prec, void *stream) {
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (neig < 0) return fail(CPWGPU_ERR_ARG, "neig = %d", neig);
    if (neig == 0) return CPWGPU_OK;
    if (!psi_dev || !hpsi_dev) return fail(CPWGPU_ERR_ARG, "null psi/hpsi");
    if (ldp < plan->mtxd || ldh < plan->mtxd) return fail(CPWGPU_ERR_ARG, "leading dimension < mtxd");
    int st = ensure_stream();
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;  // NULL: the null stream, as in HIP
    if (prec == CPWGPU_PREC_FP64) return apply_t<double>(plan, neig, psi_dev, ldp, hpsi_dev, ldh, s);
    if (prec == CPWGPU_PREC_FP32) return apply_t<float>(plan, neig, psi_dev, ldp, hpsi_dev, ldh, s);
    return fail(CPWGPU_ERR_MODE, "precision %d", prec);
}

int cpwgpu_rayleigh_ritz(int mtxd, int nb, double *psi_dev, long ldp, double *hpsi_dev, long ldh,
                         int rotate, double *eig, double *res, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (mtxd < 1 || nb < 1 || !psi_dev || !hpsi_dev || !eig || ldp < mtxd || ldh < mtxd)
        return fail(CPWGPU_ERR_ARG, "bad Rayleigh-Ritz arguments");
    if (nb > mtxd) return fail(CPWGPU_ERR_ARG, "nb = %d > mtxd = %d", nb, mtxd);
    if (nb > 2 * JAC_MAXPAIR) return fail(CPWGPU_ERR_SIZE, "nb = %d > %d", nb, 2 * JAC_MAXPAIR);
    int st = ensure_stream();
    if (st) return st;
    return rayleigh_ritz(mtxd, nb, (double2 *)psi_dev, ldp, (double2 *)hpsi_dev, ldh, rotate, eig, res,
                         (hipStream_t)stream);
}

int cpwgpu_plan_density(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, const double *occ,
                        double *rho_dev, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || neig < 1 || !psi_dev || !rho_dev || ldp < plan->mtxd)
        return fail(CPWGPU_ERR_ARG, "bad density arguments");
    int st = ensure_stream();
    if (st) return st;
    return plan_density(plan, neig, (const double2 *)psi_dev, ldp, occ, rho_dev, (hipStream_t)stream);
}

int cpwgpu_plan_forces(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp,
                       const double *occ, int natom, const double *rat, const double *avec,
                       const int *proj_atom, int nq, const int *qv, int nspec, const int *species,
                       const double *fq, double *force, double *enl) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || neig < 1 || !psi_dev || ldp < plan->mtxd || natom < 1 || !rat || !avec || !force ||
        nq < 0 || (nq > 0 && (!qv || !species || !fq || nspec < 1)))
        return fail(CPWGPU_ERR_ARG, "bad force arguments");
    for (int a = 0; nq > 0 && a < natom; a++)
        if (species[a] < 0 || species[a] >= nspec) return fail(CPWGPU_ERR_ARG, "species[%d] = %d", a, species[a]);
    int st = ensure_stream();
    if (st) return st;
    return plan_forces(plan, neig, (const double2 *)psi_dev, ldp, occ, natom, rat, avec, proj_atom, nq, qv,
                       nspec, species, fq, force, enl, g_stream);
}

}  // extern "C"

namespace {
// What a captured apply has baked into the graph for one plan: the device
// buffers it reads and writes and the projector shape.  Rebuilding the
// projectors with the same count keeps every buffer (the FP32 copy is
// converted again in place), so the graph stays valid; another count, or a
// workspace reallocated by a larger apply, changes this record.
struct PlanState {
    const cpwgpu_plan_s *p = nullptr;
    int nanl = 0;
    long lda = 0, lda32 = 0;
    bool anl32_valid = false;
    const void *buf[9] = {};
    bool operator==(const PlanState &o) const {
        if (p != o.p || nanl != o.nanl || lda != o.lda || lda32 != o.lda32 || anl32_valid != o.anl32_valid)
            return false;
        for (int i = 0; i < 9; i++)
            if (buf[i] != o.buf[i]) return false;
        return true;
    }
};

PlanState plan_state(const cpwgpu_plan_s *p) {
    PlanState s;
    s.p = p;
    s.nanl = p->nanl;
    s.lda = p->lda;
    s.lda32 = p->lda32;
    s.anl32_valid = p->anl32_valid;
    const DevBuf *b[9] = {&p->anl64, &p->anl32, &p->xnl, &p->colbuf, &p->part, &p->dhd, &p->v64, &p->v32, &p->ekpg};
    for (int i = 0; i < 9; i++) s.buf[i] = b[i]->p;
    return s;
}
}  // namespace

struct cpwgpu_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int prec = CPWGPU_PREC_FP64;
    std::vector<PlanState> state;  // per k-point, at capture
};

extern "C" {

int cpwgpu_graph_create(cpwgpu_graph_t *graph, int nk, const cpwgpu_plan_t *plans, const int *neig,
                        const double *const *psi, const long *ldp, double *const *hpsi,
                        const long *ldh, int prec) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!graph || nk < 1 || !plans || !neig || !psi || !ldp || !hpsi || !ldh)
        return fail(CPWGPU_ERR_ARG, "null graph argument");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    *graph = nullptr;
    int st = ensure_stream();
    if (st) return st;
    for (int k = 0; k < nk; k++)
        if (!plans[k] || neig[k] < 0 || ldp[k] < plans[k]->mtxd || ldh[k] < plans[k]->mtxd)
            return fail(CPWGPU_ERR_ARG, "bad arguments for k-point %d", k);
    // the k-points in sequence, each apply overlapping its own non-local GEMMs
    // with its FFT chain (a batched order that chains all GEMMs on the second
    // stream measured no faster: the pairs already fill the GPU)
    auto apply = [&](hipStream_t s) -> int {
        for (int k = 0; k < nk; k++) {
            if (neig[k] == 0) continue;
            const int st2 = prec == CPWGPU_PREC_FP64
                                ? apply_t<double>(plans[k], neig[k], psi[k], ldp[k], hpsi[k], ldh[k], s)
                                : apply_t<float>(plans[k], neig[k], psi[k], ldp[k], hpsi[k], ldh[k], s);
            if (st2) return st2;
        }
        return CPWGPU_OK;
    };
    // eager pass: every buffer, attribute and FP32 projector copy exists before capture
    const bool prof = g_prof.on;
    g_prof.on = false;
    if ((st = apply(g_stream))) { g_prof.on = prof; return st; }
    HIPCHK(hipStreamSynchronize(g_stream));
    hipStream_t cs = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    cpwgpu_graph_s *gr = new cpwgpu_graph_s();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) st = apply(cs);
    hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(cs, &gr->graph) : e;
    g_prof.on = prof;
    if (e == hipSuccess && e2 == hipSuccess && st == CPWGPU_OK)
        e2 = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (st != CPWGPU_OK || e != hipSuccess || e2 != hipSuccess) {
        if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
        if (gr->graph) (void)hipGraphDestroy(gr->graph);
        delete gr;
        if (st) return st;
        return fail(CPWGPU_ERR_HIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    gr->prec = prec;
    for (int k = 0; k < nk; k++) gr->state.push_back(plan_state(plans[k]));
    *graph = gr;
    return CPWGPU_OK;
}

int cpwgpu_graph_launch(cpwgpu_graph_t graph, void *stream) {
    if (!graph || !graph->exec) return fail(CPWGPU_ERR_ARG, "null graph");
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (size_t k = 0; k < graph->state.size(); k++) {
            const PlanState &s = graph->state[k];
            const PlanState now = plan_state(s.p);
            // an FP64 graph does not read the FP32 projector copy
            const bool same = graph->prec == CPWGPU_PREC_FP32
                                  ? now == s
                                  : (now.nanl == s.nanl && now.lda == s.lda && now.buf[0] == s.buf[0] &&
                                     std::equal(now.buf + 2, now.buf + 9, s.buf + 2));
            if (!same)
                return fail(CPWGPU_ERR_ARG,
                            "k-point %zu: the plan's projectors or buffers changed since the graph was "
                            "captured (projector count, or a larger apply); create the graph again", k);
        }
    }
    HIPCHK(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return CPWGPU_OK;
}

int cpwgpu_graph_destroy(cpwgpu_graph_t graph) {
    if (!graph) return CPWGPU_OK;
    hipDeviceSynchronize();
    if (graph->exec) HIPCHK(hipGraphExecDestroy(graph->exec));
    if (graph->graph) HIPCHK(hipGraphDestroy(graph->graph));
    delete graph;
    return CPWGPU_OK;
}

int cpwgpu_profile(int on, unsigned mask) {
    g_prof.on = on != 0;
    g_prof.mask = mask;
    return CPWGPU_OK;
}

int cpwgpu_profile_read(double *ms, long *launches, int n) {
    if (n < 0 || (n > 0 && (!ms || !launches))) return fail(CPWGPU_ERR_ARG, "bad profile buffers");
    for (int k = 0; k < n; k++) { ms[k] = 0.0; launches[k] = 0; }
    for (auto &r : g_prof.recs) {
        HIPCHK(hipEventSynchronize(r.b));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        if (r.k < n) { ms[r.k] += t; launches[r.k] += 1; }
        g_prof.pool.push_back(r.a);
        g_prof.pool.push_back(r.b);
    }
    g_prof.recs.clear();
    return CPWGPU_OK;
}

int cpwgpu_plan_set_variant(cpwgpu_plan_t plan, int prec, int what, int value) {
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    const int ix = prec == CPWGPU_PREC_FP32 ? 1 : 0;
    if (what == CPWGPU_TUNE_PLANE) {
        bool ok = true;
        if (value >= 0) {
            if (prec == CPWGPU_PREC_FP32) plane_var_resolve<float>(plan, value, &ok);
            else plane_var_resolve<double>(plan, value, &ok);
        }
        if (!ok) return fail(CPWGPU_ERR_MODE, "plane variant %d is not built for this mesh", value);
        plan->plane_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_PROJ) {
        if (value >= NPROJ_VAR) return fail(CPWGPU_ERR_MODE, "projection kernel %d", value);
        plan->proj_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_BACK) {
        if (value >= NBACK_VAR) return fail(CPWGPU_ERR_MODE, "back-projection kernel %d", value);
        plan->back_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_KIN) {
        if (value > 1) return fail(CPWGPU_ERR_MODE, "kinetic-term placement %d", value);
        plan->kin_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_SPIN) {
        const bool built = prec == CPWGPU_PREC_FP32 ? spin_reg_built<float>(plan) : spin_reg_built<double>(plan);
        if (value > 1 || (value == 1 && !built))
            return fail(CPWGPU_ERR_MODE, "two-component plane kernel %d is not built for this mesh", value);
        plan->spin_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_SCHED) {
        if (value > 2) return fail(CPWGPU_ERR_MODE, "stream schedule %d", value);
        plan->sched_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_COL) {
        bool ok = true;
        if (prec == CPWGPU_PREC_FP32) col_var_resolve<float>(plan, value, &ok);
        else col_var_resolve<double>(plan, value, &ok);
        if (value > 1 || !ok) return fail(CPWGPU_ERR_MODE, "column kernel %d is not built for this mesh", value);
        plan->col_var[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_SPLITK) {
        if (value >= 0 && (value < 64 || value > 4096))
            return fail(CPWGPU_ERR_ARG, "split-K target %d outside 64..4096", value);
        plan->splitk[ix] = value < 0 ? -1 : value;
        return CPWGPU_OK;
    }
    return fail(CPWGPU_ERR_ARG, "unknown tunable %d", what);
}

int cpwgpu_plan_get_variant(cpwgpu_plan_t plan, int prec, int what, int *value) {
    if (!plan || !value) return fail(CPWGPU_ERR_ARG, "null argument");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    const int ix = prec == CPWGPU_PREC_FP32 ? 1 : 0;
    bool ok;
    if (what == CPWGPU_TUNE_PLANE) {
        *value = prec == CPWGPU_PREC_FP32 ? plane_var_resolve<float>(plan, plan->plane_var[ix], &ok)
                                          : plane_var_resolve<double>(plan, plan->plane_var[ix], &ok);
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_SPLITK) {
        *value = plan->splitk[ix] > 0 ? plan->splitk[ix] : 256;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_PROJ) {
        *value = plan->proj_var[ix] >= 1 && plan->proj_var[ix] < NPROJ_VAR ? plan->proj_var[ix] : 0;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_BACK) {
        *value = plan->back_var[ix] >= 1 && plan->back_var[ix] < NBACK_VAR ? plan->back_var[ix] : 0;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_KIN) {
        *value = plan->kin_var[ix] == 1 ? 1 : 0;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_SPIN) {
        *value = prec == CPWGPU_PREC_FP32 ? spin_var_resolve<float>(plan) : spin_var_resolve<double>(plan);
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_SCHED) {
        *value = plan->sched_var[ix] == 1 || plan->sched_var[ix] == 2 ? plan->sched_var[ix] : 0;
        return CPWGPU_OK;
    }
    if (what == CPWGPU_TUNE_COL) {
        *value = prec == CPWGPU_PREC_FP32 ? col_var_resolve<float>(plan, plan->col_var[ix], nullptr)
                                          : col_var_resolve<double>(plan, plan->col_var[ix], nullptr);
        return CPWGPU_OK;
    }
    return fail(CPWGPU_ERR_ARG, "unknown tunable %d", what);
}

int cpwgpu_plan_tune(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, double *hpsi_dev,
                     long ldh, int prec, void *stream) {
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    if (neig <= 0) return fail(CPWGPU_ERR_ARG, "neig = %d", neig);
    return tune_plan(plan, neig, psi_dev, ldp, hpsi_dev, ldh, prec, (hipStream_t)stream);
}

int cpwgpu_plan_tune_spin(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, double *hpsi_dev,
                          long ldh, int prec, void *stream) {
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    if (neig <= 0) return fail(CPWGPU_ERR_ARG, "neig = %d", neig);
    return tune_spin(plan, neig, psi_dev, ldp, hpsi_dev, ldh, prec, (hipStream_t)stream);
}

int cpwgpu_plan_info(cpwgpu_plan_t plan, int *out) {
    if (!plan || !out) return fail(CPWGPU_ERR_ARG, "null argument");
    const FftGeom &g = plan->g;
    int v[10] = {plan->mtxd, g.n[0], g.n[1], g.n[2], g.id, g.kd[0], g.kd[1], g.kd[2], g.ncol, plan->nanl};
    std::copy(v, v + 10, out);
    return CPWGPU_OK;
}

int cpwgpu_hk_psi_c16(const int *mtxd, const int *neig, const double *psi, double *hpsi,
                      const int *lnewanl, const int *ng, const int *kgv, const double *ekpg,
                      const int *isort, const double *vscr, const int *kmscr, const double *anlga,
                      const double *xnlkb, const int *nanl, const int *mxddim, const int *mxdbnd,
                      const int *mxdanl, const int *mxdgve, const int *mxdscr) {
    (void)lnewanl;
    (void)mxdbnd;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!mtxd || !neig || !ng || !kmscr || !nanl || !mxddim || !mxdanl || !mxdgve || !mxdscr)
        return fail(CPWGPU_ERR_ARG, "null scalar argument");
    if (*ng > *mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", *ng, *mxdgve);
    int ns[3], mfft, mwrk;
    size_fft(kmscr, ns, &mfft, &mwrk);
    if (mwrk > *mxdscr)
        return fail(CPWGPU_ERR_MXDSCR, "mxdwrk = %d is greater than mxdscr = %d", mwrk, *mxdscr);
    if (*nanl > *mxdanl) return fail(CPWGPU_ERR_ARG, "nanl = %d > mxdanl = %d", *nanl, *mxdanl);
    if (*neig < 0) return fail(CPWGPU_ERR_ARG, "neig = %d", *neig);
    cpwgpu_plan_s *p = nullptr;
    int st = drop_in_plan(*mtxd, *ng, kgv, ekpg, isort, kmscr, *mxddim, *mxdgve, &p);
    if (st) return st;
    if (*neig == 0 || *mtxd == 0) return CPWGPU_OK;
    if (!psi || !hpsi || !vscr) return fail(CPWGPU_ERR_ARG, "null array argument");
    if ((st = set_potential(p, vscr, *mxdscr))) return st;
    // projectors are re-sent on every call: the reference callers never clear lnewanl
    if ((st = set_projectors_host(p, anlga, xnlkb, *nanl, *mxddim))) return st;
    const size_t colb = 16 * (size_t)(*mtxd);
    if ((st = g_drop.psi.reserve(colb * *neig))) return st;
    if ((st = g_drop.hpsi.reserve(colb * *neig))) return st;
    HIPCHK(hipMemcpy2DAsync(g_drop.psi.p, colb, psi, 16 * (size_t)*mxddim, colb, *neig,
                            hipMemcpyHostToDevice, g_stream));
    const double *dpsi = (const double *)g_drop.psi.p;
    double *dhpsi = (double *)g_drop.hpsi.p;
    st = g_prec == CPWGPU_PREC_FP32 ? apply_t<float>(p, *neig, dpsi, *mtxd, dhpsi, *mtxd, g_stream)
                                    : apply_t<double>(p, *neig, dpsi, *mtxd, dhpsi, *mtxd, g_stream);
    if (st) return st;
    HIPCHK(hipMemcpy2DAsync(hpsi, 16 * (size_t)*mxddim, g_drop.hpsi.p, colb, colb, *neig,
                            hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_hk_psi_allk_c16(const int *nrk, const int *mtxd_allk, const int *nband, const double *psi_allk,
                           double *hpsi_allk, const int *ng, const int *kgv, const double *ekpg_allk,
                           const int *isort_allk, const double *vscr, const int *kmscr, const double *anlga_allk,
                           const double *xnlkb_allk, const int *nanl_allk, const int *mxddim, const int *mxdbnd,
                           const int *mxdanl, const int *mxdgve, const int *mxdscr, const int *mxdnrk) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!nrk || !mtxd_allk || !nband || !ng || !kmscr || !nanl_allk || !mxddim || !mxdbnd || !mxdanl ||
        !mxdgve || !mxdscr || !mxdnrk)
        return fail(CPWGPU_ERR_ARG, "null scalar argument");
    if (*nrk < 0 || *nrk > *mxdnrk) return fail(CPWGPU_ERR_ARG, "nrk = %d, mxdnrk = %d", *nrk, *mxdnrk);
    if (*ng > *mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", *ng, *mxdgve);
    int ns[3], mfft, mwrk;
    size_fft(kmscr, ns, &mfft, &mwrk);
    if (mwrk > *mxdscr)
        return fail(CPWGPU_ERR_MXDSCR, "mxdwrk = %d is greater than mxdscr = %d", mwrk, *mxdscr);
    size_t maxcol = 0;
    for (int k = 0; k < *nrk; k++) {
        if (mtxd_allk[k] < 0 || mtxd_allk[k] > *mxddim)
            return fail(CPWGPU_ERR_MTXD, "k-point %d: mtxd = %d, mxddim = %d", k + 1, mtxd_allk[k], *mxddim);
        if (nband[k] < 0 || nband[k] > *mxdbnd)
            return fail(CPWGPU_ERR_ARG, "k-point %d: nband = %d, mxdbnd = %d", k + 1, nband[k], *mxdbnd);
        if (nanl_allk[k] < 0 || nanl_allk[k] > *mxdanl)
            return fail(CPWGPU_ERR_ARG, "k-point %d: nanl = %d > mxdanl = %d", k + 1, nanl_allk[k], *mxdanl);
        maxcol = std::max(maxcol, 16 * (size_t)mtxd_allk[k] * nband[k]);
    }
    if (*nrk == 0 || maxcol == 0) return CPWGPU_OK;
    if (!psi_allk || !hpsi_allk || !vscr || !kgv || !ekpg_allk || !isort_allk)
        return fail(CPWGPU_ERR_ARG, "null array argument");
    int st;
    if (g_allk.slots.size() < (size_t)*nrk) g_allk.slots.resize(*nrk);
    const size_t md = (size_t)*mxddim, mb = (size_t)*mxdbnd, ma = (size_t)*mxdanl;
    for (int k = 0; k < *nrk; k++) {
        const int mtxd = mtxd_allk[k], nb = nband[k], nanl = nanl_allk[k];
        cpwgpu_plan_s *p = nullptr;
        if ((st = slot_plan(g_allk.slots[k], mtxd, *ng, kgv, ekpg_allk + k * md, isort_allk + k * md, kmscr,
                            *mxddim, *mxdgve, &p)))
            return st;
        if (nb == 0 || mtxd == 0) continue;
        if ((st = set_potential(p, vscr, *mxdscr))) return st;
        if ((st = g_allk.psi.reserve(maxcol)) || (st = g_allk.hpsi.reserve(maxcol))) return st;
        if ((st = set_projectors_host(p, anlga_allk ? anlga_allk + 2 * k * md * ma : nullptr,
                                      xnlkb_allk ? xnlkb_allk + k * ma : nullptr, nanl, *mxddim)))
            return st;
        const size_t colb = 16 * (size_t)mtxd, off = 2 * k * md * mb;  // doubles
        HIPCHK(hipMemcpy2DAsync(g_allk.psi.p, colb, psi_allk + off, 16 * md, colb, nb, hipMemcpyHostToDevice,
                                g_stream));
        const double *dpsi = (const double *)g_allk.psi.p;
        double *dhpsi = (double *)g_allk.hpsi.p;
        st = g_prec == CPWGPU_PREC_FP32 ? apply_t<float>(p, nb, dpsi, mtxd, dhpsi, mtxd, g_stream)
                                        : apply_t<double>(p, nb, dpsi, mtxd, dhpsi, mtxd, g_stream);
        if (st) return st;
        HIPCHK(hipMemcpy2DAsync(hpsi_allk + off, 16 * md, g_allk.hpsi.p, colb, colb, nb, hipMemcpyDeviceToHost,
                                g_stream));
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_set_potential_spin(cpwgpu_plan_t plan, const double *vscr_sp, int mxdscr, int nsp) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || !vscr_sp) return fail(CPWGPU_ERR_ARG, "null argument");
    int st = set_potential_spin(plan, vscr_sp, mxdscr, nsp);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_set_projectors_spin(cpwgpu_plan_t plan, const double *anlsp, const double *xnlkbsp, int nanlsp,
                                    int mxddim) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (mxddim < plan->mtxd) return fail(CPWGPU_ERR_MTXD, "mtxd = %d, mxddim = %d", plan->mtxd, mxddim);
    int st = set_projectors_spin_host(plan, anlsp, xnlkbsp, nanlsp, mxddim);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_apply_spin(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, double *hpsi_dev,
                           long ldh, int prec, void *stream) {
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (neig < 0) return fail(CPWGPU_ERR_ARG, "neig = %d", neig);
    if (neig == 0) return CPWGPU_OK;
    if (!psi_dev || !hpsi_dev) return fail(CPWGPU_ERR_ARG, "null psi/hpsi");
    if (ldp < 2L * plan->mtxd || ldh < 2L * plan->mtxd) return fail(CPWGPU_ERR_ARG, "leading dimension < 2*mtxd");
    int st = ensure_stream();
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    if (prec == CPWGPU_PREC_FP64)
        return apply_spin<double>(plan, neig, (const double2 *)psi_dev, ldp, (double2 *)hpsi_dev, ldh, s);
    if (prec == CPWGPU_PREC_FP32)
        return apply_spin<float>(plan, neig, (const double2 *)psi_dev, ldp, (double2 *)hpsi_dev, ldh, s);
    return fail(CPWGPU_ERR_MODE, "precision %d", prec);
}

int cpwgpu_hk_psi_spin_c16(const int *mtxd, const int *neig, const double *psi_sp, double *hpsi_sp,
                           const int *lnewanl, const int *ng, const int *kgv, const double *ekpg, const int *isort,
                           const double *vscr_sp, const int *kmscr, const int *nsp, const double *anlsp,
                           const double *xnlkbsp, const int *nanlsp, const int *mxddim, const int *mxdbnd,
                           const int *mxdasp, const int *mxdgve, const int *mxdscr, const int *mxdnsp) {
    (void)lnewanl;
    (void)mxdbnd;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!mtxd || !neig || !ng || !kmscr || !nsp || !nanlsp || !mxddim || !mxdasp || !mxdgve || !mxdscr || !mxdnsp)
        return fail(CPWGPU_ERR_ARG, "null scalar argument");
    // the checks of Src/hk_psi/reference/hk_psi_spin_c16.f90:106-141, in the same order
    if (*ng > *mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", *ng, *mxdgve);
    if (*nsp > *mxdnsp) return fail(CPWGPU_ERR_ARG, "nsp = %d is greater than mxdnsp = %d", *nsp, *mxdnsp);
    if (*nsp != 1 && *nsp != 2 && *nsp != 4)
        return fail(CPWGPU_ERR_MODE, "nsp = %d is an incorrect value, should be 1,2,4", *nsp);
    int ns[3], mfft, mwrk;
    size_fft(kmscr, ns, &mfft, &mwrk);
    if (mwrk > *mxdscr) return fail(CPWGPU_ERR_MXDSCR, "mxdwrk = %d is greater than mxdscr = %d", mwrk, *mxdscr);
    if (*nanlsp > *mxdasp) return fail(CPWGPU_ERR_ARG, "nanlsp = %d > mxdasp = %d", *nanlsp, *mxdasp);
    if (*neig < 0) return fail(CPWGPU_ERR_ARG, "neig = %d", *neig);
    cpwgpu_plan_s *p = nullptr;
    int st = drop_in_plan(*mtxd, *ng, kgv, ekpg, isort, kmscr, *mxddim, *mxdgve, &p);
    if (st) return st;
    if (*neig == 0 || *mtxd == 0) return CPWGPU_OK;
    if (!psi_sp || !hpsi_sp || !vscr_sp) return fail(CPWGPU_ERR_ARG, "null array argument");
    if ((st = set_potential_spin(p, vscr_sp, *mxdscr, *nsp))) return st;
    if ((st = set_projectors_spin_host(p, anlsp, xnlkbsp, *nanlsp, *mxddim))) return st;
    const size_t colb = 32 * (size_t)(*mtxd);  // two components
    if ((st = g_drop.psi.reserve(colb * *neig))) return st;
    if ((st = g_drop.hpsi.reserve(colb * *neig))) return st;
    HIPCHK(hipMemcpy2DAsync(g_drop.psi.p, colb, psi_sp, 32 * (size_t)*mxddim, colb, *neig, hipMemcpyHostToDevice,
                            g_stream));
    const double2 *dpsi = (const double2 *)g_drop.psi.p;
    double2 *dhpsi = (double2 *)g_drop.hpsi.p;
    const long ld = 2L * *mtxd;
    st = g_prec == CPWGPU_PREC_FP32 ? apply_spin<float>(p, *neig, dpsi, ld, dhpsi, ld, g_stream)
                                    : apply_spin<double>(p, *neig, dpsi, ld, dhpsi, ld, g_stream);
    if (st) return st;
    HIPCHK(hipMemcpy2DAsync(hpsi_sp, 32 * (size_t)*mxddim, g_drop.hpsi.p, colb, colb, *neig, hipMemcpyDeviceToHost,
                            g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_hk_psi_c8(const int *mtxd, const int *neig, const float *psi, float *hpsi, const int *lnewanl,
                     const int *ng, const int *kgv, const float *ekpg, const int *isort, const float *vscr,
                     const int *kmscr, const float *anlga, const float *xnlkb, const int *nanl, const int *mxddim,
                     const int *mxdbnd, const int *mxdanl, const int *mxdgve, const int *mxdscr) {
    (void)lnewanl;
    (void)mxdbnd;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!mtxd || !neig || !ng || !kmscr || !nanl || !mxddim || !mxdanl || !mxdgve || !mxdscr)
        return fail(CPWGPU_ERR_ARG, "null scalar argument");
    if (*ng > *mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d > mxdgve = %d", *ng, *mxdgve);
    int ns[3], mfft, mwrk;
    size_fft(kmscr, ns, &mfft, &mwrk);
    if (mwrk > *mxdscr) return fail(CPWGPU_ERR_MXDSCR, "mxdwrk = %d is greater than mxdscr = %d", mwrk, *mxdscr);
    if (*nanl > *mxdanl) return fail(CPWGPU_ERR_ARG, "nanl = %d > mxdanl = %d", *nanl, *mxdanl);
    if (*neig < 0) return fail(CPWGPU_ERR_ARG, "neig = %d", *neig);
    if (*mtxd < 0 || *mtxd > *mxddim) return fail(CPWGPU_ERR_MTXD, "mtxd = %d, mxddim = %d", *mtxd, *mxddim);
    if (*mtxd > 0 && !ekpg) return fail(CPWGPU_ERR_ARG, "null ekpg");
    // the plan and the potential are kept in FP64 (exact images of the FP32 values)
    std::vector<double> ek(ekpg, ekpg + std::max(*mtxd, 0));
    cpwgpu_plan_s *p = nullptr;
    int st = drop_in_plan(*mtxd, *ng, kgv, ek.data(), isort, kmscr, *mxddim, *mxdgve, &p);
    if (st) return st;
    if (*neig == 0 || *mtxd == 0) return CPWGPU_OK;
    if (!psi || !hpsi || !vscr) return fail(CPWGPU_ERR_ARG, "null array argument");
    std::vector<double> v(vscr, vscr + std::min<size_t>(p->ntot, (size_t)*mxdscr));
    v.resize(std::max<size_t>(v.size(), p->ntot), 0.0);
    if ((st = set_potential(p, v.data(), *mxdscr))) return st;
    const int m = *mtxd, n = *neig, na = *nanl;
    const size_t c8b = 8 * (size_t)m, c16b = 16 * (size_t)m;
    const size_t stage = std::max((size_t)n, (size_t)na) * c8b;
    if ((st = g_drop.c8.reserve(stage)) || (st = g_drop.psi.reserve(c16b * n)) || (st = g_drop.hpsi.reserve(c16b * n)))
        return st;
    auto grid = [](long tot) { return dim3((unsigned)((tot + 255) / 256)); };
    // projectors: complex64 -> the plan's complex128 copy (the FP32 path rounds them back exactly);
    // host staging vectors live until the final synchronisation
    std::vector<double> x(xnlkb ? xnlkb : (const float *)nullptr, xnlkb ? xnlkb + na : (const float *)nullptr);
    if (na > 0 && (!anlga || !xnlkb)) return fail(CPWGPU_ERR_ARG, "null projector array");
    st = replace_projectors(p, na, [&](double2 *anl, long lda) -> int {
        HIPCHK(hipMemcpy2DAsync(g_drop.c8.p, c8b, anlga, 8 * (size_t)*mxddim, c8b, na, hipMemcpyHostToDevice, g_stream));
        hipLaunchKernelGGL(k_to_c128, grid((long)m * na), dim3(256), 0, g_stream, (const cplx<float> *)g_drop.c8.p,
                           (long)m, anl, lda, m, na);
        HIPCHK(hipGetLastError());
        return upload(p->xnl, x.data(), na);
    });
    if (st) return st;
    HIPCHK(hipMemcpy2DAsync(g_drop.c8.p, c8b, psi, 8 * (size_t)*mxddim, c8b, n, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(k_to_c128, grid((long)m * n), dim3(256), 0, g_stream, (const cplx<float> *)g_drop.c8.p,
                       (long)m, (double2 *)g_drop.psi.p, (long)m, m, n);
    HIPCHK(hipGetLastError());
    if ((st = apply_t<float>(p, n, (const double *)g_drop.psi.p, m, (double *)g_drop.hpsi.p, m, g_stream))) return st;
    hipLaunchKernelGGL(k_to_c64, grid((long)m * n), dim3(256), 0, g_stream, (const double2 *)g_drop.hpsi.p,
                       (long)m, (cplx<float> *)g_drop.c8.p, (long)m, m, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy2DAsync(hpsi, 8 * (size_t)*mxddim, g_drop.c8.p, c8b, c8b, n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return CPWGPU_OK;
}

int cpwgpu_plan_ditsp(cpwgpu_plan_t plan, int prec, int icmax, int iguess, double epspsi, int neig,
                      int mtxds, const double *hamsm_dev, long ldhs, const double *hdiag_dev,
                      double *psi_dev, long ldp, double *hpsi_dev, long ldh, double *ei, int *ifail,
                      int *stats, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan) return fail(CPWGPU_ERR_ARG, "null plan");
    if (prec != CPWGPU_PREC_FP64 && prec != CPWGPU_PREC_FP32) return fail(CPWGPU_ERR_MODE, "precision %d", prec);
    if (ifail) *ifail = 0;
    if (neig < 0 || neig > plan->mtxd) return fail(CPWGPU_ERR_ARG, "neig = %d, mtxd = %d", neig, plan->mtxd);
    if (neig == 0) return CPWGPU_OK;
    if (!psi_dev || !hpsi_dev || !ei || ldp < plan->mtxd || ldh < plan->mtxd)
        return fail(CPWGPU_ERR_ARG, "bad psi/hpsi/ei arguments");
    if (icmax < 0 || !(epspsi > 0.0)) return fail(CPWGPU_ERR_ARG, "icmax = %d, epspsi = %g", icmax, epspsi);
    if (hamsm_dev && (mtxds < 1 || ldhs < mtxds))
        return fail(CPWGPU_ERR_ARG, "hamsm given with mtxds = %d, ldhs = %ld", mtxds, ldhs);
    int st = ensure_stream();
    if (st) return st;
    return plan_ditsp(plan, prec, icmax, iguess, epspsi, neig, mtxds, (const double2 *)hamsm_dev, ldhs, hdiag_dev,
                      (double2 *)psi_dev, ldp, (double2 *)hpsi_dev, ldh, ei, ifail, stats, (hipStream_t)stream);
}

int cpwgpu_ditsp_c16(const int *ipr, int *ifail, const int *icmax, const int *iguess, const double *epspsi,
                     const int *lnewanl, int *neig, const int *mtxd, const int *mtxds, double *psi,
                     double *hpsi, double *ei, const double *ekpg, const int *isort, const double *vscr,
                     const int *kmscr, const int *ng, const int *kgv, const double *anlga, const double *xnlkb,
                     const int *nanl, const double *hamsm, const double *hdiag, const int *mxddim,
                     const int *mxdsml, const int *mxdbnd, const int *mxdgve, const int *mxdscr,
                     const int *mxdanl) {
    (void)lnewanl;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ipr || !ifail || !icmax || !iguess || !epspsi || !neig || !mtxd || !mtxds || !kmscr || !ng || !nanl ||
        !mxddim || !mxdsml || !mxdbnd || !mxdgve || !mxdscr || !mxdanl)
        return fail(CPWGPU_ERR_ARG, "null scalar argument");
    *ifail = 0;
    // the argument checks of Src/ditsp/ditsp_c16.f90:151-212, in the same order
    if (*neig < 1) {
        printf("\n   WARNING:   ditsp_c16 called with number of eigenvalues = %8d\n", *neig);
        return CPWGPU_OK;
    }
    if (*neig > *mtxds) return fail(CPWGPU_ERR_ARG, "neig = %d is greater than mtxds = %d", *neig, *mtxds);
    if (*neig > *mxdbnd) return fail(CPWGPU_ERR_ARG, "neig = %d is greater than mxdbnd = %d", *neig, *mxdbnd);
    if (3 * *neig > *mxdsml)
        return fail(CPWGPU_ERR_ARG, "3*neig = %d is greater than mxdsml = %d", 3 * *neig, *mxdsml);
    if (*mtxds > *mxdsml) return fail(CPWGPU_ERR_ARG, "mtxds = %d is greater than mxdsml = %d", *mtxds, *mxdsml);
    if (*mtxd > *mxddim) return fail(CPWGPU_ERR_MTXD, "mtxd = %d is greater than mxddim = %d", *mtxd, *mxddim);
    if (*nanl > *mxdanl) return fail(CPWGPU_ERR_ARG, "nanl = %d is greater than mxdanl = %d", *nanl, *mxdanl);
    if (*ng > *mxdgve) return fail(CPWGPU_ERR_NG, "ng = %d is greater than mxdgve = %d", *ng, *mxdgve);
    if (!psi || !hpsi || !ei || !ekpg || !isort || !vscr || !kgv || !hamsm || !hdiag)
        return fail(CPWGPU_ERR_ARG, "null array argument");
    if (*mtxds > DT_MAX_SMALL) return fail(CPWGPU_ERR_SIZE, "mtxds = %d > %d", *mtxds, DT_MAX_SMALL);
    cpwgpu_plan_s *p = nullptr;
    int st = drop_in_plan(*mtxd, *ng, kgv, ekpg, isort, kmscr, *mxddim, *mxdgve, &p);
    if (st) return st;
    if ((st = set_potential(p, vscr, *mxdscr))) return st;
    if ((st = set_projectors_host(p, anlga, xnlkb, *nanl, *mxddim))) return st;
    const int n = *neig, m = *mtxd, ms = *mtxds;
    const size_t colb = 16 * (size_t)m;
    if ((st = g_drop.psi.reserve(colb * n)) || (st = g_drop.hpsi.reserve(colb * n)) ||
        (st = g_drop.hs.reserve(16 * (size_t)ms * ms)) || (st = g_drop.hd.reserve(8 * (size_t)m)))
        return st;
    if (*iguess == 1)
        HIPCHK(hipMemcpy2DAsync(g_drop.psi.p, colb, psi, 16 * (size_t)*mxddim, colb, n, hipMemcpyHostToDevice,
                                g_stream));
    HIPCHK(hipMemcpy2DAsync(g_drop.hs.p, 16 * (size_t)ms, hamsm, 16 * (size_t)*mxdsml, 16 * (size_t)ms, ms,
                            hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(g_drop.hd.p, hdiag, 8 * (size_t)m, hipMemcpyHostToDevice, g_stream));
    int stats[5] = {0, 0, 0, 0, -1};
    double2 *dpsi = (double2 *)g_drop.psi.p, *dhpsi = (double2 *)g_drop.hpsi.p;
    if ((st = plan_ditsp(p, g_prec, *icmax, *iguess, *epspsi, n, ms, (const double2 *)g_drop.hs.p, ms,
                         (const double *)g_drop.hd.p, dpsi, m, dhpsi, m, ei, ifail, stats, g_stream)))
        return st;
    HIPCHK(hipMemcpy2DAsync(psi, 16 * (size_t)*mxddim, dpsi, colb, colb, n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpy2DAsync(hpsi, 16 * (size_t)*mxddim, dhpsi, colb, colb, n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    // messages of Src/ditsp/ditsp_c16.f90:386-419
    if (ms >= m) {
        if (*ipr > 1) printf("\n  Using standard diagonalization  mtxd = %8d\n\n", m);
    } else if (*ifail == 0) {
        if (*ipr > 1) printf("\n  Successful diagonalization after%8d matrix-vector multiplications\n\n", stats[1]);
    } else if (*ipr > 0) {
        printf("\n  Unsuccessful diagonalization after%8d matrix-vector multiplications (ifail = %d)\n\n",
               stats[1], *ifail);
    }
    return CPWGPU_OK;
}

}  // extern "C"
