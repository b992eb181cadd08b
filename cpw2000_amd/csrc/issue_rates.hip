// Issue cost of the VALU instruction kinds that make up the FFT kernels, on
// the GPU itself (measurement infrastructure for the plane-stage bound, not
// part of the H|psi> path).  Each kernel runs ITER x 32 instructions of one
// kind per wave in 8 independent chains, so the SIMD issues back to back.
// The cost per wave instruction and SIMD is taken from the difference of two
// launch sizes (launch overhead cancels) and the shader clock measured in the
// kernel (clock64 against the constant-rate wall clock).
//
// Buffers: out[nblk * T] (one value per thread, keeps the chains live) and
// cyc[2 * nblk] (clock and wall-clock deltas of each block's thread 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int ITER = 256;
constexpr int T = 256;
constexpr int NKIND = 11;

template <int K>
__global__ __launch_bounds__(T) void k_issue(double *out, long long *cyc, double seed) {
    const int gid = blockIdx.x * T + threadIdx.x;
    double d[8];
    float f[8];
    int u[8];
    for (int i = 0; i < 8; i++) {
        d[i] = seed + gid + i;
        f[i] = (float)d[i];
        u[i] = gid + i;
    }
    const double db = 1.0000001, dc = 1e-9;
    const float fb = 1.0000001f, fc = 1e-9f;
    __syncthreads();
    const long long w0 = wall_clock64();
    const long long c0 = clock64();
    for (int it = 0; it < ITER; it++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if constexpr (K == 0) d[i] = __builtin_fma(d[i], db, dc);   // v_fma_f64
                if constexpr (K == 1) d[i] = d[i] + dc;                     // v_add_f64
                if constexpr (K == 2) f[i] = __builtin_fmaf(f[i], fb, fc);  // v_fma_f32
                if constexpr (K == 3) asm volatile("v_add_u32 %0, %0, %1" : "+v"(u[i]) : "v"(it));
                if constexpr (K == 4) u[i] = __builtin_amdgcn_mov_dpp(u[i], 0xB1, 0xF, 0xF, true);
                if constexpr (K == 5) {  // FP64 FMA and DPP move interleaved
                    d[i] = __builtin_fma(d[i], db, dc);
                    u[i] = __builtin_amdgcn_mov_dpp(u[i], 0xB1, 0xF, 0xF, true);
                }
                if constexpr (K == 6) asm volatile("v_add_f32 %0, %0, %1" : "+v"(f[i]) : "v"(fc));
                if constexpr (K == 7)  // the partner lane's value as the DPP source of the add
                    asm volatile("v_add_f32_dpp %0, %1, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1"
                                 : "+v"(f[i]) : "v"(fc));
                if constexpr (K == 10)  // FP32 butterfly: the partner's value as the DPP multiplicand
                    asm volatile("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
                                 : "+v"(f[i]) : "v"(fc));
                if constexpr (K == 8 || K == 9) {  // gfx950 half-swaps of two registers, 32 / 16 lanes apart
                    const int i2 = (i + 4) & 7;
                    if constexpr (K == 8) {
                        auto pr = __builtin_amdgcn_permlane32_swap((unsigned)u[i], (unsigned)u[i2], false, false);
                        u[i] = (int)pr[0];
                        u[i2] = (int)pr[1];
                    } else {
                        auto pr = __builtin_amdgcn_permlane16_swap((unsigned)u[i], (unsigned)u[i2], false, false);
                        u[i] = (int)pr[0];
                        u[i2] = (int)pr[1];
                    }
                }
            }
        }
    }
    const long long c1 = clock64();
    const long long w1 = wall_clock64();
    __syncthreads();
    double s = 0;
    for (int i = 0; i < 8; i++) s += d[i] + f[i] + u[i];
    out[gid] = s;
    if (threadIdx.x == 0) {
        cyc[2 * blockIdx.x] = c1 - c0;
        cyc[2 * blockIdx.x + 1] = w1 - w0;
    }
}

template <int K> constexpr int per_wave() { return (K == 5 ? 2 : 1) * ITER * 4 * 8; }

// best of reps launches (ms); accumulates the in-kernel clock ratio
template <int K>
hipError_t timed(int nblk, double *out, long long *cyc, int reps, float *best, double *clk, double *wall) {
    hipEvent_t a, b;
    hipError_t e;
    if ((e = hipEventCreate(&a)) != hipSuccess) return e;
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return e;
    }
    *best = 3.4e38f;
    hipLaunchKernelGGL(k_issue<K>, dim3(nblk), dim3(T), 0, 0, out, cyc, 0.5);  // warm-up
    for (int r = 0; r < reps; r++) {
        (void)hipEventRecord(a, 0);
        hipLaunchKernelGGL(k_issue<K>, dim3(nblk), dim3(T), 0, 0, out, cyc, 0.5);
        (void)hipEventRecord(b, 0);
        if ((e = hipEventSynchronize(b)) != hipSuccess) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        *best = std::min(*best, ms);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) {
        std::vector<long long> h(2 * (size_t)nblk);
        e = hipMemcpy(h.data(), cyc, sizeof(long long) * h.size(), hipMemcpyDeviceToHost);
        for (int i = 0; i < nblk; i++) {
            *clk += (double)h[2 * i];
            *wall += (double)h[2 * i + 1];
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e;
}

template <int K>
hipError_t measure(int nsimd, double wall_hz, double *out, long long *cyc, double *cost, double *ghz) {
    const int n1 = 4 * nsimd / 4, n2 = 2 * n1;  // 4 and 8 waves per SIMD
    float t1, t2;
    double clk = 0, wall = 0;
    hipError_t e = timed<K>(n1, out, cyc, 3, &t1, &clk, &wall);
    if (e == hipSuccess) e = timed<K>(n2, out, cyc, 3, &t2, &clk, &wall);
    if (e != hipSuccess) return e;
    *ghz = wall > 0 ? clk / wall * wall_hz * 1e-9 : 0.0;
    // (n2 - n1) blocks of 4 waves more, spread over nsimd SIMDs
    const double dinstr = (double)(n2 - n1) * (T / 64) * per_wave<K>() / nsimd;
    *cost = (double)(t2 - t1) * 1e-3 * (*ghz * 1e9) / dinstr;
    return hipSuccess;
}

}  // namespace

extern "C" {

// res[0..9]: cycles per wave instruction per SIMD of v_fma_f64, v_add_f64,
// v_fma_f32, v_add_u32, v_mov_b32_dpp, the average of an interleaved
// v_fma_f64 / v_mov_b32_dpp stream, v_add_f32, v_add_f32 with a DPP source
// operand, v_permlane32_swap_b32 and v_permlane16_swap_b32 (each rewrites two
// registers), s_nop 1 + v_fmac_f32 with a DPP source operand (the FP32
// butterfly); res[11]: shader clock (GHz) during the runs; res[12]: number of
// SIMDs.  nres >= 13.  Returns 0, or the HIP error code.
int cpwprobe_issue_rates(double *res, int nres) {
    if (!res || nres < NKIND + 2) return (int)hipErrorInvalidValue;
    int dev = 0, ncu = 0, wall_khz = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev);
    if (e != hipSuccess) return (int)e;
    const int nsimd = 4 * ncu;
    const size_t nmax = 2 * (size_t)nsimd;  // blocks of the larger launch
    double *out = nullptr;
    long long *cyc = nullptr;
    if ((e = hipMalloc(&out, sizeof(double) * nmax * T)) != hipSuccess) return (int)e;
    if ((e = hipMalloc(&cyc, sizeof(long long) * 2 * nmax)) != hipSuccess) {
        (void)hipFree(out);
        return (int)e;
    }
    const double wall_hz = 1e3 * wall_khz;
    double ghz[NKIND] = {0};
    if (e == hipSuccess) e = measure<0>(nsimd, wall_hz, out, cyc, &res[0], &ghz[0]);
    if (e == hipSuccess) e = measure<1>(nsimd, wall_hz, out, cyc, &res[1], &ghz[1]);
    if (e == hipSuccess) e = measure<2>(nsimd, wall_hz, out, cyc, &res[2], &ghz[2]);
    if (e == hipSuccess) e = measure<3>(nsimd, wall_hz, out, cyc, &res[3], &ghz[3]);
    if (e == hipSuccess) e = measure<4>(nsimd, wall_hz, out, cyc, &res[4], &ghz[4]);
    if (e == hipSuccess) e = measure<5>(nsimd, wall_hz, out, cyc, &res[5], &ghz[5]);
    if (e == hipSuccess) e = measure<6>(nsimd, wall_hz, out, cyc, &res[6], &ghz[6]);
    if (e == hipSuccess) e = measure<7>(nsimd, wall_hz, out, cyc, &res[7], &ghz[7]);
    if (e == hipSuccess) e = measure<8>(nsimd, wall_hz, out, cyc, &res[8], &ghz[8]);
    if (e == hipSuccess) e = measure<9>(nsimd, wall_hz, out, cyc, &res[9], &ghz[9]);
    if (e == hipSuccess) e = measure<10>(nsimd, wall_hz, out, cyc, &res[10], &ghz[10]);
    double g = 0;
    for (int k = 0; k < NKIND; k++) g += ghz[k] / NKIND;
    res[NKIND] = g;
    res[NKIND + 1] = nsimd;
    (void)hipFree(out);
    (void)hipFree(cyc);
    return (int)e;
}

}  // extern "C"
