// Register-resident plane transforms: each line of length N = G*M lives in the
// registers of G adjacent lanes (M values per lane).  A transform is an
// in-register DFT_M, a lane-dependent twiddle and a DFT_G across the G lanes
// done with DPP butterflies (no LDS, no barrier).  Used for the y and x passes
// of the per-(z, band) plane stage of H|psi>: only the y/x-box sub-plane is
// kept in LDS, so more blocks fit per CU than with LDS-exchanged passes.
//
// Layouts, for a line f[0..N) and the lane at position g < G of its group:
//   DIF  in : v[j]  = f[g + G j]           out: v[k1] = F[k1 + M bitrev(g)]
//   DIT  in : v[k1] = f[k1 + M bitrev(g)]  out: v[m]  = F[g + G m]
// so a DIT transform consumes the DIF output directly, and the DIF output
// gives each lane M consecutive points (contiguous V(r) loads).
//
// DPP patterns: position xor 1 and xor 2 are quad permutations; position xor 4
// (G = 8) uses row_half_mirror (lane i <-> 7 - i within 8 lanes) with the lane
// to position map g = L for L < 4, g = 11 - L for L >= 4, under which the
// mirror pairs are exactly the positions g, g ^ 4, and xor 1 / xor 2 of the
// position remain xor 1 / xor 2 of the lane.
#pragma once

#include "fft_fast.hpp"
#include "fft_wave.hpp"

namespace cpwgpu {

template <int H> constexpr int dpp_ctrl() { return H == 1 ? 0xB1 : H == 2 ? 0x4E : 0x141; }

// Every partner lane is in range and active, so the moves need no "old" value
// (bound_ctrl set, all rows and banks enabled): no extra register copy.
template <int CTRL> __device__ __forceinline__ float xlane(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ double xlane(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

constexpr int ilog2(int g) { return g <= 1 ? 0 : 1 + ilog2(g / 2); }
constexpr int bitrev(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

// Per-lane constants of a G-lane group: position, butterfly signs and the
// upper-half twiddles W_{2h}^{(g mod 2h) - h} (backward sign) of each stage,
// and for the one-instruction FP32 butterflies (bfly_t) the same twiddles
// times the sign that the butterfly before them leaves on the upper lanes.
template <typename R, int G> struct Lanes {
    int g;    // position in the group
    int k2;   // bitrev(g): DIF output / DIT input block index
    R s1, s2, s4, t1n;
    cplx<R> w2, w4;
    cplx<R> w2s, w4s;  // DIF: s2 w2, s4 w4
    cplx<R> w2d, w4d;  // DIT: s1 w2, s2 w4
    __device__ __forceinline__ Lanes() {
        const int l = (int)(threadIdx.x % G);
        g = (G == 8 && l >= 4) ? 11 - l : l;
        k2 = 0;
        for (int i = 0; i < ilog2(G); i++) k2 |= ((g >> i) & 1) << (ilog2(G) - 1 - i);
        s1 = (g & 1) ? (R)-1 : (R)1;
        s2 = (g & 2) ? (R)-1 : (R)1;
        s4 = (g & 4) ? (R)-1 : (R)1;
        // W_4^q, q = (g mod 4) - 2 in {0, 1} for the upper half of a 4-block
        w2 = ((g & 3) == 3) ? mk<R>(0, 1) : mk<R>(1, 0);
        // W_8^q, q = (g mod 8) - 4 in {0..3} for the upper half of an 8-block
        const R h = (R)0.70710678118654752440;
        const int q = g - 4;
        w4 = q == 1 ? mk<R>(h, h) : q == 2 ? mk<R>(0, 1) : q == 3 ? mk<R>(-h, h) : mk<R>(1, 0);
        t1n = -s1;
        w2s = mk<R>(s2 * w2.x, s2 * w2.y);
        w4s = mk<R>(s4 * w4.x, s4 * w4.y);
        w2d = mk<R>(s1 * w2.x, s1 * w2.y);
        w4d = mk<R>(s2 * w4.x, s2 * w4.y);
    }
};

// A value the compiler cannot see through: twiddle loads indexed by it stay
// inside the loop that uses them (LDS reads) instead of being hoisted and held
// in registers for the whole kernel.
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

template <int H, typename R> __device__ __forceinline__ cplx<R> bfly(cplx<R> m, R s) {
    constexpr int C = dpp_ctrl<H>();
    if constexpr (sizeof(R) == 4) {
        // partner + (s * mine) with the product rounded first (no contraction
        // into an FMA; exact, s = +-1): the lane move then folds into the add
        // as its DPP source operand (a multiply and a DPP-sourced add instead
        // of a DPP move and an FMA)
#pragma clang fp contract(off)
        const R tx = s * m.x, ty = s * m.y;
        return mk<R>(xlane<C>(m.x) + tx, xlane<C>(m.y) + ty);
    } else {
        return mk<R>(xlane<C>(m.x) + s * m.x, xlane<C>(m.y) + s * m.y);  // partner + s * mine
    }
}
// FP32 butterfly in one instruction: mine + t * partner, a v_fmac_f32 whose
// multiplicand is the partner lane's value (DPP source operand) and whose
// addend is the lane's own.  t = +-1, so the product is exact and the result
// is the rounded sum or difference, as in bfly.  On the lane pair (a, b) of a
// stage, a = lower position, holding sa x and sb y (sa, sb = +-1 known per
// lane): t_a = sa sb, t_b = -sa sb give a: sa (x + y) and b: -sb (x - y), so
// with equal signs in, the upper lane's output is negated; the caller folds
// that sign into the next stage's twiddle (Lanes::w2s, w4s, w2d, w4d) or into
// the first-stage t of the transform that consumes it (line_dit_tw NEGIN).
// The s_nop gives the two wait states that a DPP read of a register written by
// the preceding VALU instruction needs (the compiler does not track hazards
// inside inline assembly); the five that follow an EXEC write are met by the
// transform code before every butterfly (loads, DFT_M, twiddles).  Both rules
// are checked on the built library by tests/test_dpp_hazards.py.
template <int H> __device__ __forceinline__ float fmac_xlane(float mine, float t) {
    if constexpr (H == 1)
        asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
            : "+v"(mine) : "v"(t));
    else if constexpr (H == 2)
        asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf"
            : "+v"(mine) : "v"(t));
    else
        asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 row_half_mirror row_mask:0xf bank_mask:0xf"
            : "+v"(mine) : "v"(t));
    return mine;
}
template <int H> __device__ __forceinline__ cplx<float> bfly_t(cplx<float> m, float t) {
    return mk<float>(fmac_xlane<H>(m.x, t), fmac_xlane<H>(m.y, t));
}
// Lines that use bfly_t: FP32 with 4 or 8 lanes, up to length 72.  Its sign
// constants and asm operands take about 5 more registers per lane; with 2
// lanes (36 values each) and the longer lines (80-128) the plane kernels then
// spill, so those keep bfly.
template <typename R, int G, int M> constexpr bool use_bfly_t() {
    return sizeof(R) == 4 && G >= 4 && G * M <= 72;
}

template <int DIR, typename R> __device__ __forceinline__ cplx<R> cmul_dir(cplx<R> a, cplx<R> w) {
    return cmul(a, DIR > 0 ? w : mk<R>(w.x, -w.y));
}

// DIF transform of a G-lane line (layouts above) with the lane twiddles
// tw(k) = exp(+2 pi i g k / N), k = 1..M-1.  IN: the register slots j that can
// hold a nonzero value (the others are zero on every lane and are not read).
// Lane twiddle sources: TwLds reads twN[g k] from LDS at each use (its index
// is made opaque where begin() is called, so the loads stay at their uses and
// nothing is held in registers across transforms); TwRegs returns values
// already in registers.
template <typename R> struct TwLds {
    const cplx<R> *twN;
    int g, gi = 0;
    __device__ __forceinline__ void begin() { gi = opaque(g); }
    __device__ __forceinline__ cplx<R> operator()(int k) const { return twN[gi * k]; }
};
template <typename R, int M> struct TwRegs {
    const cplx<R> (&tw)[M];
    __device__ __forceinline__ void begin() {}
    __device__ __forceinline__ cplx<R> operator()(int k) const { return tw[k]; }
};

// NEGOUT (lines with bfly_t, else ignored): the last stage also uses bfly_t
// and leaves the values of the odd positions negated, for a line_dit_tw with
// NEGIN that consumes them directly (a pointwise real factor in between keeps
// the pattern).
template <int DIR, int G, int M, typename R, u64mask IN, bool NEGOUT = false, typename TW>
__device__ __forceinline__ void line_dif_tw(cplx<R> (&v)[M], const Lanes<R, G> &L, TW tw) {
    DftP<M, DIR, R, IN, full_mask(M)>::run(v);
    if constexpr (G > 1) {
        tw.begin();
#pragma unroll
        for (int k = 1; k < M; k++) v[k] = cmul_dir<DIR>(v[k], tw(k));
        if constexpr (use_bfly_t<R, G, M>()) {
            // one-instruction butterflies; each upper-lane sign goes into the
            // stage twiddle that follows
            if constexpr (G >= 8) {
#pragma unroll
                for (int k = 0; k < M; k++) v[k] = cmul_dir<DIR>(bfly_t<4>(v[k], L.s4), L.w4s);
            }
            if constexpr (G >= 4) {
#pragma unroll
                for (int k = 0; k < M; k++) v[k] = cmul_dir<DIR>(bfly_t<2>(v[k], L.s2), L.w2s);
            }
            if constexpr (NEGOUT) {
#pragma unroll
                for (int k = 0; k < M; k++) v[k] = bfly_t<1>(v[k], L.s1);
            } else {
#pragma unroll
                for (int k = 0; k < M; k++) v[k] = bfly<1>(v[k], L.s1);
            }
        } else {
            if constexpr (G >= 8) {
#pragma unroll
                for (int k = 0; k < M; k++) v[k] = cmul_dir<DIR>(bfly<4>(v[k], L.s4), L.w4);
            }
            if constexpr (G >= 4) {
#pragma unroll
                for (int k = 0; k < M; k++) v[k] = cmul_dir<DIR>(bfly<2>(v[k], L.s2), L.w2);
            }
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = bfly<1>(v[k], L.s1);
        }
    }
}

// line_dif with the lane twiddles read from twN[e] = exp(+2 pi i e/N) as
// they are used (nothing held in registers across calls).
template <int DIR, int G, int M, typename R, u64mask IN = full_mask(M), bool NEGOUT = false>
__device__ __forceinline__ void line_dif(cplx<R> (&v)[M], const Lanes<R, G> &L,
                                         const cplx<R> *twN) {
    line_dif_tw<DIR, G, M, R, IN, NEGOUT>(v, L, TwLds<R>{twN, L.g});
}

// DIT transform of a G-lane line, the transpose of line_dif's structure.
// OUT: the output slots m the caller uses (the others are left undefined).
// NEGIN (lines with bfly_t, else ignored): the odd positions hold negated
// values (line_dif_tw NEGOUT).
template <int DIR, int G, int M, typename R, u64mask OUT, bool NEGIN = false, typename TW>
__device__ __forceinline__ void line_dit_tw(cplx<R> (&v)[M], const Lanes<R, G> &L, TW tw) {
    if constexpr (G > 1 && use_bfly_t<R, G, M>()) {
        // one-instruction butterflies except in the last stage (its output
        // meets the lane twiddles, of which the first is not multiplied);
        // with NEGIN, t = -s1 in the first stage gives exact outputs, else its
        // odd outputs are negated and w2d = s1 w2 restores them; the second
        // stage's upper-lane sign goes into w4d = s2 w4
#pragma unroll
        for (int k = 0; k < M; k++) v[k] = bfly_t<1>(v[k], NEGIN ? L.t1n : L.s1);
        if constexpr (G >= 8) {
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = bfly_t<2>(cmul_dir<DIR>(v[k], NEGIN ? L.w2 : L.w2d), L.s2);
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = bfly<4>(cmul_dir<DIR>(v[k], L.w4d), L.s4);
        } else if constexpr (G >= 4) {
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = bfly<2>(cmul_dir<DIR>(v[k], NEGIN ? L.w2 : L.w2d), L.s2);
        }
        tw.begin();
#pragma unroll
        for (int k = 1; k < M; k++) v[k] = cmul_dir<DIR>(v[k], tw(k));
    } else if constexpr (G > 1) {
#pragma unroll
        for (int k = 0; k < M; k++) v[k] = bfly<1>(v[k], L.s1);
        if constexpr (G >= 4) {
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = bfly<2>(cmul_dir<DIR>(v[k], L.w2), L.s2);
        }
        if constexpr (G >= 8) {
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = bfly<4>(cmul_dir<DIR>(v[k], L.w4), L.s4);
        }
        tw.begin();
#pragma unroll
        for (int k = 1; k < M; k++) v[k] = cmul_dir<DIR>(v[k], tw(k));
    }
    DftP<M, DIR, R, full_mask(M), OUT>::run(v);
}

template <int DIR, int G, int M, typename R, u64mask OUT = full_mask(M), bool NEGIN = false>
__device__ __forceinline__ void line_dit(cplx<R> (&v)[M], const Lanes<R, G> &L,
                                         const cplx<R> *twN) {
    line_dit_tw<DIR, G, M, R, OUT, NEGIN>(v, L, TwLds<R>{twN, L.g});
}

// z-column values of one plane between colbuf (cb: ncol values) and the
// sub-plane sp (row stride S; col_uy[c] = y << 16 | u), from column c_begin,
// four columns of each thread at a time with their global loads issued
// together (one round trip per four columns, not one per column).  Clamped
// loads: ncol >= 1.
template <typename R, int T>
__device__ __forceinline__ void gather_plane(cplx<R> *sp, const cplx<R> *__restrict__ cb,
                                             const int *__restrict__ col_uy, int ncol, int S, int c_begin) {
    struct Pre { int uy; cplx<R> v; };
    for (int c0 = c_begin + (int)threadIdx.x; c0 < ncol; c0 += 4 * T) {
        const auto ld = [&](int u) {
            const int c = min(c0 + u * T, ncol - 1);
            return Pre{col_uy[c], cb[c]};
        };
        const Pre q0 = ld(0), q1 = ld(1), q2 = ld(2), q3 = ld(3);
        const auto put = [&](int u, const Pre &q) {
            if (c0 + u * T < ncol) sp[(q.uy >> 16) * S + (q.uy & 0xffff)] = q.v;
        };
        put(0, q0);
        put(1, q1);
        put(2, q2);
        put(3, q3);
    }
}
template <typename R, int T>
__device__ __forceinline__ void scatter_plane(cplx<R> *__restrict__ cb, const cplx<R> *sp,
                                              const int *__restrict__ col_uy, int ncol, int S) {
    for (int c0 = threadIdx.x; c0 < ncol; c0 += 4 * T) {
        const int u0 = col_uy[min(c0, ncol - 1)], u1 = col_uy[min(c0 + T, ncol - 1)];
        const int u2 = col_uy[min(c0 + 2 * T, ncol - 1)], u3 = col_uy[min(c0 + 3 * T, ncol - 1)];
        const auto put = [&](int u, int uy) {
            if (c0 + u * T < ncol) cb[c0 + u * T] = sp[(uy >> 16) * S + (uy & 0xffff)];
        };
        put(0, u0);
        put(1, u1);
        put(2, u2);
        put(3, u3);
    }
}

// ---------------------------------------------------------------------------
// plane stage, register variant: a block takes P planes (bands b0..b0+P-1 at
// one z) with T threads and a register budget for W waves per SIMD.  LDS:
// sub[P][N][S] (y rows of the x-box: nx box columns, a zero column and a trash
// column, S = nx + 2) and the twiddle table twN[N].  Phases (barrier between each):
//   gather columns -> inverse y (GY-lane lines over u) -> per row y: inverse
//   x, V(r), forward x (GX-lane lines, no barrier) -> forward y -> scatter.
// YW: the y lines instead go through the LDS two-pass transform in place in
// sub, each line exchanged within one wavefront (line_fft_wave; GY unused).
// PR: the sphere box is |x|, |y| <= N/4 (checked by the caller), so the
// register slots that hold only x (y) outside the box on every lane are known
// zeros of the inverse transforms and unused outputs of the forward ones:
// they are neither loaded, transformed nor stored.  The box then has at most
// N/2 + 1 columns, and the sub-plane row stride is the compile-time
// pr_sub_stride<N>() instead of the argument S: every LDS offset of the y
// lines and x rows folds into the instructions' immediate offsets.
// ---------------------------------------------------------------------------
template <int N> constexpr int pr_sub_stride() { return N / 2 + 3; }

template <typename R, int N, int GX, int GY, int P, int T, int W, bool YW = false, bool PR = false>
__global__ __launch_bounds__(T, W) void k_plane_reg(cplx<R> *__restrict__ colbuf,
                                                 const int *__restrict__ col_uy,
                                                 const R *__restrict__ vr,
                                                 const double2 *__restrict__ twg, FftGeom g,
                                                 int nb, int S_arg) {
    static_assert(N % GX == 0 && N % GY == 0 && T % 64 == 0, "bad plane shape");
    constexpr int MX = N / GX, MY = N / GY;
    const int S = PR ? pr_sub_stride<N>() : S_arg;
    // slots that can hold box points (all of them unless PR)
    constexpr u64mask XB = PR ? box_slots(N, GX, MX, N / 4) : full_mask(MX);
    constexpr u64mask YB = PR ? box_slots(N, GY, MY, N / 4) : full_mask(MY);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R> *twN = reinterpret_cast<cplx<R> *>(smem_raw);  // [N]
    cplx<R> *sub = twN + N;                                 // [P][N][S]
    const int n3 = g.n[2];
    const int z = blockIdx.x % n3;
    const int b0 = (blockIdx.x / n3) * P;
    const int np = min(P, nb - b0);  // planes of this block that exist
    const int nx = g.nx, kd1 = g.kd[0], ushift = N - nx, xhi = N - kd1;
    const int tid = threadIdx.x;

    // the first GPRE z-column values of each thread (first plane) are loaded
    // before the twiddles and the clearing of the sub-plane, which they
    // overlap (clamped: the kernel runs only with g.ncol >= 1); four named
    // values, not an array, so that FP64 keeps them in registers too
    constexpr int GPRE = 4, NPRE = GPRE;
    struct Pre { int uy; cplx<R> v; };
    const cplx<R> *cb0 = colbuf + ((long)b0 * n3 + z) * g.ncolp;
    const auto pre = [&](int u) {
        const int c = min(tid + u * T, g.ncol - 1);
        return Pre{col_uy[c], cb0[c]};
    };
    const Pre q0 = pre(0), q1 = pre(1), q2 = pre(2), q3 = pre(3);
    load_twiddles(twN, twg, N);
    for (int q = tid; q < P * N * S; q += T) sub[q] = mk<R>(0, 0);
    __syncthreads();
    const auto put = [&](int u, const Pre &q) {
        if (tid + u * T < g.ncol) sub[(q.uy >> 16) * S + (q.uy & 0xffff)] = q.v;
    };
    put(0, q0);
    put(1, q1);
    put(2, q2);
    put(3, q3);
    for (int p = 0; p < np; p++)
        gather_plane<R, T>(sub + p * N * S, colbuf + ((long)(b0 + p) * n3 + z) * g.ncolp, col_uy, g.ncol, S,
                           p == 0 ? NPRE * T : 0);
    __syncthreads();

    if constexpr (YW) {
        for (int p = 0; p < P; p++) line_fft_wave<R, N, +1, T>(sub + p * N * S, nx, 1, S, twN);
    } else {  // inverse y on the P * nx box lines
        const Lanes<R, GY> L;
        const int ntask = P * nx;
        for (int task = tid / GY; task < ntask; task += T / GY) {
            const int p = task / nx, u = task - p * nx;
            cplx<R> *col = sub + p * N * S + u;
            cplx<R> v[MY];
#pragma unroll
            for (int j = 0; j < MY; j++)
                if (has_bit(YB, j)) v[j] = col[(L.g + GY * j) * S];
            line_dif<+1, GY, MY, R, YB>(v, L, twN);
#pragma unroll
            for (int k = 0; k < MY; k++) col[(k + MY * L.k2) * S] = v[k];
        }
    }
    __syncthreads();

    {  // rows: inverse x, V(r), forward x
        const Lanes<R, GX> L;
        const int ntask = P * N;
        // the lane twiddles of every row: in FP64, read once into registers
        // where they fit (up to 9 per lane; all loads in flight together),
        // else read from LDS as each transform uses them.  (FP64 blocks are
        // held to 3 per CU by LDS anyway; in FP32 the 16 extra registers cost
        // a wave per SIMD and measured 5% slower.)
        constexpr bool HOIST = sizeof(R) == 8 && MX <= 9;
        cplx<R> twx[MX];
        if constexpr (HOIST) {
#pragma unroll
            for (int k = 1; k < MX; k++) twx[k] = twN[L.g * k];
        }
        const auto tw = [&]() {
            if constexpr (HOIST) return TwRegs<R, MX>{twx};
            else return TwLds<R>{twN, L.g};
        }();
        for (int task = tid / GX; task < ntask; task += T / GX) {
            const int p = task / N, y = task - p * N;
            cplx<R> *row = sub + (p * N + y) * S;
            cplx<R> v[MX];
            // x = g + GX j: box columns, or the zero column nx for the gap
#pragma unroll
            for (int j = 0; j < MX; j++) {
                const int x = L.g + GX * j;
                if (has_bit(XB, j)) v[j] = row[x <= kd1 ? x : x >= xhi ? x - ushift : nx];
            }
            // (with bfly_t the odd positions carry negated values between the two)
            line_dif_tw<+1, GX, MX, R, XB, true>(v, L, tw);
            const R *vy = vr + ((long)z * g.n[1] + y) * g.id + MX * L.k2;
#pragma unroll
            for (int k = 0; k < MX; k++) v[k] = scale(v[k], vy[k]);
            line_dit_tw<-1, GX, MX, R, XB, true>(v, L, tw);
            // gap columns go to the trash column nx + 1
#pragma unroll
            for (int m = 0; m < MX; m++) {
                const int x = L.g + GX * m;
                if (has_bit(XB, m)) row[x <= kd1 ? x : x >= xhi ? x - ushift : nx + 1] = v[m];
            }
        }
    }
    __syncthreads();

    if constexpr (YW) {
        for (int p = 0; p < P; p++) line_fft_wave<R, N, -1, T>(sub + p * N * S, nx, 1, S, twN);
    } else {  // forward y
        const Lanes<R, GY> L;
        const int ntask = P * nx;
        for (int task = tid / GY; task < ntask; task += T / GY) {
            const int p = task / nx, u = task - p * nx;
            cplx<R> *col = sub + p * N * S + u;
            cplx<R> v[MY];
#pragma unroll
            for (int k = 0; k < MY; k++) v[k] = col[(k + MY * L.k2) * S];
            line_dit<-1, GY, MY, R, YB>(v, L, twN);
            // rows outside the y box are not gathered back
#pragma unroll
            for (int m = 0; m < MY; m++)
                if (has_bit(YB, m)) col[(L.g + GY * m) * S] = v[m];
        }
    }
    __syncthreads();
    // scatter back, the map entries of four columns per thread loaded together
    for (int p = 0; p < np; p++)
        scatter_plane<R, T>(colbuf + ((long)(b0 + p) * n3 + z) * g.ncolp, sub + p * N * S, col_uy, g.ncol, S);
}

// LDS bytes of k_plane_reg
template <typename R, int N, int P> constexpr size_t plane_reg_lds(int S) {
    return sizeof(cplx<R>) * ((size_t)N + (size_t)P * N * S);
}

}  // namespace cpwgpu
