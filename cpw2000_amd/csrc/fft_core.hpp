// Register-resident small DFTs with compile-time twiddles (host and device).
//
//   dft<M, DIR>(v): v[m] <- sum_n v[n] exp(DIR * 2 pi i n m / M), natural order
//                   in and out, DIR = -1 forward, +1 backward, unnormalised.
//
// M in {1,2,3,4,5} are written out; a composite M = P*Q is split into P
// transforms of length Q and Q transforms of length P (Cooley-Tukey, twiddles
// exp(DIR 2 pi i s k / M) taken from a constexpr table, so after unrolling
// every twiddle is a literal).
//
// A length-N line is transformed by two such passes, N = A*B:
//   pass 1: "thread" s < A holds x[s + A j], j < B: DFT_B, then times
//           exp(DIR 2 pi i s k / N)        -> Y[s][k], stored at s + A k
//   pass 2: "thread" k < B holds Y[s][k], s < A:   DFT_A -> X[k + B q]
// and the transposed split (inner index k < B, outer q < A) runs the same
// two passes from the distribution pass 2 leaves behind.
#pragma once

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CPW_HD __host__ __device__ __forceinline__
#else
#define CPW_HD inline
#endif

namespace cpwgpu {

template <typename R> struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

template <typename R> CPW_HD cplx<R> mk(R a, R b) { cplx<R> c; c.x = a; c.y = b; return c; }
template <typename R> CPW_HD cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> CPW_HD cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> CPW_HD cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> CPW_HD cplx<R> scale(cplx<R> a, R s) { return mk<R>(a.x * s, a.y * s); }
// i * dir * a  (dir = +-1)
template <typename R> CPW_HD cplx<R> rot(cplx<R> a, int dir) {
    return dir > 0 ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// constexpr cos / sin of 2 pi k / n  (octant reduction + Taylor series)
// ---------------------------------------------------------------------------
constexpr double kTwoPi = 6.283185307179586476925286766559005768;

constexpr double taylor_sin(double x) {
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 14; i++) {
        term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; i++) {
        term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
// returns cos (want_sin = false) or sin of 2 pi num/den, 0 <= num/den reduced
constexpr double trig_frac(long num, long den, bool want_sin) {
    num %= den;
    if (num < 0) num += den;
    // [1/2, 1): cos(2pi t) = cos(2pi(1-t)), sin(2pi t) = -sin(2pi(1-t))
    if (2 * num > den) {
        double v = trig_frac(den - num, den, want_sin);
        return want_sin ? -v : v;
    }
    // (1/4, 1/2]: cos = -cos(2pi(1/2 - t)), sin = sin(2pi(1/2 - t))
    if (4 * num > den) {
        double v = trig_frac(den - 2 * num, 2 * den, want_sin);
        return want_sin ? v : -v;
    }
    // (1/8, 1/4]: cos = sin(2pi(1/4 - t)), sin = cos(2pi(1/4 - t))
    if (8 * num > den) return trig_frac(den - 4 * num, 4 * den, !want_sin);
    const double x = kTwoPi * (double)num / (double)den;
    return want_sin ? taylor_sin(x) : taylor_cos(x);
}

template <int M> struct Roots {
    double c[M], s[M];
    constexpr Roots() : c(), s() {
        for (int k = 0; k < M; k++) {
            c[k] = trig_frac(k, M, false);
            s[k] = trig_frac(k, M, true);
        }
    }
};

constexpr int first_factor(int m) {
    return m % 4 == 0 && m != 4 ? 4 : (m % 2 == 0 ? 2 : (m % 3 == 0 ? 3 : (m % 5 == 0 ? 5 : m)));
}

// ---------------------------------------------------------------------------
// small DFTs
// ---------------------------------------------------------------------------
template <int M, int DIR, typename R> struct Dft;

template <int DIR, typename R> struct Dft<1, DIR, R> {
    static CPW_HD void run(cplx<R> *) {}
};
template <int DIR, typename R> struct Dft<2, DIR, R> {
    static CPW_HD void run(cplx<R> *v) {
        cplx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};
template <int DIR, typename R> struct Dft<3, DIR, R> {
    static CPW_HD void run(cplx<R> *v) {
        const R h = (R)0.86602540378443864676;  // sin(2 pi/3)
        cplx<R> t1 = v[1] + v[2];
        cplx<R> t2 = v[0] - scale(t1, (R)0.5);
        cplx<R> t3 = rot(scale(v[1] - v[2], h), DIR);
        v[0] = v[0] + t1;
        v[1] = t2 + t3;
        v[2] = t2 - t3;
    }
};
template <int DIR, typename R> struct Dft<4, DIR, R> {
    static CPW_HD void run(cplx<R> *v) {
        cplx<R> a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = rot(v[1] - v[3], DIR);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + d;
        v[3] = b - d;
    }
};
template <int DIR, typename R> struct Dft<5, DIR, R> {
    static CPW_HD void run(cplx<R> *v) {
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        cplx<R> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
        cplx<R> a1 = v[0] + scale(t1, c1) + scale(t2, c2);
        cplx<R> a2 = v[0] + scale(t1, c2) + scale(t2, c1);
        cplx<R> b1 = rot(scale(t3, s1) + scale(t4, s2), DIR);
        cplx<R> b2 = rot(scale(t3, s2) - scale(t4, s1), DIR);
        v[0] = v[0] + t1 + t2;
        v[1] = a1 + b1;
        v[4] = a1 - b1;
        v[2] = a2 + b2;
        v[3] = a2 - b2;
    }
};

// composite M = P*Q
template <int M, int DIR, typename R> struct Dft {
    static CPW_HD void run(cplx<R> *v) {
        constexpr int P = first_factor(M);
        constexpr int Q = M / P;
        static_assert(P > 1 && P <= 5 && Q > 1, "unsupported DFT length");
        constexpr Roots<M> W{};
        cplx<R> t[M];
#pragma unroll
        for (int s = 0; s < P; s++) {
            cplx<R> u[Q];
#pragma unroll
            for (int j = 0; j < Q; j++) u[j] = v[s + P * j];
            Dft<Q, DIR, R>::run(u);
#pragma unroll
            for (int k = 0; k < Q; k++) {
                const int e = (s * k) % M;
                if (e == 0) {
                    t[s * Q + k] = u[k];
                } else {
                    const cplx<R> w = mk<R>((R)W.c[e], (R)(DIR * W.s[e]));
                    t[s * Q + k] = cmul(u[k], w);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < Q; k++) {
            cplx<R> u[P];
#pragma unroll
            for (int s = 0; s < P; s++) u[s] = t[s * Q + k];
            Dft<P, DIR, R>::run(u);
#pragma unroll
            for (int q = 0; q < P; q++) v[k + Q * q] = u[q];
        }
    }
};

template <int M, int DIR, typename R> CPW_HD void dft(cplx<R> *v) { Dft<M, DIR, R>::run(v); }

// ---------------------------------------------------------------------------
// pruned DFTs
//
//   DftP<M, DIR, R, IN, OUT>::run(v): the transform of dft<M, DIR> when the
//   inputs outside the bit mask IN are zero (they are never read) and only the
//   outputs in the bit mask OUT are used (the others are left undefined).
//   M <= 64.  Zero sub-transforms are skipped, zero terms are left out of the
//   sums, and the work behind unused outputs is never written, so the compiler
//   drops it.  Quarter-turn twiddles are exact rotations, not products.
// ---------------------------------------------------------------------------
typedef unsigned long long u64mask;

constexpr u64mask full_mask(int m) { return m >= 64 ? ~0ull : (1ull << m) - 1; }
constexpr bool has_bit(u64mask m, int b) { return (m >> b) & 1ull; }
// bits t < Q with s + P t in mask
constexpr u64mask stride_mask(u64mask m, int s, int P, int Q) {
    u64mask r = 0;
    for (int t = 0; t < Q; t++)
        if (has_bit(m, s + P * t)) r |= 1ull << t;
    return r;
}
// bits s < P whose stride-P sub-sequence has a bit in mask
constexpr u64mask live_subs(u64mask m, int P, int Q) {
    u64mask r = 0;
    for (int s = 0; s < P; s++)
        if (stride_mask(m, s, P, Q)) r |= 1ull << s;
    return r;
}
// bits k < Q with some k + Q q (q < P) in mask
constexpr u64mask needed_cols(u64mask m, int P, int Q) {
    u64mask r = 0;
    for (int k = 0; k < Q; k++)
        for (int q = 0; q < P; q++)
            if (has_bit(m, k + Q * q)) r |= 1ull << k;
    return r;
}

constexpr int lowest_bit(u64mask m) {
    int b = 0;
    while (b < 64 && !has_bit(m, b)) b++;
    return b;
}

template <int I, int N, typename F> CPW_HD void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

template <int M> struct RootTable {
    static constexpr Roots<M> W{};
};

// a * exp(DIR 2 pi i E / M), E a compile-time exponent
template <int M, int E, int DIR, typename R> CPW_HD cplx<R> twiddle(cplx<R> a) {
    constexpr int e = ((E % M) + M) % M;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (2 * e == M) {
        return mk<R>(-a.x, -a.y);
    } else if constexpr (4 * e == M) {
        return rot(a, DIR);
    } else if constexpr (4 * e == 3 * M) {
        return rot(a, -DIR);
    } else {
        return cmul(a, mk<R>((R)RootTable<M>::W.c[e], (R)(DIR * RootTable<M>::W.s[e])));
    }
}

// leaf: direct sums over the nonzero inputs (the optimised leaf when all are)
template <int P, int DIR, typename R, u64mask IN> CPW_HD void dft_leaf_masked(cplx<R> *u) {
    if constexpr (IN == full_mask(P)) {
        Dft<P, DIR, R>::run(u);
    } else {
        cplx<R> out[P];
        static_for<0, P>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            cplx<R> acc = mk<R>(0, 0);
            static_for<0, P>([&](auto s_) {
                constexpr int s = decltype(s_)::value;
                if constexpr (has_bit(IN, s)) {
                    const cplx<R> t = twiddle<P, s * q, DIR, R>(u[s]);
                    if constexpr (s == lowest_bit(IN)) acc = t;
                    else acc = acc + t;
                }
            });
            out[q] = acc;
        });
#pragma unroll
        for (int q = 0; q < P; q++) u[q] = out[q];
    }
}

template <int M, int DIR, typename R, u64mask IN, u64mask OUT> struct DftP {
    static CPW_HD void run(cplx<R> *v) {
        static_assert(M <= 64, "pruned DFT longer than 64");
        if constexpr (IN == full_mask(M) && OUT == full_mask(M)) {
            Dft<M, DIR, R>::run(v);
        } else if constexpr (M <= 5) {
            dft_leaf_masked<M, DIR, R, IN>(v);
        } else {
            constexpr int P = first_factor(M);
            constexpr int Q = M / P;
            constexpr u64mask KNEED = needed_cols(OUT, P, Q);  // stage-2 transforms that are used
            constexpr u64mask SLIVE = live_subs(IN, P, Q);     // stage-1 transforms with nonzero input
            cplx<R> t[M];
            static_for<0, P>([&](auto s_) {
                constexpr int s = decltype(s_)::value;
                constexpr u64mask SIN = stride_mask(IN, s, P, Q);
                if constexpr (SIN != 0) {
                    cplx<R> u[Q];
#pragma unroll
                    for (int j = 0; j < Q; j++)
                        if (has_bit(SIN, j)) u[j] = v[s + P * j];
                    DftP<Q, DIR, R, SIN, KNEED>::run(u);
                    static_for<0, Q>([&](auto k_) {
                        constexpr int k = decltype(k_)::value;
                        if constexpr (has_bit(KNEED, k)) t[s * Q + k] = twiddle<M, s * k, DIR, R>(u[k]);
                    });
                }
            });
            static_for<0, Q>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                if constexpr (has_bit(KNEED, k)) {
                    constexpr u64mask QOUT = stride_mask(OUT, k, Q, P);
                    cplx<R> u[P];
#pragma unroll
                    for (int s = 0; s < P; s++)
                        if (has_bit(SLIVE, s)) u[s] = t[s * Q + k];
                    DftP<P, DIR, R, SLIVE, QOUT>::run(u);
#pragma unroll
                    for (int q = 0; q < P; q++)
                        if (has_bit(QOUT, q)) v[k + Q * q] = u[q];
                }
            });
        }
    }
};

// bit mask of the slots j < M (x = g + G j, every lane position g < G) whose
// x lies inside the box |x| <= KD of a length-N line
constexpr u64mask box_slots(int N, int G, int M, int KD) {
    u64mask r = 0;
    for (int j = 0; j < M; j++)
        for (int g = 0; g < G; g++) {
            const int x = g + G * j;
            if (x <= KD || x >= N - KD) r |= 1ull << j;
        }
    return r;
}

// Split of a line length N into A*B for the two-pass transform (A <= B, both
// register-sized); 0 if N is not served by the specialised kernels.
constexpr int split_a(int n) {
    return n == 24 ? 4 : n == 32 ? 4 : n == 36 ? 6 : n == 40 ? 5 : n == 48 ? 6 : n == 54 ? 6
         : n == 60 ? 6 : n == 64 ? 8 : n == 72 ? 8 : n == 80 ? 8 : n == 96 ? 8 : n == 100 ? 10
         : n == 108 ? 9 : n == 120 ? 10 : n == 128 ? 8 : n == 144 ? 12 : 0;
}

}  // namespace cpwgpu
