// CPU check of cpw2000_amd/csrc/fft_core.hpp: register DFTs and the two-pass
// line transforms (direct and transposed split) against a long-double DFT.
// Prints "max_rel_err <value>" and exits nonzero above tolerance.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../cpw2000_amd/csrc/fft_core.hpp"

using namespace cpwgpu;
typedef cplx<double> C;

static std::vector<C> naive(const std::vector<C> &x, int dir) {
    const int n = (int)x.size();
    std::vector<C> y(n);
    for (int m = 0; m < n; m++) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            long double a = dir * 2.0L * 3.141592653589793238462643383279502884L * (long double)((long)j * m % n) / n;
            sr += x[j].x * cosl(a) - x[j].y * sinl(a);
            si += x[j].x * sinl(a) + x[j].y * cosl(a);
        }
        y[m] = mk<double>((double)sr, (double)si);
    }
    return y;
}

static double relerr(const std::vector<C> &a, const std::vector<C> &b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += (a[i].x - b[i].x) * (a[i].x - b[i].x) + (a[i].y - b[i].y) * (a[i].y - b[i].y);
        den += b[i].x * b[i].x + b[i].y * b[i].y;
    }
    return std::sqrt(num / den);
}

static std::mt19937_64 rng(12345);
static std::vector<C> rnd(int n) {
    std::normal_distribution<double> g;
    std::vector<C> v(n);
    for (auto &c : v) c = mk<double>(g(rng), g(rng));
    return v;
}

template <int M, int DIR> double check_dft() {
    auto x = rnd(M);
    C v[M];
    for (int i = 0; i < M; i++) v[i] = x[i];
    dft<M, DIR, double>(v);
    std::vector<C> y(v, v + M);
    return relerr(y, naive(x, DIR));
}

// two-pass transform of one line, "threads" emulated by loops
template <int N, int DIR> double check_two_pass() {
    constexpr int A = split_a(N), B = N / A;
    auto x = rnd(N);
    std::vector<C> buf = x, tw(N);
    for (int j = 0; j < N; j++) {
        long double a = DIR * 2.0L * 3.141592653589793238462643383279502884L * j / N;
        tw[j] = mk<double>((double)cosl(a), (double)sinl(a));
    }
    for (int s = 0; s < A; s++) {  // pass 1
        C v[B];
        for (int j = 0; j < B; j++) v[j] = buf[s + A * j];
        dft<B, DIR, double>(v);
        for (int k = 0; k < B; k++) buf[s + A * k] = cmul(v[k], tw[s * k]);
    }
    std::vector<C> out(N);
    C hold[B][A];
    for (int k = 0; k < B; k++) {  // pass 2
        for (int s = 0; s < A; s++) hold[k][s] = buf[s + A * k];
        dft<A, DIR, double>(hold[k]);
        for (int q = 0; q < A; q++) out[k + B * q] = hold[k][q];
    }
    double e1 = relerr(out, naive(x, DIR));
    // transposed split from the pass-2 distribution: thread k holds x'[k + B q]
    auto x2 = rnd(N);
    for (int k = 0; k < B; k++)
        for (int q = 0; q < A; q++) hold[k][q] = x2[k + B * q];
    std::vector<C> buf2(N), out2(N);
    for (int k = 0; k < B; k++) {  // pass 1'
        dft<A, DIR, double>(hold[k]);
        for (int m = 0; m < A; m++) buf2[k + B * m] = cmul(hold[k][m], tw[k * m]);
    }
    for (int m = 0; m < A; m++) {  // pass 2'
        C v[B];
        for (int k = 0; k < B; k++) v[k] = buf2[k + B * m];
        dft<B, DIR, double>(v);
        for (int p = 0; p < B; p++) out2[m + A * p] = v[p];
    }
    double e2 = relerr(out2, naive(x2, DIR));
    return e1 > e2 ? e1 : e2;
}

// pruned transform: inputs outside IN hold NaN (must never be read), only the
// outputs in OUT are compared with the DFT of the input with zeros outside IN
template <int M, int DIR, u64mask IN, u64mask OUT> double check_pruned() {
    auto x = rnd(M);
    for (int j = 0; j < M; j++)
        if (!has_bit(IN, j)) x[j] = mk<double>(0.0, 0.0);
    C v[M];
    for (int j = 0; j < M; j++) v[j] = has_bit(IN, j) ? x[j] : mk<double>(NAN, NAN);
    DftP<M, DIR, double, IN, OUT>::run(v);
    auto y = naive(x, DIR);
    std::vector<C> got, want;
    for (int m = 0; m < M; m++)
        if (has_bit(OUT, m)) {
            got.push_back(v[m]);
            want.push_back(y[m]);
        }
    const double e = relerr(got, want);
    return e == e ? e : 1.0;  // NaN read -> failure
}

// the masks of the register line transforms: slots of a length-N line on G
// lanes whose x can fall in the box |x| <= N/4
template <int N, int G, int DIR> double check_line_masks() {
    constexpr int M = N / G;
    constexpr u64mask B = box_slots(N, G, M, N / 4);
    const double a = check_pruned<M, DIR, B, full_mask(M)>();
    const double b = check_pruned<M, DIR, full_mask(M), B>();
    const double c = check_pruned<M, DIR, B, B>();
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

int main() {
    double e = 0;
    auto up = [&](double v, const char *what, int m) {
        if (v > 1e-14) printf("FAIL %s %d err %.3e\n", what, m, v);
        if (v > e) e = v;
    };
#define D(M) up(check_dft<M, -1>(), "dft", M); up(check_dft<M, 1>(), "dft", M);
    D(2) D(3) D(4) D(5) D(6) D(8) D(9) D(10) D(12) D(15) D(16) D(18) D(20)
#define T(N) up(check_two_pass<N, -1>(), "two_pass", N); up(check_two_pass<N, 1>(), "two_pass", N);
    T(24) T(32) T(36) T(40) T(48) T(54) T(60) T(64) T(72) T(80) T(96) T(100) T(108) T(120) T(128) T(144)
#define PM(N, G) up(check_line_masks<N, G, -1>(), "pruned", N * 100 + G); \
    up(check_line_masks<N, G, 1>(), "pruned", N * 100 + G);
    PM(24, 2) PM(24, 4) PM(48, 2) PM(48, 4) PM(48, 8) PM(64, 2) PM(64, 4) PM(64, 8) PM(72, 2) PM(72, 4)
    PM(72, 8) PM(80, 2) PM(80, 4) PM(80, 8) PM(96, 2) PM(96, 4) PM(96, 8) PM(108, 2) PM(108, 4) PM(128, 2)
    PM(128, 4) PM(128, 8)
    // irregular masks
    up(check_pruned<9, 1, 0x1C7ull, 0x1FFull>(), "pruned", 9);
    up(check_pruned<12, -1, 0x00Full, 0xF0Full>(), "pruned", 12);
    up(check_pruned<36, 1, 0xE000003FFull, 0x3Full>(), "pruned", 36);
    up(check_pruned<20, -1, 0x1ull, 0xFFFFFull>(), "pruned", 20);
    printf("max_rel_err %.3e\n", e);
    return e < 1e-14 ? 0 : 1;
}
