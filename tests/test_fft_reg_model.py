"""Lane-level model of the register-resident line transforms (csrc/fft_reg.hpp)
checked against numpy's FFT: G lanes of M values, the lane -> position map,
the DPP partner patterns (quad xor 1, quad xor 2, row_half_mirror), the
butterfly signs and stage twiddles, and the DIF/DIT layouts
    DIF  in : lane g holds f[g + G j]           out: F[k1 + M bitrev(g)]
    DIT  in : lane g holds f[k1 + M bitrev(g)]  out: F[g + G m]
for every (N, G) the plane kernels are built with.  A wrong partner lane, sign
or twiddle index fails these checks."""

import numpy as np
import pytest


def pos_of_lane(l, G):
    return 11 - l if (G == 8 and l >= 4) else l


def partner_lane(l, h):
    # dpp_ctrl<H>: 1 -> quad_perm [1,0,3,2], 2 -> quad_perm [2,3,0,1], 4 -> row_half_mirror
    if h == 1:
        return l ^ 1
    if h == 2:
        return l ^ 2
    return 7 - l


def bitrev(v, bits):
    return int(format(v, f"0{bits}b")[::-1], 2) if bits else 0


def stage_tw(g, h, direction):
    # W_{2h}^{(g mod 2h) - h} for the upper half (backward sign), 1 otherwise
    p = g % (2 * h)
    if p < h or h == 1:
        return 1.0 + 0j
    return np.exp(direction * 2j * np.pi * (p - h) / (2 * h))


def dft(v, direction):
    m = len(v)
    k = np.arange(m)
    return np.exp(direction * 2j * np.pi * np.outer(k, k) / m) @ v


def bfly(vals, G, h):
    """vals[l] (lanes) -> partner + s * mine per lane."""
    out = np.empty_like(vals)
    for l in range(G):
        g = pos_of_lane(l, G)
        s = -1.0 if g & h else 1.0
        out[l] = vals[partner_lane(l, h)] + s * vals[l]
    return out


def line_dif(v, G, N, direction):
    M = N // G
    v = np.array([dft(v[l], direction) for l in range(G)])
    tw = np.exp(2j * np.pi * np.arange(N) / N)
    for l in range(G):
        g = pos_of_lane(l, G)
        w = tw[g * np.arange(M)]
        v[l] = v[l] * (w if direction > 0 else np.conj(w))
    for h in (4, 2, 1):
        if h >= G:
            continue
        v = bfly(v, G, h)
        for l in range(G):
            v[l] *= stage_tw(pos_of_lane(l, G), h, direction)
    return v


def line_dit(v, G, N, direction):
    M = N // G
    v = v.copy()
    for h in (1, 2, 4):
        if h >= G:
            continue
        for l in range(G):
            v[l] *= stage_tw(pos_of_lane(l, G), h, direction)
        v = bfly(v, G, h)
    tw = np.exp(2j * np.pi * np.arange(N) / N)
    for l in range(G):
        g = pos_of_lane(l, G)
        w = tw[g * np.arange(M)]
        v[l] = v[l] * (w if direction > 0 else np.conj(w))
    return np.array([dft(v[l], direction) for l in range(G)])


def sgn(g, h):
    return -1.0 if g & h else 1.0


def bfly_t(vals, G, h, t):
    """One-instruction FP32 butterfly (fmac_xlane): mine + t[lane] * partner."""
    out = np.empty_like(vals)
    for l in range(G):
        out[l] = vals[l] + t[l] * vals[partner_lane(l, h)]
    return out


def lane_consts(G, fn):
    return [fn(pos_of_lane(l, G)) for l in range(G)]


def line_dif_t(v, G, N, direction, negout):
    """line_dif_tw with bfly_t (FP32, G = 4 or 8): the upper-lane sign left by
    each butterfly is folded into the stage twiddle after it (Lanes::w4s =
    s4 w4, w2s = s2 w2); with negout the last stage leaves the odd positions
    negated."""
    M = N // G
    v = np.array([dft(v[l], direction) for l in range(G)])
    tw = np.exp(2j * np.pi * np.arange(N) / N)
    for l in range(G):
        w = tw[pos_of_lane(l, G) * np.arange(M)]
        v[l] = v[l] * (w if direction > 0 else np.conj(w))
    for h in (4, 2):
        if h >= G:
            continue
        v = bfly_t(v, G, h, lane_consts(G, lambda g: sgn(g, h)))
        for l in range(G):
            g = pos_of_lane(l, G)
            v[l] *= sgn(g, h) * stage_tw(g, h, direction)
    return bfly_t(v, G, 1, lane_consts(G, lambda g: sgn(g, 1))) if negout else bfly(v, G, 1)


def line_dit_t(v, G, N, direction, negin):
    """line_dit_tw with bfly_t (FP32, G = 4 or 8): first stage t = -s1 on
    negated odd inputs (exact outputs) or s1 (odd outputs negated, restored by
    w2d = s1 w2); the second stage's sign goes into w4d = s2 w4; the last stage
    is the exact bfly."""
    M = N // G
    v = bfly_t(v.copy(), G, 1, lane_consts(G, lambda g: -sgn(g, 1) if negin else sgn(g, 1)))
    for l in range(G):
        g = pos_of_lane(l, G)
        v[l] *= (1.0 if negin else sgn(g, 1)) * stage_tw(g, 2, direction)
    if G == 8:
        v = bfly_t(v, G, 2, lane_consts(G, lambda g: sgn(g, 2)))
        for l in range(G):
            g = pos_of_lane(l, G)
            v[l] *= sgn(g, 2) * stage_tw(g, 4, direction)
        v = bfly(v, G, 4)
    else:
        v = bfly(v, G, 2)
    tw = np.exp(2j * np.pi * np.arange(N) / N)
    for l in range(G):
        w = tw[pos_of_lane(l, G) * np.arange(M)]
        v[l] = v[l] * (w if direction > 0 else np.conj(w))
    return np.array([dft(v[l], direction) for l in range(G)])


SHAPES = [(n, g) for n in (24, 48, 64, 72, 80, 96, 108, 128) for g in (1, 2, 4, 8) if n % g == 0]


@pytest.mark.parametrize("N,G", SHAPES)
def test_dif_then_dit_layouts(N, G):
    rng = np.random.default_rng(N * 10 + G)
    f = rng.normal(size=N) + 1j * rng.normal(size=N)
    M, bits = N // G, G.bit_length() - 1
    # DIF, backward sign: lane at position g holds f[g + G j]
    lanes = np.empty((G, M), complex)
    for l in range(G):
        g = pos_of_lane(l, G)
        lanes[l] = f[g + G * np.arange(M)]
    out = line_dif(lanes, G, N, +1)
    F = np.fft.ifft(f) * N  # sum_x f[x] exp(+2 pi i x k / N)
    for l in range(G):
        g = pos_of_lane(l, G)
        np.testing.assert_allclose(out[l], F[np.arange(M) + M * bitrev(g, bits)], atol=1e-9)
    # DIT, forward sign, straight from the DIF output layout
    back = line_dit(out, G, N, -1)
    for l in range(G):
        g = pos_of_lane(l, G)
        np.testing.assert_allclose(back[l], N * f[g + G * np.arange(M)], atol=1e-9)


@pytest.mark.parametrize("N,G", [(72, 8), (72, 4), (72, 2), (80, 8)])
def test_dit_forward_transform_alone(N, G):
    rng = np.random.default_rng(7)
    f = rng.normal(size=N) + 1j * rng.normal(size=N)
    M, bits = N // G, G.bit_length() - 1
    lanes = np.empty((G, M), complex)
    for l in range(G):
        lanes[l] = f[np.arange(M) + M * bitrev(pos_of_lane(l, G), bits)]
    out = line_dit(lanes, G, N, -1)
    F = np.fft.fft(f)
    for l in range(G):
        g = pos_of_lane(l, G)
        np.testing.assert_allclose(out[l], F[g + G * np.arange(M)], atol=1e-9)


def test_half_mirror_pairs_are_position_xor4():
    for l in range(8):
        assert pos_of_lane(partner_lane(l, 4), 8) == pos_of_lane(l, 8) ^ 4
        assert pos_of_lane(partner_lane(l, 2), 8) == pos_of_lane(l, 8) ^ 2
        assert pos_of_lane(partner_lane(l, 1), 8) == pos_of_lane(l, 8) ^ 1


@pytest.mark.parametrize("N,G", [(n, g) for n in (24, 48, 64, 72, 80, 96, 108, 128) for g in (4, 8)
                                 if n % g == 0])
def test_one_instruction_butterflies(N, G):
    """The FP32 form (bfly_t with sign-carrying twiddles): exact DIF, DIF with
    negated odd positions into DIT with NEGIN (the x rows, V(r) in between),
    and DIT alone on exact inputs, all against numpy's FFT."""
    rng = np.random.default_rng(N * 7 + G)
    f = rng.normal(size=N) + 1j * rng.normal(size=N)
    vr = rng.normal(size=N)
    M, bits = N // G, G.bit_length() - 1
    lanes = np.empty((G, M), complex)
    for l in range(G):
        lanes[l] = f[pos_of_lane(l, G) + G * np.arange(M)]
    F = np.fft.ifft(f) * N
    exact = line_dif_t(lanes, G, N, +1, negout=False)
    neg = line_dif_t(lanes, G, N, +1, negout=True)
    for l in range(G):
        g = pos_of_lane(l, G)
        want = F[np.arange(M) + M * bitrev(g, bits)]
        np.testing.assert_allclose(exact[l], want, atol=1e-9)
        np.testing.assert_allclose(neg[l], sgn(g, 1) * want, atol=1e-9)
    # V(r) on the lane layout (M consecutive points per lane), then forward
    for l in range(G):
        neg[l] *= vr[np.arange(M) + M * bitrev(pos_of_lane(l, G), bits)]
    back = line_dit_t(neg, G, N, -1, negin=True)
    g_r = np.fft.fft(F * vr)
    for l in range(G):
        np.testing.assert_allclose(back[l], g_r[pos_of_lane(l, G) + G * np.arange(M)], atol=1e-8)
    # DIT alone on exact inputs (the forward y and z lines)
    ins = np.empty((G, M), complex)
    for l in range(G):
        ins[l] = f[np.arange(M) + M * bitrev(pos_of_lane(l, G), bits)]
    out = line_dit_t(ins, G, N, -1, negin=False)
    Ff = np.fft.fft(f)
    for l in range(G):
        np.testing.assert_allclose(out[l], Ff[pos_of_lane(l, G) + G * np.arange(M)], atol=1e-9)


def test_one_instruction_butterfly_catches_a_wrong_sign():
    """Dropping the sign carried into w4s (using w4) must fail the check above."""
    N, G = 72, 8
    rng = np.random.default_rng(3)
    f = rng.normal(size=N) + 1j * rng.normal(size=N)
    M, bits = N // G, 3
    lanes = np.array([f[pos_of_lane(l, G) + G * np.arange(M)] for l in range(G)])
    v = np.array([dft(lanes[l], +1) for l in range(G)])
    tw = np.exp(2j * np.pi * np.arange(N) / N)
    for l in range(G):
        v[l] = v[l] * tw[pos_of_lane(l, G) * np.arange(M)]
    v = bfly_t(v, G, 4, lane_consts(G, lambda g: sgn(g, 4)))
    for l in range(G):
        v[l] *= stage_tw(pos_of_lane(l, G), 4, +1)  # sign not folded in
    v = bfly(v, G, 2)
    for l in range(G):
        v[l] *= stage_tw(pos_of_lane(l, G), 2, +1)
    v = bfly(v, G, 1)
    F = np.fft.ifft(f) * N
    err = max(np.abs(v[l] - F[np.arange(M) + M * bitrev(pos_of_lane(l, G), bits)]).max() for l in range(G))
    assert err > 1e-3
