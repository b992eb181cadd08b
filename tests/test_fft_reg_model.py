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
