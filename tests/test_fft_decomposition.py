"""The index algebra of the line-parallel FFT kernels, checked on the CPU
against numpy.fft: a line of N points held by L lanes of R = N / L points
(lane n1 holds x[n1 + L r]), an R-point DFT per lane, the twiddles
W_N^(n1 k1), one exchange (lane n1 writes its k1-th value to slot
k1 L + n1, then takes slots (n1 + L j) L + n for j < R / L, n < L) and
L-point DFTs; element o of a lane is X[n1 + L (o / L) + R (o % L)]
(voce_kernels.hip: line_dft, line_out_index, P2Plan).  Also the 12-point
butterfly's split n = n1 + 4 n2, k = 3 k1 + k2 (radix 3, twiddle
W12^(n1 k2), radix 4)."""
import numpy as np
import pytest

# (N, L) of every line plan: plane kernels 16, 32, 48, 64, 128; split passes 256
PLANS = [(16, 4), (32, 4), (48, 4), (64, 8), (128, 8), (256, 16)]


def line_dft(x, L, inverse=False):
    N = len(x)
    R = N // L
    assert R % L == 0
    sgn = 1.0 if inverse else -1.0
    dft = np.fft.ifft if inverse else np.fft.fft
    scale = (lambda a: a * len(a)) if inverse else (lambda a: a)
    v = [scale(dft(np.array([x[n1 + L * r] for r in range(R)]))) for n1 in range(L)]
    for n1 in range(L):
        v[n1] = v[n1] * np.exp(sgn * 2j * np.pi * n1 * np.arange(R) / N)
    slots = np.zeros(N, complex)
    for n1 in range(L):
        for k1 in range(R):
            slots[k1 * L + n1] = v[n1][k1]
    out = np.zeros(N, complex)
    for n1 in range(L):
        w = np.zeros(R, complex)
        for j in range(R // L):
            w[j * L:(j + 1) * L] = scale(dft(np.array([slots[(n1 + L * j) * L + n] for n in range(L)])))
        for o in range(R):
            out[n1 + L * (o // L) + R * (o % L)] = w[o]
    return out


@pytest.mark.parametrize("N,L", PLANS)
@pytest.mark.parametrize("inverse", [False, True])
def test_line_dft_matches_numpy(N, L, inverse):
    rng = np.random.default_rng(N + L)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ref = np.fft.ifft(x) * N if inverse else np.fft.fft(x)
    got = line_dft(x, L, inverse)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_line_out_index_is_a_permutation():
    for N, L in PLANS:
        R = N // L
        idx = sorted(n1 + L * (o // L) + R * (o % L) for n1 in range(L) for o in range(R))
        assert idx == list(range(N))


@pytest.mark.parametrize("inverse", [False, True])
def test_12_point_butterfly_split(inverse):
    rng = np.random.default_rng(12)
    v = rng.standard_normal(12) + 1j * rng.standard_normal(12)
    sgn = 1.0 if inverse else -1.0
    W = lambda n, k: np.exp(sgn * 2j * np.pi * k / n)
    y = np.zeros((3, 4), complex)
    for n1 in range(4):
        t = [v[n1], v[n1 + 4], v[n1 + 8]]
        for k2 in range(3):
            y[k2, n1] = sum(t[n2] * W(3, n2 * k2) for n2 in range(3)) * W(12, n1 * k2)
    out = np.zeros(12, complex)
    for k2 in range(3):
        for k1 in range(4):
            out[3 * k1 + k2] = sum(y[k2, n1] * W(4, n1 * k1) for n1 in range(4))
    ref = np.fft.ifft(v) * 12 if inverse else np.fft.fft(v)
    assert np.abs(out - ref).max() < 1e-13
