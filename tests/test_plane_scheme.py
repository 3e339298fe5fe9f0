"""The in-place four-step scheme of the whole-plane convolution kernel
(conv_plane_kernel in xatu_amd/csrc/device/convolution.hip), restated stage by stage with
the same index maps and checked against a 2-D FFT convolution. The kernel computes
G <- IDFT2(DFT2(G) * P) (both unnormalised) per n x n plane, P indexed [xi1][xi0]; its GPU
parity is covered by the matrix-free tests against the oracle's matrix."""
import numpy as np
import pytest

LENGTHS = [1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16]


def factors(n):
    """(A, B): A >= B, both DFT lengths of the kernel, A B = n, B as large as possible."""
    A = B = 0
    for b in LENGTHS:
        for a in LENGTHS:
            if a >= b and a * b == n and b > B:
                A, B = a, b
    return A, B


def dft(v, sign):
    j = np.arange(len(v))
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / len(v)) @ v


def plane_scheme(x, P):
    n = x.shape[0]
    A, B = factors(n)
    w = np.exp(-2j * np.pi * np.arange(n) / n)
    pl = np.zeros((n, n + 1), complex)

    def stage_a(get, put, s, forward):
        v = dft(np.array([get(B * j + s) for j in range(A)]), -1 if forward else 1)
        for k in range(A):
            put(B * k + s, v[k] * (w[s * k] if forward and k else 1))

    def stage_b(get, put, t, mode, spec=None):
        v = np.array([get(B * t + j) for j in range(B)])
        if mode >= 0:
            v = dft(v, -1)
        if mode == 1:
            v = v * np.array([spec(t + A * k) for k in range(B)])
        if mode != 0:
            v = dft(v, 1)
        for j in range(B):
            put(B * t + j, v[j] * (np.conj(w[j * t]) if mode != 0 and j * t else 1))

    def row(o1):
        return (lambda p: pl[o1, p]), (lambda p, val: pl.__setitem__((o1, p), val))

    def col(p):
        return (lambda r: pl[r, p]), (lambda r, val: pl.__setitem__((r, p), val))

    for o1 in range(n):  # rows, forward A-point stage from G
        for s in range(B):
            stage_a(lambda q: x[o1, q], row(o1)[1], s, True)
    for o1 in range(n):  # rows, forward B-point stage
        for t in range(A):
            stage_b(*row(o1), t, 0)
    for p in range(n):  # columns, forward A-point stage
        for s in range(B):
            stage_a(*col(p), s, True)
    for p in range(n):  # columns: forward B-point, spectrum, inverse
        xi0 = p // B + A * (p % B)
        for t in range(A):
            stage_b(*col(p), t, 1, spec=lambda xi1: P[xi1, xi0])
    for p in range(n):  # columns, inverse A-point stage
        for s in range(B):
            stage_a(*col(p), s, False)
    for o1 in range(n):  # rows, inverse B-point stage
        for t in range(A):
            stage_b(*row(o1), t, -1)
    out = np.zeros((n, n), complex)
    for o1 in range(n):  # rows, inverse A-point stage to G
        for s in range(B):
            stage_a(row(o1)[0], lambda q, val: out.__setitem__((o1, q), val), s, False)
    return out


@pytest.mark.parametrize("n", [2, 3, 6, 8, 9, 12, 16, 20, 24])
def test_plane_scheme_equals_fft_convolution(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    P = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    ref = np.fft.ifft2(np.fft.fft2(x) * P) * n * n
    got = plane_scheme(x, P)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_plane_scheme_catches_a_wrong_spectrum_order():
    """Reading the spectrum in natural rather than the permuted row-frequency order breaks it."""
    n = 12
    rng = np.random.default_rng(1)
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    P = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A, B = factors(n)
    perm = [p // B + A * (p % B) for p in range(n)]
    ref = np.fft.ifft2(np.fft.fft2(x) * P) * n * n
    # feeding the scheme a spectrum whose columns are already permuted is equivalent to the
    # kernel reading column p instead of column xi0(p)
    wrong = plane_scheme(x, P[:, np.argsort(perm)])
    assert np.abs(wrong - ref).max() > 1e-3 * np.abs(ref).max()
