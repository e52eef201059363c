"""Pins the oracle's own FFT (the stand-in for FFTW, which is not vendored)
against numpy.fft and a brute-force O(n^2) DFT."""
import numpy as np
import pytest


def brute_dft(a, sign):
    n = a.size
    k = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return w @ a


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 30, 64, 96, 384, 512, 1024])
def test_fft1d_matches_numpy(O, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = np.abs(np.fft.fft(a)).max()
    assert np.abs(O.fft1d(a, -1) - np.fft.fft(a)).max() <= 1e-13 * scale
    assert np.abs(O.fft1d(a, +1) - np.fft.ifft(a) * n).max() <= 1e-13 * scale


@pytest.mark.parametrize("n", [5, 8, 12, 16, 27, 48])
def test_fft1d_matches_brute_force(O, n):
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, +1):
        ref = brute_dft(a, sign)
        assert np.abs(O.fft1d(a, sign) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(16, 16), (64, 48), (12, 20), (128, 64)])
def test_fft2d_matches_numpy_and_inverts(O, shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = O.fft2d(a, -1)
    assert np.abs(f - np.fft.fft2(a)).max() <= 1e-12 * np.abs(f).max()
    back = O.fft2d(f, +1) / a.size
    assert np.abs(back - a).max() <= 1e-13 * np.abs(a).max()


def test_fft2d_is_row_major_x_rows(O):
    # a plane wave along y (the contiguous axis) lands in row 0
    nx, ny = 8, 16
    a = np.exp(2j * np.pi * 3 * np.arange(ny) / ny)[None, :].repeat(nx, 0)
    f = O.fft2d(a, -1)
    assert abs(f[0, 3] - nx * ny) < 1e-9
    f[0, 3] = 0
    assert np.abs(f).max() < 1e-9
