"""Pins the oracle's scalar physics, grids, scattering factors, templates,
slicer and transmission functions with closed forms, invariants and an
independent (convolution-theorem) route."""
import os
import re

import numpy as np
import pytest

REF_HEADER = "/root/reference/src/scatterer_param_LUT.hpp"
M_E0, CC, EE = 9.109e-31, 299792458.0, 1.60217662e-19


def test_wavelength_known_values(O):
    # Kirkland (2010) Table 2.2: 0.02508 A at 200 kV, 0.01969 A at 300 kV
    assert O.wavelength(200e3)[0] == pytest.approx(0.02508, abs=5e-6)
    assert O.wavelength(300e3)[0] == pytest.approx(0.01969, abs=5e-6)
    assert O.wavelength(100e3)[0] == pytest.approx(0.03701, abs=1e-5)


@pytest.mark.parametrize("v", [80e3, 200e3, 300e3])
def test_gamma_is_relativistic_mass_ratio(O, v):
    # gamma = (h/lambda)^2 (m0c^2+eV)/((2m0c^2+eV) eV m0) reduces to 1 + eV/(m0 c^2)
    assert O.wavelength(v)[1] == pytest.approx(1.0 + EE * v / (M_E0 * CC * CC), rel=1e-12)


def test_bandwidth_cutoffs(O):
    pap, t = O.bwcutoffs(512, 512, 40.0, 40.0)
    assert pap == 512 / 80.0 and t == pytest.approx(2.0 / 3.0 * 6.4, rel=1e-15)
    pap, _ = O.bwcutoffs(256, 128, 20.0, 20.0)
    assert pap == 128 / 40.0  # the smaller of the two axes


def test_domains_fft_order(O):
    kx, ky, xx, yy = O.domains(8, 6, 4.0, 3.0, 0.5, -1.0)
    assert np.allclose(kx, np.fft.fftfreq(8, d=4.0 / 8), atol=1e-15)
    assert np.allclose(ky, np.fft.fftfreq(6, d=3.0 / 6), atol=1e-15)
    assert np.allclose(xx, 0.5 + np.arange(8) * 0.5)
    assert np.allclose(yy, -1.0 + np.arange(6) * 0.5)


def test_scherzer(O):
    lam = O.wavelength(200e3)[0]
    assert O.scherzer_defocus_uncorrected(lam, 1.2e7) == pytest.approx(477.3, abs=0.1)
    assert O.scherzer_alphamax_uncorrected(lam, 1.2e7) == pytest.approx(
        1.34 * (lam / 1.2e7) ** 0.25, rel=1e-14)


def _params_from_reference():
    txt = open(REF_HEADER).read()
    pat = re.compile(r"\(get_paramPtr\(\s*(\d+)\s*\)\)->([abcd])\[([0-2])\]\s*=\s*([-+0-9.eE]+)\s*;")
    tab = {}
    for m in pat.finditer(txt):
        tab[(int(m.group(1)), m.group(2), int(m.group(3)))] = float(m.group(4))
    return tab


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference sources not mounted")
def test_parameter_table_matches_reference_source(O):
    tab = _params_from_reference()
    assert len(tab) == 103 * 12
    for z in range(1, 104):
        a = [tab[(z, "a", i)] for i in range(3)]
        b = [tab[(z, "b", i)] for i in range(3)]
        c = [tab[(z, "c", i)] for i in range(3)]
        d = [tab[(z, "d", i)] for i in range(3)]
        for q2 in (0.0, 0.37, 4.0):
            ref = sum(ai / (q2 + bi) for ai, bi in zip(a, b)) + sum(
                ci * np.exp(-di * q2) for ci, di in zip(c, d))
            assert O.fe(z, q2) == pytest.approx(ref, rel=1e-14), (z, q2)


def test_fe_limits(O):
    for z in (1, 14, 29, 40, 79):
        f0 = O.fe(z, 0.0)
        assert f0 > 0.0
        # large q: the Gaussians vanish, the Lorentzians fall as sum(a)/q^2
        q2 = 1e6
        assert O.fe(z, q2) * q2 == pytest.approx(O.fe(z, 1e7) * 1e7, rel=1e-4)
        assert O.fe(z, 1.0) < f0
    assert np.isnan(O.fe(0, 0.0)) and np.isnan(O.fe(104, 0.0))


def _template_setup(O, n=64, L=12.0, v=300e3):
    kx, ky, xx, yy = O.domains(n, n, L, L)
    lam, gam = O.wavelength(v)
    pap, bwt = O.bwcutoffs(n, n, L, L)
    return kx, ky, xx, yy, lam, gam, pap, bwt


def test_template_is_real_symmetric_and_matches_k_space(O):
    n, L = 64, 12.0
    kx, ky, xx, yy, lam, gam, pap, bwt = _template_setup(O, n, L)
    t = O.template(14, lam, gam, 1 / (L * L), pap, kx, ky)
    assert np.abs(t.imag).max() <= 1e-12 * np.abs(t.real).max()
    r = t.real
    assert np.allclose(r, r.T, rtol=0, atol=1e-12 * np.abs(r).max())
    assert np.allclose(r, np.roll(r[::-1, :], 1, axis=0), rtol=0, atol=1e-12 * np.abs(r).max())
    # FFT(template) * sqrt(M) / M must give lambda gamma/(ab) f_e(k^2) on the support
    M = n * n
    tk = np.fft.fft2(t) / np.sqrt(M)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    fe = np.vectorize(lambda q: O.fe(14, q))(k2)
    expect = np.where(k2 < pap * pap, lam * gam / (L * L) * fe, 0.0)
    assert np.abs(tk - expect).max() <= 1e-12 * np.abs(expect).max()


def test_template_off_origin_is_shifted(O):
    n, L = 32, 8.0
    kx, ky, xx, yy, lam, gam, pap, bwt = _template_setup(O, n, L)
    t0 = O.template(29, lam, gam, 1 / (L * L), pap, kx, ky)
    t1 = O.template(29, lam, gam, 1 / (L * L), pap, kx, ky, xmin=3 * L / n, ymin=-2 * L / n)
    # e^{+2 pi i k.x0} in k-space is a shift by -x0 in real space; the phase uses
    # the reference's PI = 3.14159265369, off by 1e-10, hence the looser bound
    assert np.abs(t1 - np.roll(np.roll(t0, -3, 0), 2, 1)).max() <= 1e-9 * np.abs(t0).max()


@pytest.mark.parametrize("z,minth,lb,th", [
    ([0.0, 0.5, 1.0, 1.01, 2.5], 1.0, [0.5, 1.01, 2.5], [0.51, 1.49, 1.0]),
    ([2.5, 1.0, 0.0, 1.01, 0.5], 1.0, [0.5, 1.01, 2.5], [0.51, 1.49, 1.0]),
    ([3.0], 2.0, [3.0], [2.0]),
    ([1.0, 1.0, 1.0, 4.0, 4.0], 0.5, [1.0, 4.0], [3.0, 0.5]),
])
def test_autoslice_hand_cases(O, z, minth, lb, th):
    q = np.zeros((len(z), 3))
    q[:, 2] = z
    sl, lower, thick = O.autoslice(q, 0.0, 10.0, minth)
    assert np.allclose(lower, lb, rtol=0, atol=1e-15)
    assert np.allclose(thick, th, rtol=0, atol=1e-15)
    # every slice's atoms lie within min_thickness of its first atom
    for s in range(len(lb)):
        zs = np.asarray(z)[sl == s]
        assert zs.max() - zs.min() <= minth


def test_autoslice_empty_specimen(O):
    sl, lower, thick = O.autoslice(np.zeros((0, 3)), 1.5, 20.0, 2.0)
    assert list(lower) == [1.5] and list(thick) == [20.0]


def test_shift_rule_edges(O):
    xx = np.arange(8) * 0.5  # dx = 0.5, half = 0.25
    assert O.shift_index(xx, 0.0) == 0          # first qualifying i = 1
    assert O.shift_index(xx, 0.2499) == 0
    assert O.shift_index(xx, 0.25) == 0         # 0.5 - 0.25 >= 0.25 at i = 1: half rounds down
    assert O.shift_index(xx, 0.2501) == 1
    assert O.shift_index(xx, 1.3) == 3          # nearest pixel 1.5
    assert O.shift_index(xx, 3.2, previous=5) == 6
    assert O.shift_index(xx, 3.3, previous=5) == 5   # beyond x[N-1] - dx/2: keeps previous
    assert O.shift_index(xx, -0.3) == -1             # below the domain


def _numpy_transmission(O, q, Z, zlist, kx, ky, xx, yy, lam, gam, L, pap, bwt):
    """Convolution-theorem route: sigma_v = IFFT(sum_sp T_sp * FFT(count_sp)) / sqrt(M)."""
    n = kx.size
    M = n * n
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    sig_k = np.zeros((n, n), complex)
    for z in zlist:
        cnt = np.zeros((n, n))
        for a in np.nonzero(Z == z)[0]:
            sx = O.shift_index(xx, q[a, 0])
            sy = O.shift_index(yy, q[a, 1])
            cnt[sx, sy] += 1
        fe = np.vectorize(lambda v: O.fe(int(z), v))(k2)
        tk = np.where(k2 < pap * pap, lam * gam / (L * L) * fe, 0.0)
        sig_k += tk * np.fft.fft2(cnt)
    sv = np.fft.ifft2(sig_k) * M / np.sqrt(M)
    t = np.exp(-sv.imag) * (np.cos(sv.real) + 1j * np.sin(sv.real))
    tk = np.fft.fft2(t)
    tk[k2 >= bwt * bwt] = 0
    return np.fft.ifft2(tk)


def test_transmission_equals_convolution_theorem(O):
    n, L = 64, 12.0
    kx, ky, xx, yy, lam, gam, pap, bwt = _template_setup(O, n, L)
    rng = np.random.default_rng(5)
    q = rng.random((25, 3)) * L
    Z = rng.choice(np.array([14, 29], np.int32), 25)
    zlist = np.array([14, 29], np.int32)
    tm = np.stack([O.template(z, lam, gam, 1 / (L * L), pap, kx, ky) for z in zlist])
    atoms = np.arange(25, dtype=np.int32)
    t = O.transmission(q, Z, atoms, zlist, tm, kx, ky, xx, yy, bwt)
    ref = _numpy_transmission(O, q, Z, zlist, kx, ky, xx, yy, lam, gam, L, pap, bwt)
    assert np.abs(t - ref).max() <= 1e-12
    # vacuum slice: t == 1
    t0 = O.transmission(q, Z, np.zeros(0, np.int32), zlist, tm, kx, ky, xx, yy, bwt)
    assert np.abs(t0 - 1.0).max() <= 1e-14


def test_transmission_is_permutation_invariant(O):
    n, L = 32, 8.0
    kx, ky, xx, yy, lam, gam, pap, bwt = _template_setup(O, n, L)
    rng = np.random.default_rng(9)
    q = rng.random((10, 3)) * L
    Z = np.full(10, 14, np.int32)
    tm = O.template(14, lam, gam, 1 / (L * L), pap, kx, ky)[None]
    zl = np.array([14], np.int32)
    a = O.transmission(q, Z, np.arange(10, dtype=np.int32), zl, tm, kx, ky, xx, yy, bwt)
    b = O.transmission(q, Z, rng.permutation(10).astype(np.int32), zl, tm, kx, ky, xx, yy, bwt)
    assert np.abs(a - b).max() <= 1e-14
