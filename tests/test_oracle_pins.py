"""Further pins of the oracle (and of the constants the library shares with
it), each chosen so that a plausible slip fails:

- one slice step of orc_slice_step against a numpy restatement of
  adfstem.cpp:1866-2132 (IFFT, the transmission update of :1922-1952 in both
  modes, FFT, bw_limit.cpp:30-47, slice.cpp:1366-1421) with a random,
  non-uniform transmission function, random waves and random propagators on
  16^2 and 32^2 grids: a transposed, shifted or conjugated t, a dropped 1/s
  or a swapped real/imaginary update fails;
- a vacuum specimen through the whole raster (orc_adfstem): |psi_p(k)| =
  n_ap^-1/2 inside the aperture for every position (the probe shift and the
  propagators are pure phases), so V_omega = 0 on every populated bin and
  every ADF pixel equals (aperture pixels inside the detector ring) / n_ap,
  counted here independently;
- the Kirkland table and the Scherzer constants against values read from the
  reference sources (tests/golden/reference_constants.json, written by
  tests/golden/make_reference_fixtures.py), so they stay pinned where the
  reference is not mounted; the Scherzer helpers at inputs where each
  formula's power is exact, so a wrong exponent or a swapped argument fails.
"""
import json
import os
import re
import struct
import hashlib
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIX = json.load(open(os.path.join(HERE, "golden", "reference_constants.json")))
REF_SRC = "/root/reference/src"


def _step_numpy(psi_k, t, px, py, kx, ky, bwt, mode):
    """adfstem.cpp:1866-2132 for one slice, unitary FFTs as FFTW's
    unnormalised transforms divided by s = sqrt(NxNy)."""
    nx, ny = psi_k.shape
    s = np.sqrt(nx * ny)
    u = np.fft.ifft2(psi_k) * (nx * ny)             # IFFT_u (:1891)
    a = (t.real * u.real - t.imag * u.imag) / s     # Re written first (:1922-1952)
    other = a if mode == 0 else u.real
    b = (t.real * u.imag + t.imag * other) / s
    w = np.fft.fft2(a + 1j * b) / s                 # FFT_u / s (:1986-1992)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    inside = k2 < bwt * bwt                         # bw_limit.cpp:39, slice.cpp:1389
    return np.where(inside, w * (px[:, None] * py[None, :]), 0.0)


@pytest.mark.parametrize("n, L", [(16, 4.0), (32, 6.0)])
@pytest.mark.parametrize("mode", [0, 1])
def test_slice_step_against_numpy_restatement(O, n, L, mode):
    rng = np.random.default_rng(n + 7 * mode)
    kx, ky, _, _ = O.domains(n, n, L, L)
    _, bwt = O.bwcutoffs(n, n, L, L)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    assert np.abs(k2 - bwt * bwt).min() > 1e-9 * bwt * bwt     # no pixel on the band edge
    psi = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    t = (1.0 + 0.3 * rng.normal(size=(n, n))) * np.exp(1j * rng.uniform(-3, 3, size=(n, n)))
    px = np.exp(1j * rng.uniform(-3, 3, n)) * rng.uniform(0.5, 1.5, n)
    py = np.exp(1j * rng.uniform(-3, 3, n)) * rng.uniform(0.5, 1.5, n)
    got = O.slice_step(psi, t, px, py, kx, ky, bwt, transmit_mode=mode)
    want = _step_numpy(psi, t, px, py, kx, ky, bwt, mode)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-13, err
    # the same step with t transposed, or with the other transmission mode, is far off
    off = _step_numpy(psi, t.T, px, py, kx, ky, bwt, mode)
    assert np.abs(off - want).max() / np.abs(want).max() > 1e-2
    other = _step_numpy(psi, t, px, py, kx, ky, bwt, 1 - mode)
    assert np.abs(other - want).max() / np.abs(want).max() > 1e-3


def test_vacuum_raster_gives_flat_v_omega_and_counted_adf(O):
    n, L, v = 64, 8.0, 300e3
    alpha, inner, outer = 0.045, 0.02, 0.06
    job = O.make_job(nx=n, ny=n, xperiod=L, yperiod=L, zperiod=6.0, voltage=v, min_slice=1.0,
                     alpha_max=alpha, adf=1, det_inner=inner, det_outer=outer, fem=1, dr=0.5)
    spec = O.Specimen(job, np.zeros((0, 3)), np.zeros(0, np.int32))
    xp = np.array([0.3, 2.9, 5.55, 7.7])
    yp = np.array([1.1, 4.0, 6.6])
    img, vom, _ = O.adfstem(job, spec, xp, yp)
    kx, ky, _, _ = O.domains(n, n, L, L)
    lam = O.wavelength(v)[0]
    _, bwt = O.bwcutoffs(n, n, L, L)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    k1, k2r = inner / lam, outer / lam          # compared against k^2 (func.cpp:2633)
    for thr in (alpha * alpha, k1, k2r):        # no pixel on a threshold
        assert np.abs(lam * lam * k2 - thr if thr == alpha * alpha else k2 - thr).min() > 1e-9
    aperture = lam * lam * k2 < alpha * alpha
    assert aperture.sum() > 50 and (k2[aperture] < bwt * bwt).all()
    ring = (k2 > k1) & (k2 <= k2r)
    expect = np.sum(aperture & ring) / aperture.sum()
    assert 0 < expect < 1
    np.testing.assert_allclose(img, expect, rtol=1e-12)
    # V_omega: bins inside the aperture are the direct beam, zeroed
    # (adfstem.cpp:2518-2522, V = 0); the bin across the aperture edge holds
    # the same ring mean at every position, so V = 0 up to rounding; bins
    # outside hold only rounding noise of the transforms
    nb, dk = O.fem_nbins(kx, ky, 0.5, bwt)
    assert vom.size == nb
    k_ap = alpha / lam
    upper = (np.arange(nb) + 1) * dk
    direct = lam * upper <= alpha
    across = (~direct) & (np.arange(nb) * dk < k_ap)
    assert direct.sum() > 10 and across.sum() >= 1
    assert (vom[direct] == 0.0).all()
    assert np.abs(vom[across]).max() < 1e-12


def _inc_table():
    txt = open(os.path.join(ROOT, "ms-stem-fem_amd", "csrc", "kirkland_params.inc")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    rows = re.findall(r"\{([^{}]*)\}", txt)
    return [[float(x) for x in r.split(",") if x.strip()] for r in rows]


def test_parameter_table_matches_reference_values(O):
    tab = _inc_table()
    assert len(tab) == 103 and all(len(r) == 12 for r in tab)
    blob = b"".join(struct.pack("<12d", *r) for r in tab)
    assert hashlib.sha256(blob).hexdigest() == FIX["kirkland_sha256_f64le"]
    for z in range(1, 104):
        for k, q2 in enumerate(FIX["fe_q2"]):
            assert O.fe(z, q2) == pytest.approx(FIX["fe"][z - 1][k], rel=1e-14), (z, q2)


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not mounted")
def test_fixture_reads_the_reference_sources():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_reference_fixtures as M
    tab = M.kirkland_table()
    blob = b"".join(struct.pack("<12d", *r) for r in tab)
    assert hashlib.sha256(blob).hexdigest() == FIX["kirkland_sha256_f64le"]
    assert M.scherzer_constants() == FIX["scherzer"]


def test_scherzer_helpers_against_reference_constants():
    import msf
    sc = FIX["scherzer"]
    # uncorrected (scherzer.cpp:61-79): 0.87 sqrt(Cs3 lambda), 1.34 (lambda/Cs3)^(1/4)
    assert msf.scherzer_defocus(1.0, 4.0) == pytest.approx(
        2 * sc["scherzer_defocus_adfstem_uncorrected"], rel=1e-15)
    assert msf.scherzer_alphamax(16.0, 1.0) == pytest.approx(
        2 * sc["scherzer_alphamax_adfstem_uncorrected"], rel=1e-15)
    # corrected to Cs5 (scherzer.cpp:131-168): lambda^2 Cs5 = 8 (lambda Cs5^2 = 64)
    assert msf.scherzer_defocus(1.0, 8.0, corrected=1) == pytest.approx(
        2 * sc["scherzer_defocus_adfstem_correctedtoCs5"], rel=1e-15)
    # lambda / Cs5 = 64: the sixth root is 2 (the fourth root would be 2.83)
    assert msf.scherzer_alphamax(64.0, 1.0, corrected=1) == pytest.approx(
        2 * sc["scherzer_alphamax_adfstem_correctedtoCs5"], rel=1e-15)
    # lambda Cs5^2 = 27 (lambda^2 Cs5 = 729)
    assert msf.scherzer_cs3(27.0, 1.0) == pytest.approx(
        3 * sc["scherzer_Cs3_adfstem_correctedtoCs5"], rel=1e-15)
