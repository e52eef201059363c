"""The oracle's detector MTF (adfstem.cpp:300-439, :2272-2385; tem.cpp:276-285)
against what fixes it without the reference: a flat MTF with a vanishing sinc
argument is the identity (the 2-D MTF is 1 everywhere and the application
returns |psi| outside the aperture), a single-point table gives zero, the
2-D MTF of an even table is real and even, and a numpy restatement of the
construction (its real-space band limit by k index included) agrees on a
small grid."""
import numpy as np
import pytest

from helpers import Case, random_atoms, rel_l2

FLAT = dict(mtf=[1.0, 1.0], mtf_domain=[0.0, 1e-10], mtf_rate=1e12)


def _job(O, nx=32, ny=48, L=(8.0, 10.0), **kw):
    return O.make_job(nx=nx, ny=ny, xperiod=L[0], yperiod=L[1], zperiod=5.0, voltage=200e3,
                      alpha_max=0.015, **kw)


def test_flat_mtf_is_one(O):
    j = _job(O)
    m = O.mtf_2d(j, FLAT["mtf"], FLAT["mtf_domain"], FLAT["mtf_rate"])
    np.testing.assert_allclose(m, 1.0, rtol=0, atol=1e-13)


def test_single_point_table_is_zero(O):
    j = _job(O)
    assert (O.mtf_2d(j, [0.7], [0.5], 3.0) == 0).all()


def test_flat_mtf_application_is_masked_magnitude(O):
    j = _job(O)
    kx, ky, _, _ = O.domains(j.nx, j.ny, j.xperiod, j.yperiod)
    _, bwt = O.bwcutoffs(j.nx, j.ny, j.xperiod, j.yperiod)
    lam = O.wavelength(j.voltage)[0]
    rng = np.random.default_rng(3)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    psi = (rng.normal(size=k2.shape) + 1j * rng.normal(size=k2.shape)) * (k2 < 0.9 * bwt ** 2)
    out = O.apply_mtf(j, psi, np.ones_like(psi))
    want = np.where(lam ** 2 * k2 > j.alpha_max ** 2, np.abs(psi), 0.0)
    assert (want == 0).sum() > 5 and (want > 0).sum() > 100
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-12)


def _numpy_mtf(j, mtf, dom, rate):
    """The construction restated with numpy's FFT (unnormalised backward =
    ifft2 * NxNy), the band limit applied to the real-space array by k index."""
    O_kx = np.fft.fftfreq(j.nx, j.xperiod / j.nx)
    O_ky = np.fft.fftfreq(j.ny, j.yperiod / j.ny)
    xxx = j.xperiod ** 2 * O_kx[:, None] / j.nx
    yyy = j.yperiod ** 2 * O_ky[None, :] / j.ny
    rsqr = xxx ** 2 + yyy ** 2
    d2 = (0.5 * rate * np.asarray(dom)) ** 2
    m = np.zeros(rsqr.shape)
    for k in range(len(mtf) - 1):
        sel = (rsqr >= d2[k]) & (rsqr < d2[k + 1])
        lin = mtf[k] + (mtf[k + 1] - mtf[k]) * (np.sqrt(rsqr) - np.sqrt(d2[k])) / (
            np.sqrt(d2[k + 1]) - np.sqrt(d2[k]))
        m = np.where(sel, lin * np.sinc(xxx / rate / np.pi) * np.sinc(yyy / rate / np.pi), m)
    a = np.fft.ifft2(m) * (j.nx * j.ny)
    bwt = (2.0 / 3.0) * min(j.nx / (2 * j.xperiod), j.ny / (2 * j.yperiod))
    a[O_kx[:, None] ** 2 + O_ky[None, :] ** 2 >= bwt ** 2] = 0
    return m, np.fft.fft2(a) / (j.nx * j.ny)


def test_mtf_2d_against_numpy_restatement(O):
    j = _job(O, nx=16, ny=24, L=(6.0, 9.0))
    mtf = [1.0, 0.8, 0.55, 0.3, 0.2]
    dom = [0.0, 0.25, 0.5, 0.75, 1.0]
    rate = 2 * 4.5              # the table reaches rsqr = 4.5^2: part of the grid lies beyond
    m_np, want = _numpy_mtf(j, mtf, dom, rate)
    assert (m_np == 0).sum() > 10 and (m_np != 0).sum() > 100
    got = O.mtf_2d(j, mtf, dom, rate)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    # an even table gives a real, even 2-D MTF
    assert np.abs(got.imag).max() < 1e-15
    neg = got[(-np.arange(16)) % 16][:, (-np.arange(24)) % 24]
    np.testing.assert_allclose(neg, got, rtol=0, atol=1e-15)
    # and the real-space band limit matters: without it the result is m itself
    assert np.abs(got - m_np).max() > 1e-3


def test_fem_with_flat_mtf_equals_fem_on_masked_magnitude(O):
    """A whole raster with the flat MTF: the oracle's exit waves are |psi|
    outside the aperture and V_omega is that of those magnitudes."""
    q, Z = random_atoms(25, (8.0, 8.0, 5.0), seed=12, species=(14, 29))
    c = Case(nx=64, ny=64, cell=(8.0, 8.0, 5.0), voltage=200e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.015, cs3=2e5, defocus=20.0, fem=True, dr=0.5)
    xp, yp = np.array([0.5, 3.0, 6.5]), np.array([1.0, 4.0])
    _, (img0, v0, ex0) = c.run_oracle(O, xp, yp, exit_waves=True)
    cm = Case(**{**c.__dict__, **FLAT})
    _, (img1, v1, ex1) = cm.run_oracle(O, xp, yp, exit_waves=True)
    np.testing.assert_array_equal(img1, img0)            # the ADF is taken before the MTF
    kx, ky, _, _ = O.domains(64, 64, 8.0, 8.0)
    lam = O.wavelength(200e3)[0]
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    A = np.where(lam ** 2 * k2 > 0.015 ** 2, np.abs(ex0), 0.0)
    np.testing.assert_allclose(ex1, A, rtol=0, atol=1e-12 * np.abs(ex0).max())
    _, bwt = O.bwcutoffs(64, 64, 8.0, 8.0)
    nb, dk = O.fem_nbins(kx, ky, 0.5, bwt)
    rr = [O.fem_radial(a.astype(np.complex128), kx, ky, dk, nb, lam, 0.015)
          for a in A.reshape(-1, 64, 64)]
    r = np.stack([x for x, _ in rr])
    cnt = rr[0][1]
    with np.errstate(divide="ignore", invalid="ignore"):
        want = O.fem_finalize((r / cnt).sum(0), (r ** 2 / (cnt * cnt)).sum(0), r.shape[0])
    # rings whose pixels all lie inside the aperture hold only FFT round-off
    # after the MTF (a flat MTF is a delta in k): V is undefined there
    live = (r > 0).any(0)
    assert (~live).sum() >= 1 and live.sum() > 10
    assert rel_l2(v1[live], want[live]) < 1e-10
    assert rel_l2(v1, v0) > 1e-6                          # the direct beam really is removed
