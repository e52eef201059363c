"""Pins the oracle's probe, slice loop and detectors with closed forms:
the k-space probe, vacuum propagation, the reference's in-place transmission
update on a uniform transmission function, ring sums and V_omega algebra."""
import numpy as np
import pytest


def _setup(O, n=64, L=12.0, v=300e3, **kw):
    job = O.make_job(nx=n, ny=n, xperiod=L, yperiod=L, zperiod=L, voltage=v, **kw)
    kx, ky, xx, yy = O.domains(n, n, L, L)
    lam, gam = O.wavelength(v)
    pap, bwt = O.bwcutoffs(n, n, L, L)
    return job, kx, ky, xx, yy, lam, bwt


def _H(kx, ky, lam, alpha, chi):
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    inside = lam * lam * k2 < alpha * alpha
    return np.where(inside, np.exp(-1j * chi(k2)), 0.0), inside.sum()


@pytest.mark.parametrize("corrected", [0, 1])
def test_probe_is_normalised_aperture_function(O, corrected):
    pi_ref = 3.14159265369
    job, kx, ky, xx, yy, lam, _ = _setup(O, alpha_max=0.03, cs3=2.0e6, cs5=3.0e7,
                                         defocus=150.0, corrected=corrected)
    p = O.probe(job)
    assert abs(np.sum(np.abs(p) ** 2) - 1.0) < 1e-12
    if corrected:
        chi = lambda k2: pi_ref * lam * k2 * (-150.0 + 0.5 * 2.0e6 * lam ** 2 * k2
                                              + (1 / 3) * 3.0e7 * (lam ** 2 * k2) ** 2)
    else:
        chi = lambda k2: pi_ref * lam * k2 * (0.5 * 2.0e6 * lam ** 2 * k2 - 150.0)
    H, nap = _H(kx, ky, lam, 0.03, chi)
    n = kx.size
    # FFT_u(probe) = N^2 H / (N sqrt(n_ap)) for the unshifted probe at (0, 0)
    pk = np.fft.fft2(p) / n * np.sqrt(nap)
    assert np.abs(pk - H).max() < 1e-11


def test_probe_shifts_persist_and_follow_raster_order(O):
    xx = np.arange(8) * 0.5
    yy = np.arange(8) * 0.5
    sx, sy = O.probe_shifts(xx, yy, [0.6, 3.1, 3.4], [1.1, 3.3, 0.0])
    # x outer, y inner; 3.3 and 3.4 lie beyond x[N-1] - dx/2 and keep the previous index
    assert list(sx) == [1, 1, 1, 6, 6, 6, 6, 6, 6]
    assert list(sy) == [2, 2, 0, 2, 2, 0, 2, 2, 0]


def _uniform_stack(n, S, value):
    return np.full((S, n, n), value, np.complex128)


def test_vacuum_propagation_is_masked_propagator_product(O):
    job, kx, ky, xx, yy, lam, bwt = _setup(O, alpha_max=0.05, defocus=40.0)
    p = O.probe(job)
    S = 3
    th = [1.5, 2.0, 0.7]
    px = np.stack([O.propagator(lam, t, kx, bwt) for t in th])
    py = np.stack([O.propagator(lam, t, ky, bwt) for t in th])
    n = kx.size
    psi = O.propagate_position(p, 5, 9, _uniform_stack(n, S, 1.0), px, py, kx, ky, bwt)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    expect = np.fft.fft2(np.roll(np.roll(p, 5, 0), 9, 1)) / n
    for s in range(S):
        expect = expect * px[s][:, None] * py[s][None, :] * (k2 < bwt * bwt)
    assert np.abs(psi - expect).max() < 1e-13
    # propagators are pure phases inside the band limit: the norm is kept
    assert abs(np.sum(np.abs(psi) ** 2) - np.sum(np.abs(expect) ** 2)) < 1e-13


@pytest.mark.parametrize("mode", [0, 1])
def test_uniform_transmission_single_slice(O, mode):
    """t = c + i d everywhere.  The reference (adfstem.cpp:1922-1952) writes
    Re(psi) first and then uses it for Im(psi); mode 1 is the exact product."""
    job, kx, ky, xx, yy, lam, bwt = _setup(O, alpha_max=0.02)  # aperture inside the band limit
    n = kx.size
    p = O.probe(job)
    c, d = 0.8, 0.6
    px = O.propagator(lam, 1.7, kx, bwt)[None]
    py = O.propagator(lam, 1.7, ky, bwt)[None]
    psi = O.propagate_position(p, 0, 0, _uniform_stack(n, 1, c + 1j * d), px, py, kx, ky, bwt,
                               transmit_mode=mode)
    u = n * p  # IFFT_u(FFT_u(probe)/s) = s * probe
    a = (c * u.real - d * u.imag) / n
    b = (c * u.imag + d * (a if mode == 0 else u.real)) / n
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    expect = np.fft.fft2(a + 1j * b) / n * px[0][:, None] * py[0][None, :] * (k2 < bwt * bwt)
    assert np.abs(psi - expect).max() < 1e-13
    if mode == 1:  # |t| = 1: the exact product keeps the norm of the band-limited probe
        assert abs(np.sum(np.abs(psi) ** 2) - 1.0) < 1e-12


def test_adf_ring_is_half_open_in_k_squared(O):
    n, L = 16, 4.0
    kx, ky, _, _ = O.domains(n, n, L, L)
    psi = np.zeros((n, n), complex)
    psi[1, 0] = 1.0   # k^2 = 1/16
    psi[2, 0] = 2.0   # k^2 = 4/16
    psi[0, 3] = 3.0   # k^2 = 9/16
    k1, k2 = 1 / 16, 9 / 16   # (k1, k2]: excludes 1/16, includes 9/16
    assert O.adf(psi, kx, ky, k1, k2) == pytest.approx(4.0 + 9.0, rel=1e-15)
    assert O.adf(psi, kx, ky, 0.0, 1.0) == pytest.approx(14.0, rel=1e-15)


def test_fem_bins_membership_counts_and_direct_beam(O):
    n, L = 32, 8.0
    kx, ky, _, _ = O.domains(n, n, L, L)
    lam = O.wavelength(300e3)[0]
    _, bwt = O.bwcutoffs(n, n, L, L)
    nb, dk = O.fem_nbins(kx, ky, 1 / np.sqrt(2), bwt)
    assert dk == pytest.approx(1 / L, rel=1e-12)
    assert nb == int(bwt / dk)
    psi = np.ones((n, n), complex) * 2.0
    r, c = O.fem_radial(psi, kx, ky, dk, nb, lam, 0.0)
    # bin ii holds the pixels with ii*dk < |k| <= (ii+1)*dk (bin 0 includes k = 0)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    edges2 = ((np.arange(nb + 1)) * dk) ** 2
    for ii in range(nb):
        lo = -1.0 if ii == 0 else edges2[ii]
        cnt = np.sum((k2 > lo) & (k2 <= edges2[ii + 1]))
        assert c[ii] == cnt
        assert r[ii] == pytest.approx(2.0 * cnt, rel=1e-14)
    # bins whose outer edge satisfies lambda * k <= alpha_max are zeroed
    amax = 2.5 * dk * lam
    r2, _ = O.fem_radial(psi, kx, ky, dk, nb, lam, amax)
    assert list(r2[:2]) == [0.0, 0.0] and r2[2] > 0


def test_fem_variance_closed_forms(O):
    # r_p in {a, b} alternately: V = N sum r^2 / (sum r)^2 - 1 = ((a-b)/(a+b))^2
    a, b, n = 3.0, 1.0, 10
    s1 = np.array([n / 2 * (a + b), 7.0, 0.0, np.nan])
    s2 = np.array([n / 2 * (a * a + b * b), 7.0 * 7.0 / n, 0.0, np.nan])
    v = O.fem_finalize(s1, s2, n)
    assert v[0] == pytest.approx(((a - b) / (a + b)) ** 2, rel=1e-14)
    assert v[1] == pytest.approx(0.0, abs=1e-15)   # constant r
    assert v[2] == 0.0                              # sum r == 0 -> 0
    assert np.isnan(v[3])                           # empty ring (0/0) stays NaN


def test_factored_step_reproduces_the_position_loop(O):
    """orc_initial_wave followed by orc_slice_step over every slice is the
    whole-position loop orc_propagate_position, bit for bit."""
    from helpers import Case, random_atoms
    q, Z = random_atoms(20, (6.0, 6.0, 4.0), seed=12, species=(14, 29))
    c = Case(nx=32, ny=32, cell=(6.0, 6.0, 4.0), voltage=300e3, q=q, Z=Z, min_slice=1.0,
             alpha_max=0.02)
    job = c.oracle_job(O)
    spec = O.Specimen(job, c.q, c.Z.astype(np.int32))
    kx, ky, xx, yy = O.domains(32, 32, 6.0, 6.0)
    _, bwt = O.bwcutoffs(32, 32, 6.0, 6.0)
    probe = O.probe(job)
    _, _, ex = O.adfstem(job, spec, np.array([xx[5]]), np.array([yy[9]]), exit_waves=True)
    sx, sy = (np.asarray(a)[0] for a in O.probe_shifts(xx, yy, np.array([xx[5]]), np.array([yy[9]])))
    w = O.initial_wave(probe, int(sx), int(sy))
    for n in range(spec.n_slices):
        w = O.slice_step(w, spec.t_stack[n], spec.px[n], spec.py[n], kx, ky, bwt)
    np.testing.assert_array_equal(w, ex[0])
