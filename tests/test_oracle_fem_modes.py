"""The oracle's other FEM modes (Vbar_r, V_re, Omega_vimage, V_gt17, V_rbar;
adfstem.cpp:2688-2781, :2896-3605) against closed forms on synthetic
intensities: a constant field, a field alternating between two values over
positions, and a field that varies over pixels but not over positions."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def geom(O):
    nx = ny = 32
    kx, ky, _, _ = O.domains(nx, ny, 8.0, 8.0)
    _, bwt = O.bwcutoffs(nx, ny, 8.0, 8.0)
    nb, dk = O.fem_nbins(kx, ky, 1 / np.sqrt(2), bwt)
    lam = 0.025
    alpha = lam * dk              # bin 0 (upper edge dk) is the direct beam; the aperture
                                  # (lambda k <= alpha) then lies inside bin 0 as well
    return dict(nx=nx, ny=ny, kx=kx, ky=ky, nb=nb, dk=dk, lam=lam, alpha=alpha)


def _run(O, g, mag):
    """mag: per-position |psi| [P, nx, ny]"""
    P = mag.shape[0]
    r1 = np.zeros((P, g["nb"]))
    c1 = np.zeros((P, g["nb"]), np.int32)
    for p in range(P):
        r1[p], c1[p] = O.fem_radial(mag[p].astype(np.complex128), g["kx"], g["ky"], g["dk"],
                                    g["nb"], g["lam"], g["alpha"])
    out = O.fem_modes(mag ** 2, g["kx"], g["ky"], g["dk"], g["nb"], g["lam"], g["alpha"], r1, c1)
    return out, c1[0]


def _live(counts, g):
    zeroed = g["lam"] * (np.arange(1, g["nb"] + 1) * g["dk"]) <= g["alpha"]
    return (counts > 0) & ~zeroed, zeroed


def test_constant_field(O, geom):
    mag = np.full((3, geom["nx"], geom["ny"]), 0.37)
    out, counts = _run(O, geom, mag)
    live, zeroed = _live(counts, geom)
    assert zeroed[0] and not zeroed[1] and live.sum() > 3
    for m in ("vbar_r", "v_re", "v_rbar", "omega_vimage"):
        np.testing.assert_allclose(out[m][live], 0.0, atol=1e-12)
    np.testing.assert_allclose(out["v_gt17"][live], 1.0, rtol=1e-12)
    # direct-beam bins: ring sums zeroed -> 0/0 in the ratio modes, 0 in the
    # annular mean of the (zeroed) variance image
    for m in ("vbar_r", "v_re", "v_rbar", "v_gt17"):
        assert np.isnan(out[m][zeroed]).all(), m
    assert (out["omega_vimage"][zeroed] == 0.0).all()


def test_two_valued_over_positions(O, geom):
    a, b = 0.2, 0.7
    mag = np.stack([np.full((geom["nx"], geom["ny"]), v) for v in (a, b, a, b)])
    out, counts = _run(O, geom, mag)
    live, _ = _live(counts, geom)
    v = (a - b) ** 2 / (a + b) ** 2
    np.testing.assert_allclose(out["v_re"][live], v, rtol=1e-12)
    np.testing.assert_allclose(out["omega_vimage"][live], v, rtol=1e-12)
    np.testing.assert_allclose(out["v_gt17"][live], v + 1.0, rtol=1e-12)
    np.testing.assert_allclose(out["v_rbar"][live], 0.0, atol=1e-12)
    np.testing.assert_allclose(out["vbar_r"][live], 0.0, atol=1e-12)


def test_pixel_pattern_constant_over_positions(O, geom):
    rng = np.random.default_rng(5)
    pattern = 0.1 + rng.random((geom["nx"], geom["ny"]))
    mag = np.stack([pattern] * 3)
    out, counts = _run(O, geom, mag)
    live, _ = _live(counts, geom)
    # the ring's own spread: n sum a^2 / (sum a)^2 - 1, the same at every position
    assert (out["v_re"][live] > 1e-3).all()
    np.testing.assert_allclose(out["v_rbar"][live], out["v_re"][live], rtol=1e-12)
    np.testing.assert_allclose(out["vbar_r"][live], out["v_re"][live], rtol=1e-12)
    np.testing.assert_allclose(out["v_gt17"][live], 1.0, rtol=1e-12)
    np.testing.assert_allclose(out["omega_vimage"][live], 0.0, atol=1e-12)


def test_correlograph_uniform_field(O):
    """A uniform |psi| gives, per ring ii and phi sector jj, count_ii *
    (n_s A) / (count_ii A) - 1 = n_s - 1 (n_s = pixels of the sector), 0 on
    empty sectors and on direct-beam rings; the sectors of a ring hold all of
    its pixels."""
    nx = ny = 48
    kx, ky, _, _ = O.domains(nx, ny, 6.0, 6.0)
    _, bwt = O.bwcutoffs(nx, ny, 6.0, 6.0)
    nb, dk = O.fem_nbins(kx, ky, 0.5, bwt)
    nphi = O.lib().orc_correlograph_nphi(dk, nb)
    assert nphi > 4
    lam, alpha = 0.025, 0.025 * dk
    P, A = 3, 0.4
    mag = np.full((P, nx, ny), A)
    r1 = np.zeros((P, nb)); c1 = np.zeros((P, nb), np.int32)
    for p in range(P):
        r1[p], c1[p] = O.fem_radial(mag[p].astype(np.complex128), kx, ky, dk, nb, lam, alpha)
    avg = np.zeros(nb * nphi); sqr = np.zeros(nb * nphi)
    each = np.zeros(P * nb * nphi)
    assert O.lib().orc_correlograph((mag ** 2).reshape(-1), P, kx, ky, nx, ny, dk, nb, lam,
                                    alpha, r1.reshape(-1), c1.reshape(-1), avg, sqr,
                                    each.ctypes.data) == 0
    # the per-position values (every position alike here) average to avg
    each = each.reshape(P, nphi, nb)
    np.testing.assert_allclose(each.mean(0), avg.reshape(nphi, nb), atol=1e-12)
    np.testing.assert_allclose((each ** 2).sum(0), sqr.reshape(nphi, nb), atol=1e-10)
    avg = avg.reshape(nphi, nb)
    sqr = sqr.reshape(nphi, nb)
    # sector populations, counted independently
    K2 = kx[:, None] ** 2 + ky[None, :] ** 2
    phi = 3.14159265369 + np.arctan2(ky[None, :], kx[:, None])
    dphi = 2 * 3.14159265369 / nphi
    sec = np.clip(np.ceil(phi / dphi).astype(int) - 1, 0, nphi - 1)
    order = np.argsort(K2, axis=None, kind="stable")
    ring = np.full(nx * ny, -1)
    for ii in range(nb):
        lo = (ii * dk) ** 2 if ii else -1.0
        m = (K2.reshape(-1) <= ((ii + 1) * dk) ** 2) & (K2.reshape(-1) > lo)
        ring[m] = ii
    for ii in range(nb):
        zeroed = lam * (ii + 1) * dk <= alpha
        ns = np.bincount(sec.reshape(-1)[ring == ii], minlength=nphi)
        want = np.where(ns > 0, ns - 1.0, 0.0) if not zeroed else np.zeros(nphi)
        np.testing.assert_allclose(avg[:, ii], want, atol=1e-12, err_msg=str(ii))
        np.testing.assert_allclose(sqr[:, ii], P * want ** 2, atol=1e-10)
