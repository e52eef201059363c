"""GPU parity of the FEM modes (V_omega, Vbar_r, V_re, Omega_vimage, V_gt17,
V_rbar; adfstem.cpp:2489-3605) against the FP64 oracle: 1e-4 relative L2
over the bins, identical NaN patterns, and bit-identical results for any
batch depth."""
import numpy as np
import pytest

from helpers import Case, random_atoms, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _case():
    q, Z = random_atoms(70, (12.0, 12.0, 8.0), seed=41, species=(14, 29, 40))
    return Case(nx=128, ny=128, cell=(12.0, 12.0, 8.0), voltage=200e3, q=q, Z=Z, min_slice=1.5,
                alpha_max=0.012, cs3=5e5, defocus=30.0, inner=0.02, outer=0.15, fem=True, dr=0.5)


def test_fem_modes_match_oracle(msf, ctx, O):
    c = _case()
    xp = np.linspace(0.3, 11.2, 5)
    yp = np.linspace(0.8, 11.6, 4)
    img_o, v_o = O.adfstem_fem(c.oracle_job(O), O.Specimen(c.oracle_job(O), c.q,
                                                             c.Z.astype(np.int32)), xp, yp)
    spec = c.gpu_specimen(msf, ctx)
    img_g, v_g = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf),
                                 detector=c.detector(msf))
    assert rel_l2(img_g, img_o) < 1e-4
    for m in msf.FEM_MODES:
        assert v_g[m].shape == v_o[m].shape, m
        err = rel_l2(v_g[m], v_o[m])
        print(m, err, flush=True)
        assert err < 1e-4, (m, err)
    # the V_omega of the plain entry point is the same computation
    _, v1 = ctx.adfstem(spec, c.optics(msf), xp, yp, fem=c.femcfg(msf))
    np.testing.assert_array_equal(v1, v_g["v_omega"])


def test_fem_modes_do_not_depend_on_batch(msf, ctx):
    c = _case()
    xp = np.linspace(0.0, 11.0, 6)
    yp = np.linspace(0.5, 10.5, 5)
    spec = c.gpu_specimen(msf, ctx)
    outs = []
    try:
        for b, graph, ch in ((0, True, 1), (1, False, 1), (7, True, 1), (7, False, 1),
                             (7, True, 2), (7, False, 2), (4, True, 2)):
            ctx.configure(batch=b, graph=graph, chains=ch)
            outs.append(ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf),
                                        correlograph=True)[1])
    finally:
        ctx.configure()
    for o in outs[1:]:
        for m in o:
            np.testing.assert_array_equal(o[m], outs[0][m], err_msg=m)


def test_fem_mode_subsets(msf, ctx):
    """Requesting a subset computes the same values as requesting all."""
    c = _case()
    xp = np.array([1.0, 5.5])
    yp = np.array([2.0, 9.0, 10.0])
    spec = c.gpu_specimen(msf, ctx)
    _, full = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf))
    for sub in (("vbar_r",), ("omega_vimage", "v_rbar"), ("v_omega", "v_gt17")):
        _, part = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf), modes=sub)
        for m in sub:
            np.testing.assert_array_equal(part[m], full[m], err_msg=m)


def test_correlograph_matches_oracle(msf, ctx, O):
    """The correlograph (adfstem.cpp:2603-2650, :3620-3632): mean and sum of
    squares over positions of the per-position sector ratios, [n_phi, n_bins],
    against the oracle at 1e-4 relative L2; identical for any batch depth."""
    c = _case()
    xp = np.linspace(0.3, 11.2, 4)
    yp = np.linspace(0.8, 11.6, 5)
    job = c.oracle_job(O)
    _, o = O.adfstem_fem(job, O.Specimen(job, c.q, c.Z.astype(np.int32)), xp, yp,
                         modes=("v_omega",), correlograph=True, correlograph_each=True)
    spec = c.gpu_specimen(msf, ctx)
    nphi, k_edges, phi_edges = msf.correlograph_bins(spec.grid, c.dr)
    nb = o["v_omega"].size
    assert nphi == O.lib().orc_correlograph_nphi(k_edges[1], nb) and nphi > 4
    print("correlograph", nb, nphi, flush=True)
    _, g = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf), modes=("v_omega",),
                           correlograph=True, correlograph_each=True)
    # per position (--correlograph_everyimage / _everytxt)
    assert g["correlograph_each"].shape == o["correlograph_each"].shape == (xp.size * yp.size,
                                                                             nphi, nb)
    err = rel_l2(g["correlograph_each"], o["correlograph_each"])
    print("correlograph_each", err, flush=True)
    assert err < 1e-4
    np.testing.assert_allclose(g["correlograph_each"].mean(0), g["correlograph_avg"],
                               rtol=1e-12, atol=1e-15)
    for m in ("correlograph_avg", "correlograph_sqr_sum"):
        assert g[m].shape == o[m].shape == (nphi, nb), m
        assert (o[m] != 0).sum() > nb, m          # a populated correlograph
        err = rel_l2(g[m], o[m])
        print(m, err, flush=True)
        assert err < 1e-4, (m, err)
    np.testing.assert_array_equal(g["v_omega"], ctx.adfstem_fem(
        spec, c.optics(msf), xp, yp, c.femcfg(msf), modes=("v_omega",))[1]["v_omega"])
    try:
        ctx.configure(batch=7, graph=False)
        _, g7 = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf), modes=(),
                                correlograph=True)
    finally:
        ctx.configure()
    for m in ("correlograph_avg", "correlograph_sqr_sum"):
        np.testing.assert_array_equal(g7[m], g[m], err_msg=m)


@pytest.mark.parametrize("mtf", [False, True])
def test_pixel_images_match_oracle_exit_waves(msf, ctx, O, mtf):
    """Per-pixel averages of |psi| and |psi|^2 over the raster and the 2-D
    variance <|psi|^2>/<|psi|>^2 - 1, zeroed inside the aperture
    (adfstem.cpp:2966-3000, :3468-3492), against the same quantities taken
    from the oracle's exit waves; NaN (0/0) exactly where the oracle's waves
    vanish at every position, outside the band limit."""
    c = _case()
    if mtf:
        d = np.linspace(0.0, 1.0, 11)
        c = Case(**{**c.__dict__, "mtf": 1.0 / (1.0 + (d / 0.4) ** 2), "mtf_domain": d,
                    "mtf_rate": 12.0})
    xp = np.linspace(0.3, 11.2, 4)
    yp = np.linspace(0.8, 11.6, 3)
    job = c.oracle_job(O)
    _, r_o = O.adfstem_fem(job, O.Specimen(job, c.q, c.Z.astype(np.int32)), xp, yp,
                           modes=("omega_vimage",), exit_waves=True)
    ex = r_o["exit_waves"]
    a_o = np.abs(ex).mean(axis=0)
    s_o = (np.abs(ex) ** 2).mean(axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        v_o = s_o / (a_o * a_o) - 1.0
    kx, ky, _, _ = O.domains(c.nx, c.ny, c.cell[0], c.cell[1])
    lam = O.wavelength(c.voltage)[0]
    v_o[lam ** 2 * (kx[:, None] ** 2 + ky[None, :] ** 2) <= c.alpha_max ** 2] = 0.0
    spec = c.gpu_specimen(msf, ctx)
    _, r_g = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf),
                             detector=c.detector(msf), pixel_images=True)
    for name, got, want in (("mean_abs", r_g["diffraction_mean_abs"], a_o),
                            ("mean_sqr", r_g["diffraction_mean_sqr"], s_o)):
        assert got.shape == (c.nx, c.ny)
        err = rel_l2(got, want)
        print(name, mtf, err, flush=True)
        assert err < 1e-4, (name, err)
    v_g = r_g["variance_2d"]
    np.testing.assert_array_equal(np.isnan(v_g), np.isnan(v_o))
    assert np.isnan(v_o).any() and np.isfinite(v_o).any()
    fin = np.isfinite(v_o)
    err = rel_l2(v_g[fin], v_o[fin])
    print("variance_2d", mtf, err, flush=True)
    assert err < 1e-4
    # the Omega_vimage mode is the ring average of this image
    assert rel_l2(r_g["omega_vimage"], r_o["omega_vimage"]) < 1e-4
