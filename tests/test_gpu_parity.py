"""GPU parity: libmsf.so (HIP, gfx950) against the FP64 oracle on the same
seeded inputs, at sizes the oracle finishes in seconds.

Tolerances (relative L2 over the compared array):
  transmission functions, probe (FP64 on both sides)        1e-11
  exit-wave intensity |psi(k)|^2 (FP32 wave)                 1e-5
  ADF image and V_omega (FP32 wave, FP64 accumulators)       1e-4
"""
import numpy as np
import pytest

from helpers import Case, random_atoms, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _case_small(**kw):
    q, Z = random_atoms(60, (12.0, 12.0, 8.0), seed=7, species=(14, 29))
    base = dict(nx=64, ny=64, cell=(12.0, 12.0, 8.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
                alpha_max=0.02, inner=0.03, outer=0.2)
    base.update(kw)
    return Case(**base)


def test_slices_and_transmission_match_oracle(msf, ctx, O):
    c = _case_small()
    spec_o = O.Specimen(c.oracle_job(O), c.q, c.Z.astype(np.int32))
    spec_g = c.gpu_specimen(msf, ctx)
    lb, th, na = spec_g.slice_info()
    assert spec_g.n_slices == spec_o.n_slices
    np.testing.assert_array_equal(lb, spec_o.lower_bound)
    np.testing.assert_array_equal(th, spec_o.thickness)
    assert na.sum() == c.Z.size
    for s in range(spec_g.n_slices):
        err = rel_l2(spec_g.transmission(s), spec_o.t_stack[s])
        assert err < 1e-11, (s, err)


def test_probe_matches_oracle(msf, ctx, O):
    for corrected, cs3, cs5, df in ((0, 1.2e7, 0.0, 477.0), (1, -2.0e5, 5.0e7, -30.0)):
        c = _case_small(corrected=corrected, cs3=cs3, cs5=cs5, defocus=df, voltage=200e3,
                        alpha_max=0.03)
        pg = ctx.probe_realspace(c.grid(msf), c.optics(msf))
        po = O.probe(c.oracle_job(O))
        assert rel_l2(pg, po) < 1e-11
        assert abs(np.sum(np.abs(pg) ** 2) - 1.0) < 1e-12


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", [64, 128])
def test_exit_waves_match_oracle(msf, ctx, O, mode, n):
    c = _case_small(nx=n, ny=n, transmit_mode=mode)
    xp = np.array([0.0, 3.3, 6.1, 11.9])
    yp = np.array([0.4, 7.77])
    spec_o, (_, _, ex_o) = c.run_oracle(O, xp, yp, exit_waves=True)
    ctx.configure(batch=3, transmit_mode=mode)
    spec_g = c.gpu_specimen(msf, ctx)
    ex_g = ctx.exit_waves(spec_g, c.optics(msf), xp, yp)
    ctx.configure()
    for p in range(xp.size * yp.size):
        err = rel_l2(np.abs(ex_g[p]) ** 2, np.abs(ex_o[p]) ** 2)
        assert err < 1e-5, (p, err)


def test_non_square_grid(msf, ctx, O):
    q, Z = random_atoms(40, (12.0, 6.0, 6.0), seed=3)
    c = Case(nx=128, ny=64, cell=(12.0, 6.0, 6.0), voltage=200e3, q=q, Z=Z, min_slice=1.0,
             alpha_max=0.025, defocus=20.0, cs3=1e6)
    xp = np.array([1.0, 9.5])
    yp = np.array([2.2, 5.9])
    _, (img_o, _, ex_o) = c.run_oracle(O, xp, yp, exit_waves=True)
    spec_g = c.gpu_specimen(msf, ctx)
    ex_g = ctx.exit_waves(spec_g, c.optics(msf), xp, yp)
    for p in range(4):
        assert rel_l2(np.abs(ex_g[p]) ** 2, np.abs(ex_o[p]) ** 2) < 1e-5
    img_g, _ = ctx.adfstem(spec_g, c.optics(msf), xp, yp, detector=c.detector(msf))
    assert rel_l2(img_g, img_o) < 1e-4


def test_adf_and_fem_match_oracle(msf, ctx, O):
    c = _case_small(nx=128, ny=128, fem=True, dr=0.5, alpha_max=0.01, inner=0.02, outer=0.15)
    xp = np.linspace(0.0, 11.0, 5)
    yp = np.linspace(0.5, 11.5, 4)
    _, (img_o, v_o, _) = c.run_oracle(O, xp, yp)
    spec_g = c.gpu_specimen(msf, ctx)
    img_g, v_g = ctx.adfstem(spec_g, c.optics(msf), xp, yp, detector=c.detector(msf),
                             fem=c.femcfg(msf))
    assert rel_l2(img_g, img_o) < 1e-4
    assert v_g.shape == v_o.shape
    assert rel_l2(v_g, v_o) < 1e-4


def test_batch_composition_does_not_change_results(msf, ctx, O):
    c = _case_small(nx=64, ny=64, fem=True, dr=0.7)
    xp = np.linspace(0.0, 11.0, 7)
    yp = np.linspace(0.0, 11.0, 5)
    spec_g = c.gpu_specimen(msf, ctx)
    outs = []
    for b, ch in ((0, 1), (1, 1), (6, 1), (35, 1), (6, 2), (1, 2)):
        ctx.configure(batch=b, chains=ch)
        outs.append(ctx.adfstem(spec_g, c.optics(msf), xp, yp, detector=c.detector(msf),
                                fem=c.femcfg(msf)))
    ctx.configure()
    for img, v in outs[1:]:
        np.testing.assert_array_equal(img, outs[0][0])
        np.testing.assert_array_equal(v, outs[0][1])
