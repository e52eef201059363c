"""The two radix plans of the slice passes (msf_config.fft_plan): the primary
plan of every length and the alternate plans of 256 (16 x 16), 512
(16 x 8 x 4) and 1024 (16 x 16 x 4) with 16 values per lane.  Each plan
against the FP64 oracle at the usual tolerances (|psi|^2 1e-5, ADF and
V_omega 1e-4); the two plans differ by FP32 rounding only; the automatic
choice per length (alternate at 256 and 1024) and the autotuner's choice
reproduce bit for bit when set explicitly."""
import numpy as np
import pytest

from helpers import Case, random_atoms, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _case(n, seed):
    q, Z = random_atoms(50, (10.0, 10.0, 6.0), seed=seed, species=(14, 29))
    return Case(nx=n, ny=n, cell=(10.0, 10.0, 6.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
                alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.5)


def _plans(ctx):
    st = ctx.stats()
    return st["row_fft_plan"], st["col_fft_plan"]


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_both_plans_match_oracle(msf, ctx, O, n):
    c = _case(n, 60 + n // 256)
    xp = np.array([0.0, 4.3, 7.9])
    yp = np.array([1.1, 6.6])
    _, (img_o, v_o, ex_o) = c.run_oracle(O, xp, yp, exit_waves=True)
    spec = c.gpu_specimen(msf, ctx)
    run = lambda: ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                              fem=c.femcfg(msf))
    res = {}
    try:
        for fp in (1, 2):
            print("grid", n, "fft_plan", fp, flush=True)
            ctx.configure(fft_plan=fp)
            ex = ctx.exit_waves(spec, c.optics(msf), xp, yp)
            img, v = run()
            assert _plans(ctx) == (fp, fp)
            for p in range(xp.size * yp.size):
                err = rel_l2(np.abs(ex[p]) ** 2, np.abs(ex_o[p]) ** 2)
                assert err < 1e-5, (fp, p, err)
            assert rel_l2(img, img_o) < 1e-4
            assert rel_l2(v, v_o) < 1e-4
            res[fp] = (ex, img, v)
        # the automatic choice: the alternate at 256 and 1024, the primary at 512
        ctx.configure()
        auto = 1 if n == 512 else 2
        img, v = run()
        assert _plans(ctx) == (auto, auto)
        np.testing.assert_array_equal(img, res[auto][1])
        np.testing.assert_array_equal(v, res[auto][2])
    finally:
        ctx.configure()
    assert not np.array_equal(res[1][0], res[2][0])     # the alternate arithmetic ran
    d = rel_l2(res[2][0], res[1][0])
    print("plan 2 vs plan 1 exit waves", d, flush=True)
    assert d < 1e-5


def test_lengths_without_alternate_use_the_primary_plan(msf, ctx):
    c = _case(128, 70)
    xp = np.array([0.5, 5.5])
    yp = np.array([2.0, 8.0])
    spec = c.gpu_specimen(msf, ctx)
    outs = []
    try:
        for fp in (1, 2, 3, 4):
            ctx.configure(fft_plan=fp)
            outs.append(ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                                    fem=c.femcfg(msf)))
            assert _plans(ctx) == (1, 1)
    finally:
        ctx.configure()
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0][0], o[0])
        np.testing.assert_array_equal(outs[0][1], o[1])
    with pytest.raises(msf.MsfError):
        ctx.configure(fft_plan=5)


def test_mixed_lengths_follow_each_axis(msf, ctx, O):
    """A 512 x 256 grid: automatically the primary plan along x (512) and the
    alternate along y (256), within the oracle's tolerances."""
    q, Z = random_atoms(40, (12.0, 6.0, 6.0), seed=72, species=(14, 29))
    c = Case(nx=512, ny=256, cell=(12.0, 6.0, 6.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.5)
    xp = np.array([1.0, 9.5])
    yp = np.array([2.2, 5.9])
    _, (img_o, v_o, _) = c.run_oracle(O, xp, yp)
    spec = c.gpu_specimen(msf, ctx)
    img, v = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                         fem=c.femcfg(msf))
    assert _plans(ctx) == (2, 1)
    assert rel_l2(img, img_o) < 1e-4
    assert rel_l2(v, v_o) < 1e-4


@pytest.mark.parametrize("fp,plans", [(3, (2, 1)), (4, (1, 2))])
def test_per_axis_plans(msf, ctx, O, fp, plans):
    """fft_plan 3 (alternate rows, primary columns) and 4 (the reverse) at
    1024^2, within the oracle's tolerances; the mixes differ from the
    all-alternate plan by FP32 rounding only."""
    q, Z = random_atoms(30, (40.0, 40.0, 3.0), seed=74, species=(29,))
    c = Case(nx=1024, ny=1024, cell=(40.0, 40.0, 3.0), voltage=200e3, q=q, Z=Z, min_slice=2.0,
             alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.5)
    xp = np.array([5.0, 31.3])
    yp = np.array([20.0])
    _, (img_o, v_o, _) = c.run_oracle(O, xp, yp)
    spec = c.gpu_specimen(msf, ctx)
    try:
        ctx.configure(fft_plan=fp)
        img, v = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                             fem=c.femcfg(msf))
        assert _plans(ctx) == plans
        ctx.configure(fft_plan=2)
        img2, v2 = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                               fem=c.femcfg(msf))
    finally:
        ctx.configure()
        spec.free()
    assert rel_l2(img, img_o) < 1e-4
    assert rel_l2(v, v_o) < 1e-4
    assert rel_l2(img, img2) < 1e-6


def test_autotuned_plan_reproduces_when_set(msf, ctx):
    c = _case(256, 71)
    xp = np.linspace(0.0, 9.0, 8)
    yp = np.linspace(0.3, 9.3, 6)
    spec = c.gpu_specimen(msf, ctx)
    run = lambda: ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                              fem=c.femcfg(msf))
    try:
        ctx.configure(fft_plan=1)
        img1, v1 = run()
        msf.tuning_clear()
        ctx.configure(autotune=True)
        img_t, v_t = run()
        st = ctx.stats()
        print("tuned:", st["batch"], st["col_tile"], st["chains"], _plans(ctx), flush=True)
        assert st["autotuned"] == 1 and st["row_fft_plan"] == st["col_fft_plan"]
        ctx.configure(batch=st["batch"], col_tile=st["col_tile"], chains=st["chains"],
                      fft_plan=st["row_fft_plan"])
        img_e, v_e = run()
    finally:
        ctx.configure()
    np.testing.assert_array_equal(img_e, img_t)
    np.testing.assert_array_equal(v_e, v_t)
    assert rel_l2(img_t, img1) < 1e-5 and rel_l2(v_t, v1) < 1e-5
