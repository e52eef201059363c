"""GPU parity of the detector MTF (adfstem.cpp:238-453, :2272-2385): the 2-D
MTF and its application on given waves (FP64 entry points) against the
oracle, and whole rasters whose FEM statistics see the MTF-filtered waves
(FP32 slice loop) within 1e-4; results independent of batch depth and graph
replay, and the run plan rebuilt when the table changes."""
import numpy as np
import pytest

from helpers import Case, random_atoms, rel_l2

pytestmark = pytest.mark.gpu

DOMAIN = np.linspace(0.0, 1.0, 11)
TABLE = 1.0 / (1.0 + (DOMAIN / 0.4) ** 2)


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _case(rate=12.0, table=TABLE):
    q, Z = random_atoms(70, (12.0, 12.0, 8.0), seed=43, species=(14, 29, 40))
    return Case(nx=128, ny=128, cell=(12.0, 12.0, 8.0), voltage=200e3, q=q, Z=Z, min_slice=1.5,
                alpha_max=0.012, cs3=5e5, defocus=30.0, inner=0.02, outer=0.15, fem=True, dr=0.5,
                mtf=table, mtf_domain=DOMAIN, mtf_rate=rate)


def test_mtf_entry_points_match_oracle(msf, ctx, O):
    c = _case()
    job = c.oracle_job(O)
    g = c.grid(msf)
    fem = c.femcfg(msf)
    m_o = O.mtf_2d(job, TABLE, DOMAIN, 12.0)
    m_g = ctx.mtf_2d(g, fem)
    print("mtf2d", rel_l2(m_g, m_o), flush=True)
    assert rel_l2(m_g, m_o) < 1e-12
    assert np.abs(m_o).max() > 0.1
    _, (_, _, ex) = c.run_oracle(O, np.array([1.0, 7.5]), np.array([2.0]), exit_waves=True)
    w_o = np.stack([O.apply_mtf(job, e, m_o) for e in ex])
    w_g = ctx.apply_mtf(g, fem, c.alpha_max, ex)
    print("apply", rel_l2(w_g, w_o), flush=True)
    assert rel_l2(w_g, w_o) < 1e-11
    with pytest.raises(msf.MsfError):
        ctx.apply_mtf(g, msf.make_fem(0.5, TABLE, DOMAIN, 0.0), c.alpha_max, ex)


def test_fem_with_mtf_matches_oracle(msf, ctx, O):
    c = _case()
    xp = np.linspace(0.3, 11.2, 4)
    yp = np.linspace(0.8, 11.6, 5)
    job = c.oracle_job(O)
    img_o, v_o = O.adfstem_fem(job, O.Specimen(job, c.q, c.Z.astype(np.int32)), xp, yp,
                               correlograph=True)
    spec = c.gpu_specimen(msf, ctx)
    img_g, v_g = ctx.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf),
                                 detector=c.detector(msf), correlograph=True)
    assert rel_l2(img_g, img_o) < 1e-4           # the ADF is taken before the MTF
    for m in msf.FEM_MODES + ("correlograph_avg", "correlograph_sqr_sum"):
        err = rel_l2(v_g[m], v_o[m])
        print(m, err, flush=True)
        assert err < 1e-4, (m, err)
    # the MTF changes the statistics: the same run without it differs
    c0 = Case(**{**c.__dict__, "mtf": None, "mtf_domain": None, "mtf_rate": 0.0})
    _, v0 = ctx.adfstem_fem(spec, c0.optics(msf), xp, yp, c0.femcfg(msf))
    assert rel_l2(v0["v_omega"], v_g["v_omega"]) > 1e-2


def test_mtf_runs_do_not_depend_on_batch_graph_or_cache(msf, ctx):
    c = _case()
    c2 = _case(rate=9.0, table=TABLE ** 2)
    xp = np.linspace(0.0, 11.0, 6)
    yp = np.linspace(0.5, 10.5, 3)
    spec = c.gpu_specimen(msf, ctx)
    run = lambda cc: ctx.adfstem_fem(spec, cc.optics(msf), xp, yp, cc.femcfg(msf),
                                     modes=("v_omega", "v_re"))[1]
    first = run(c)
    other = run(c2)                               # a different table rebuilds the plan
    assert rel_l2(other["v_omega"], first["v_omega"]) > 1e-3
    again = run(c)
    outs = [again]
    try:
        for b, graph, ch in ((1, False, 1), (7, True, 1), (7, False, 1), (7, True, 2), (5, False, 2)):
            ctx.configure(batch=b, graph=graph, chains=ch)
            outs.append(run(c))
    finally:
        ctx.configure()
    for o in outs:
        for m in o:
            np.testing.assert_array_equal(o[m], first[m], err_msg=m)
