"""The serving queue (msf_submit / msf_poll / msf_wait): jobs of mixed sizes
submitted together, each equal bit for bit to the same job run directly and
within 1e-4 of the FP64 oracle; a failing job does not stop the others."""
import numpy as np
import pytest

from helpers import Case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _jobs(n=6, seed=3000):
    """A small mixed workload in the manner of a batched submission: N in
    {64, 96, 128}, cell side L with N / L >= 8 (so a 0.05-0.18 rad ADF ring
    fits inside the band limit at 300 kV), Si at 0.05 A^-3, minslice 2.0,
    explicit rasters of side 3..5 at spacing L / side."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n):
        N = int(rng.choice([64, 96, 128]))
        L = float(rng.uniform(N / 12.0, N / 8.0))
        n_at = max(1, int(round(0.05 * L * L * 4.0)))
        q = rng.random((n_at, 3)) * [L, L, 4.0]
        side = int(rng.integers(3, 6))
        xp = np.arange(side) * (L / side)
        c = Case(nx=N, ny=N, cell=(L, L, 4.0), voltage=300e3, q=q,
                 Z=np.full(n_at, 14, np.uint32), min_slice=2.0, alpha_max=0.015,
                 inner=0.05, outer=0.18, fem=True, dr=0.7)
        jobs.append((c, xp))
    return jobs


def test_submitted_jobs_match_direct_runs_and_oracle(msf, ctx, O):
    jobs = _jobs()
    tickets = []
    for c, xp in jobs:
        tickets.append(ctx.submit(c.grid(msf), c.q, c.Z, c.min_slice, c.optics(msf), xp, xp,
                                  detector=c.detector(msf), fem=c.femcfg(msf),
                                  modes=("v_omega", "v_re")))
    results = [None] * len(jobs)
    for k in reversed(range(len(jobs))):          # waiting order does not matter
        results[k] = ctx.wait(tickets[k])
        assert ctx.poll(tickets[k]) == -1          # forgotten after wait
    for (c, xp), (img, fem) in zip(jobs, results):
        spec = c.gpu_specimen(msf, ctx)
        img_d, fem_d = ctx.adfstem_fem(spec, c.optics(msf), xp, xp, c.femcfg(msf),
                                       modes=("v_omega", "v_re"), detector=c.detector(msf))
        np.testing.assert_array_equal(img, img_d)
        for m in fem:
            np.testing.assert_array_equal(fem[m], fem_d[m])
        _, (img_o, v_o, _) = c.run_oracle(O, xp, xp)
        assert rel_l2(img, img_o) < 1e-4
        assert rel_l2(fem["v_omega"], v_o) < 1e-4
        spec.free()


def test_failing_job_does_not_stop_the_queue(msf, ctx):
    (c, xp), (c2, xp2) = _jobs(2, seed=7)
    bad = ctx.submit(c.grid(msf), c.q, np.full(c.Z.size, 200, np.uint32), c.min_slice,
                     c.optics(msf), xp, xp, detector=c.detector(msf))
    good = ctx.submit(c2.grid(msf), c2.q, c2.Z, c2.min_slice, c2.optics(msf), xp2, xp2,
                      detector=c2.detector(msf))
    assert ctx.poll(good) in (0, msf.PENDING)
    with pytest.raises(msf.MsfError) as e:
        ctx.wait(bad)
    assert e.value.code == -1 and "atomic number" in str(e.value)
    img, _ = ctx.wait(good)
    assert np.isfinite(img).all() and img.max() > 0


def test_destroy_finishes_the_queue(msf):
    (c, xp), = _jobs(1, seed=11)
    own = msf.Context(0)
    t = own.submit(c.grid(msf), c.q, c.Z, c.min_slice, c.optics(msf), xp, xp,
                   detector=c.detector(msf))
    own.close()                                   # joins the worker after the job ran
    assert np.isfinite(t.img).all() and t.img.max() > 0


def _submit_all(ctx, msf, jobs):
    return [ctx.submit(c.grid(msf), c.q, c.Z, c.min_slice, c.optics(msf), xp, xp,
                       detector=c.detector(msf), fem=c.femcfg(msf), modes=("v_omega", "v_re"))
            for c, xp in jobs]


def test_continuous_batching_overlaps_jobs_and_matches_alone_runs(msf):
    """Jobs submitted together run at the same time on several slots, and
    each result is bit-identical to the same job served alone."""
    jobs = _jobs(9, seed=3100)
    mix = msf.Context(0)
    alone = msf.Context(0)
    try:
        mix.serve_configure(slots=3)
        alone.serve_configure(slots=1)
        got = [mix.wait(t, info=True) for t in _submit_all(mix, msf, jobs)]
        st = mix.serve_stats()
        print("mixed:", st, flush=True)
        assert st["jobs_done"] == 9 and st["running"] == 0 and st["queued"] == 0
        assert st["max_running"] >= 2                      # the work overlapped
        assert len({info["slot"] for _, _, info in got}) >= 2
        for (c, xp), (img, fem, info) in zip(jobs, got):
            t = _submit_all(alone, msf, [(c, xp)])[0]
            img_a, fem_a = alone.wait(t)
            np.testing.assert_array_equal(img, img_a)
            for m in fem:
                np.testing.assert_array_equal(fem[m], fem_a[m])
            assert info["run_ms"] > 0 and info["queue_ms"] >= 0
        assert alone.serve_stats()["max_running"] == 1
    finally:
        mix.close()
        alone.close()


def test_admission_by_footprint(msf):
    """With a memory limit of about two jobs' footprints, at most two run at
    once and the admitted footprints never exceed the limit; with a limit
    below any job, jobs run one at a time."""
    jobs = _jobs(6, seed=3200)
    fps = [msf.job_footprint(c.grid(msf), c.q, c.Z, c.min_slice, xp, xp, fem=c.femcfg(msf))
           for c, xp in jobs]
    assert min(fps) > 0
    limit = int(2.05 * max(fps))
    ctx2 = msf.Context(0)
    ctx1 = msf.Context(0)
    try:
        ctx2.serve_configure(slots=4, mem_limit=limit)
        ts = _submit_all(ctx2, msf, jobs)
        assert [t.footprint for t in ts] == fps
        for t in ts:
            ctx2.wait(t)
        st = ctx2.serve_stats()
        print("limited:", st, flush=True)
        assert st["max_running"] <= 2 and st["peak_admitted_bytes"] <= limit
        ctx1.serve_configure(slots=4, mem_limit=1)
        for t in _submit_all(ctx1, msf, jobs[:3]):
            img, _ = ctx1.wait(t)
            assert np.isfinite(img).all()
        st1 = ctx1.serve_stats()
        assert st1["max_running"] == 1 and st1["jobs_done"] == 3
        with pytest.raises(msf.MsfError):
            ctx1.serve_configure(slots=2)               # only before the first submission
    finally:
        ctx2.close()
        ctx1.close()


def test_runs_on_a_resident_specimen_share_one_build(msf):
    """Four served runs on one resident specimen (other optics, rasters and
    detectors each): the engine builds no specimen for them, and each result
    equals the same job submitted with its atoms (a build per job) and a
    direct run, bit for bit.  Freeing the specimen while runs on it are
    queued waits for them."""
    (c, xp), = _jobs(1, seed=3300)
    eng = msf.Context(0)
    try:
        eng.serve_configure(slots=3)
        spec = c.gpu_specimen(msf, eng)
        fem = c.femcfg(msf)
        variants = [(c.optics(msf), xp, xp, c.detector(msf)),
                    (msf.Optics(0, 0.0, 0.0, 30.0, 0.012), xp[:2], xp, c.detector(msf)),
                    (c.optics(msf), xp, xp[::-1].copy(), msf.Detector(1, 0.04, 0.15, 0)),
                    (msf.Optics(0, 1.0e5, 0.0, -20.0, 0.018), xp, xp[1:], c.detector(msf))]
        runs = [eng.submit_run(spec, o, x, y, detector=d, fem=fem) for o, x, y, d in variants]
        got = [eng.wait(t) for t in runs]
        st = eng.serve_stats()
        print("resident:", st, flush=True)
        assert st["specimen_builds"] == 0 and st["jobs_done"] == 4 and st["resident_bytes"] == 0
        own = [eng.submit(c.grid(msf), c.q, c.Z, c.min_slice, o, x, y, detector=d, fem=fem)
               for o, x, y, d in variants]
        got_own = [eng.wait(t) for t in own]
        assert eng.serve_stats()["specimen_builds"] == 4
        for (o, x, y, d), (img, v), (img2, v2) in zip(variants, got, got_own):
            np.testing.assert_array_equal(img, img2)
            np.testing.assert_array_equal(v["v_omega"], v2["v_omega"])
            img3, v3 = eng.adfstem(spec, o, x, y, detector=d, fem=fem)
            np.testing.assert_array_equal(img, img3)
            np.testing.assert_array_equal(v["v_omega"], v3)
        # runs still queued when the specimen is freed: the free waits for them
        late = [eng.submit_run(spec, o, x, y, detector=d, fem=fem) for o, x, y, d in variants[:2]]
        spec.free()
        for t, (img, v) in zip(late, got[:2]):
            img4, v4 = eng.wait(t)
            np.testing.assert_array_equal(img4, img)
            np.testing.assert_array_equal(v4["v_omega"], v["v_omega"])
    finally:
        eng.close()


def test_grouped_jobs_match_direct_runs(msf):
    """With grouping (msf_serve_group), queued jobs of one grid size run as a
    group with shared slice-pass launches; every job, built from atoms or on a
    resident specimen, equals the same run made directly, bit for bit, and a
    job with bad atoms fails alone."""
    own = msf.Context(0)
    try:
        own.serve_configure(slots=1)
        own.serve_group(8)
        # a larger head job keeps the one slot busy while the rest queue up
        (c0, _), = _jobs(1, seed=21)
        head_xp = np.linspace(0.0, c0.cell[0] * 0.9, 12)
        jobs = _jobs(9, seed=22)
        spec_r = jobs[0][0].gpu_specimen(msf, own)
        tickets = [own.submit(c0.grid(msf), c0.q, c0.Z, c0.min_slice, c0.optics(msf), head_xp,
                              head_xp, detector=c0.detector(msf), fem=c0.femcfg(msf))]
        for c, xp in jobs:
            tickets.append(own.submit(c.grid(msf), c.q, c.Z, c.min_slice, c.optics(msf), xp, xp,
                                      detector=c.detector(msf), fem=c.femcfg(msf),
                                      modes=("v_omega", "v_re")))
        c, xp = jobs[0]
        t_res = own.submit_run(spec_r, c.optics(msf), xp[::-1].copy(), xp,
                               detector=c.detector(msf), fem=c.femcfg(msf))
        bad = own.submit(c.grid(msf), c.q, np.full(c.Z.size, 200, np.uint32), c.min_slice,
                         c.optics(msf), xp, xp, detector=c.detector(msf))
        results = [own.wait(t) for t in tickets[1:]]
        img_res, fem_res = own.wait(t_res)
        with pytest.raises(msf.MsfError):
            own.wait(bad)
        own.wait(tickets[0])
        st = own.serve_stats()
        print("groups", st["groups"], "grouped jobs", st["grouped_jobs"], "launches",
              st["group_launches"], flush=True)
        assert st["groups"] >= 1 and st["grouped_jobs"] >= 2 and st["group_launches"] > 0
        for (c, xp), (img, fem) in zip(jobs, results):
            spec = c.gpu_specimen(msf, own)
            img_d, fem_d = own.adfstem_fem(spec, c.optics(msf), xp, xp, c.femcfg(msf),
                                           modes=("v_omega", "v_re"), detector=c.detector(msf))
            np.testing.assert_array_equal(img, img_d)
            for m in fem:
                np.testing.assert_array_equal(fem[m], fem_d[m])
            spec.free()
        c, xp = jobs[0]
        img_d, fem_d = own.adfstem_fem(spec_r, c.optics(msf), xp[::-1].copy(), xp, c.femcfg(msf),
                                       modes=("v_omega",), detector=c.detector(msf))
        np.testing.assert_array_equal(img_res, img_d)
        np.testing.assert_array_equal(fem_res["v_omega"], fem_d["v_omega"])
        spec_r.free()
    finally:
        own.close()
