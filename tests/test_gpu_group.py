"""Grouped launches (msf_run_group): runs on specimens of one grid size share
their slice-pass launches, each run keeping its own specimen, vacuum waves,
kept columns, batches and raster-order statistics.  Every run's ADF image
and FEM outputs must equal those of the same run on its own, bit for bit."""
import numpy as np
import pytest

from helpers import random_atoms

pytestmark = pytest.mark.gpu

MODES = ("v_omega", "vbar_r", "v_re")


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _runs(msf, ctx, n, count, seed):
    """`count` runs of grid n: cells 0.045-0.07 A per pixel (different kept
    columns), depths of 1-4 slices, rasters of 2x1 to 7x6 positions, ADF
    rings and FEM bin widths varied."""
    rng = np.random.default_rng(seed)
    runs = []
    for k in range(count):
        L = n * float(rng.uniform(0.045, 0.07))
        depth = 2.0 * int(rng.integers(1, 5))
        q, Z = random_atoms(int(0.02 * L * L * depth), (L, L, depth), seed=seed + k,
                            species=(14, 29))
        spec = ctx.specimen(msf.make_grid(n, n, L, L, depth, 300e3), q, Z, 2.0)
        nxp, nyp = int(rng.integers(2, 8)), int(rng.integers(1, 7))
        runs.append(dict(spec=spec, optics=msf.Optics(0, 0.0, 0.0, float(rng.uniform(-20, 20)),
                                                      float(rng.uniform(0.012, 0.025))),
                         x_p=np.linspace(0.1, L - 0.3, nxp), y_p=np.linspace(0.2, L - 0.5, nyp),
                         detector=msf.Detector(1, float(rng.uniform(0.03, 0.05)), 0.11, 0),
                         fem=msf.Fem(1, float(rng.uniform(0.4, 0.8)))))
    return runs


def _alone(ctx, r):
    img, fem = ctx.adfstem_fem(r["spec"], r["optics"], r["x_p"], r["y_p"], r["fem"],
                               modes=MODES, detector=r["detector"])
    return img, fem


def _check_equal(alone, grouped):
    img_a, fem_a = alone
    img_g, fem_g = grouped
    np.testing.assert_array_equal(img_g, img_a)
    for m in MODES:
        np.testing.assert_array_equal(fem_g[m], fem_a[m])


@pytest.mark.parametrize("n,count,batch", [(128, 5, 4), (256, 4, 0), (512, 3, 8), (64, 19, 3)])
def test_group_equals_runs_alone(msf, ctx, n, count, batch):
    """Grouped runs of one grid size (19 at 64^2: more runs than one launch
    takes) give each run's results bit for bit, with several batches per run
    (batch 3, 4, 8) or one (batch 0: the rule)."""
    runs = _runs(msf, ctx, n, count, seed=100 * n + count)
    try:
        ctx.configure(batch=batch)
        alone = [_alone(ctx, r) for r in runs]
        print(f"n {n}: {count} runs alone done", flush=True)
        grouped, status = ctx.run_group(runs, modes=MODES)
        st = ctx.stats()
        assert status == [0] * count
        assert st["group_runs"] == count
        assert st["group_launches"] > 0
        for a, g in zip(alone, grouped):
            _check_equal(a, g)
        assert all(np.isfinite(g[0]).all() for g in grouped)
        # a second group on the same context reuses the run plans
        grouped2, status2 = ctx.run_group(runs[::-1], modes=MODES)
        assert status2 == [0] * count
        for a, g in zip(alone[::-1], grouped2):
            _check_equal(a, g)
    finally:
        ctx.configure()
        for r in runs:
            r["spec"].free()


def test_group_reports_failing_runs(msf, ctx):
    """A run whose detector lies beyond the band limit, and a run on another
    grid size, fail alone; the other runs of the group still finish with
    their own results."""
    runs = _runs(msf, ctx, 128, 3, seed=77)
    other = _runs(msf, ctx, 64, 1, seed=78)[0]
    bad = dict(runs[1])
    bad["detector"] = msf.Detector(1, 0.9, 1.0, 0)
    try:
        alone = [_alone(ctx, r) for r in (runs[0], runs[2])]
        grouped, status = ctx.run_group([runs[0], bad, other, runs[2]], modes=MODES)
        assert status[0] == 0 and status[3] == 0
        assert status[1] == -1 and status[2] == -1
        _check_equal(alone[0], grouped[0])
        _check_equal(alone[1], grouped[3])
    finally:
        for r in runs + [other]:
            r["spec"].free()


def test_group_with_every_fem_output(msf, ctx):
    """Grouped runs with the other FEM modes, the correlograph (per position
    too), the per-pixel images and the detector MTF, equal to each run alone."""
    runs = _runs(msf, ctx, 128, 3, seed=91)
    dom = np.linspace(0.0, 1.0, 11)
    mtf = 1.0 / (1.0 + (dom / 0.4) ** 2)
    extra = dict(modes=msf.FEM_MODES, correlograph=True, correlograph_each=True,
                 pixel_images=True)
    try:
        ctx.configure(batch=4)
        for r in runs:
            r.update(extra)
        runs[1]["fem"] = msf.make_fem(runs[1]["fem"].dr, mtf, dom, 12.0)
        alone = [ctx.adfstem_fem(r["spec"], r["optics"], r["x_p"], r["y_p"], r["fem"],
                                 modes=msf.FEM_MODES, detector=r["detector"], correlograph=True,
                                 correlograph_each=True, pixel_images=True) for r in runs]
        grouped, status = ctx.run_group(runs)
        assert status == [0, 0, 0]
        for (img_a, out_a), (img_g, out_g) in zip(alone, grouped):
            np.testing.assert_array_equal(img_g, img_a)
            assert sorted(out_a) == sorted(out_g)
            for k in out_a:
                np.testing.assert_array_equal(out_g[k], out_a[k])
    finally:
        ctx.configure()
        for r in runs:
            r["spec"].free()
