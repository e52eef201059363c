"""GPU parity over the remaining transform lengths, optics, detector modes and
edge cases of the raster, against the FP64 oracle (same tolerances as
test_gpu_parity.py: exit intensity 1e-5, ADF / V_omega 1e-4 relative L2)."""
import numpy as np
import pytest

from helpers import Case, random_atoms, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _exit_parity(msf, ctx, O, c, xp, yp, batch=0, tol=1e-5):
    _, (_, _, ex_o) = c.run_oracle(O, xp, yp, exit_waves=True)
    ctx.configure(batch=batch, transmit_mode=c.transmit_mode)
    spec_g = c.gpu_specimen(msf, ctx)
    ex_g = ctx.exit_waves(spec_g, c.optics(msf), xp, yp)
    ctx.configure()
    errs = [rel_l2(np.abs(ex_g[p]) ** 2, np.abs(ex_o[p]) ** 2) for p in range(ex_o.shape[0])]
    assert max(errs) < tol, errs
    # the band limit leaves nothing outside k^2 < bwcutoff_t^2
    assert np.count_nonzero(ex_g[0]) <= np.count_nonzero(np.abs(ex_o[0]) > 0)


@pytest.mark.parametrize("n", [16, 32, 96, 192, 256, 384, 768])
def test_exit_waves_other_lengths(msf, ctx, O, n):
    """The plans with radix-6 first stages (96, 192, 384, 768; 12 or 24
    values per lane, column passes bounded by their tile size), the 256-point
    plan and the smallest lengths (16 and 32: 2 and 4 lanes per transform)."""
    cell = (max(n / 8.0, 6.0), max(n / 8.0, 6.0), 6.0)
    q, Z = random_atoms(50 if n > 32 else 8, cell, seed=n, species=(14, 40))
    c = Case(nx=n, ny=n, cell=cell, voltage=300e3, q=q, Z=Z, min_slice=2.0, alpha_max=0.015)
    _exit_parity(msf, ctx, O, c, np.array([0.3, cell[0] / 2 + 0.11]), np.array([1.7]), batch=2)


def test_exit_waves_1024(msf, ctx, O):
    """1024^2 with the automatic plan: 16 x 16 x 4 on 64 lanes, lean row
    passes and 4-column lean tiles of the column pass."""
    cell = (40.0, 40.0, 3.0)
    q, Z = random_atoms(30, cell, seed=11, species=(29,))
    c = Case(nx=1024, ny=1024, cell=cell, voltage=200e3, q=q, Z=Z, min_slice=2.0,
             alpha_max=0.02)
    _exit_parity(msf, ctx, O, c, np.array([5.0, 31.3]), np.array([20.0]))


@pytest.mark.parametrize("n", [96, 384, 512, 1024])
def test_shifts_at_the_row_pair_edges(msf, ctx, O, n):
    """The row pass reads the transmission and u0 rows in pairs (elements j
    and j + n/2 of a lane in one read, u0 at r = (j - sy) mod n and its
    partner (r + n/2) mod n): probe shifts sy of 0, 1, n/2 - 1, n/2, n/2 + 1
    and n - 2 (the largest the shift rule yields) put the wrap of both at
    every place, on the row passes that interleave reads and products (96,
    384) and those that read ahead (512 primary, 1024 alternate); exit
    intensities within 1e-5 of the oracle."""
    L = n * 0.05
    cell = (L, L, 3.0)
    q, Z = random_atoms(30, cell, seed=300 + n, species=(14, 29))
    c = Case(nx=n, ny=n, cell=cell, voltage=300e3, q=q, Z=Z, min_slice=2.0, alpha_max=0.02)
    d = L / n
    want = [0, 1, n // 2 - 1, n // 2, n // 2 + 1, n - 2]
    yp = d * np.array(want, float)
    xp = np.array([d * (n // 2 + 3)])
    sh = msf.probe_shifts(c.grid(msf), xp, yp)
    assert list(sh[:, 1]) == want and list(sh[:, 0]) == [n // 2 + 3] * len(want)
    print("grid", n, "shifts", want, flush=True)
    _exit_parity(msf, ctx, O, c, xp, yp, batch=4)


def test_non_square_lengths(msf, ctx, O):
    """nx != ny with different plans on the two axes (192 x 96)."""
    cell = (18.0, 9.0, 4.0)
    q, Z = random_atoms(25, cell, seed=5, species=(14,))
    c = Case(nx=192, ny=96, cell=cell, voltage=300e3, q=q, Z=Z, min_slice=1.5, alpha_max=0.02)
    _exit_parity(msf, ctx, O, c, np.array([2.0, 17.0]), np.array([4.4]))


def test_corrected_probe_full_run(msf, ctx, O):
    """adfstemcorr: Cs3 + Cs5 aberration function in the probe, ADF + FEM."""
    q, Z = random_atoms(60, (12.0, 12.0, 8.0), seed=9, species=(14, 29))
    c = Case(nx=128, ny=128, cell=(12.0, 12.0, 8.0), voltage=200e3, q=q, Z=Z, min_slice=1.5,
             corrected=1, cs3=-1.5e5, cs5=4.0e7, defocus=-12.0, alpha_max=0.025, inner=0.03,
             outer=0.18, fem=True, dr=0.5)
    xp = np.linspace(0.2, 11.0, 4)
    yp = np.linspace(0.9, 11.3, 3)
    _, (img_o, v_o, _) = c.run_oracle(O, xp, yp)
    spec_g = c.gpu_specimen(msf, ctx)
    img_g, v_g = ctx.adfstem(spec_g, c.optics(msf), xp, yp, detector=c.detector(msf),
                             fem=c.femcfg(msf))
    assert rel_l2(img_g, img_o) < 1e-4
    assert rel_l2(v_g, v_o) < 1e-4


def test_physical_detector_mode(msf, ctx, O):
    """MSF_ADF_PHYSICAL squares the thresholds; the ring differs from the
    reference's k^2-vs-k ring and both match the oracle in their mode."""
    q, Z = random_atoms(60, (12.0, 12.0, 8.0), seed=21, species=(29,))
    imgs = {}
    for mode in (0, 1):
        c = Case(nx=128, ny=128, cell=(12.0, 12.0, 8.0), voltage=300e3, q=q, Z=Z,
                 min_slice=1.5, alpha_max=0.02, inner=0.04, outer=0.2, adf_mode=mode)
        xp = np.array([1.0, 6.0, 10.5])
        yp = np.array([2.0, 8.0])
        _, (img_o, _, _) = c.run_oracle(O, xp, yp)
        spec_g = c.gpu_specimen(msf, ctx)
        img_g, _ = ctx.adfstem(spec_g, c.optics(msf), xp, yp, detector=c.detector(msf))
        assert rel_l2(img_g, img_o) < 1e-4, mode
        imgs[mode] = img_g
    assert rel_l2(imgs[0], imgs[1]) > 1e-2


def test_si256_geometry_subset(msf, ctx, O):
    """The 256^2 Si job (400 atoms in a 20 A cube, 300 kV, 10 mrad, ADF
    40-200 mrad, minslice 2.0) on a 4 x 4 subset of its central raster."""
    import specimens as SP
    job, q, Z, xp, yp = SP.config_si_256(seed=1001)
    c = Case(nx=256, ny=256, cell=(20.0, 20.0, 20.0), voltage=job["voltage"], q=q, Z=Z,
             min_slice=job["min_slice"], cs3=job["cs3"], defocus=job["defocus"],
             alpha_max=job["alpha_max"], inner=job["det_inner"], outer=job["det_outer"])
    sub_x, sub_y = xp[::4], yp[5:9]
    _, (img_o, _, _) = c.run_oracle(O, sub_x, sub_y)
    spec_g = c.gpu_specimen(msf, ctx)
    assert 8 <= spec_g.n_slices <= 12
    img_g, _ = ctx.adfstem(spec_g, c.optics(msf), sub_x, sub_y, detector=c.detector(msf))
    assert rel_l2(img_g, img_o) < 1e-4


def test_positions_beyond_last_pixel_keep_previous_shift(msf, ctx, O):
    """No grid index qualifies for x >= x[N-1] + dx/2: the previous position's
    shift persists (the raster loop never resets it)."""
    q, Z = random_atoms(40, (12.0, 12.0, 6.0), seed=13, species=(14,))
    c = Case(nx=64, ny=64, cell=(12.0, 12.0, 6.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02)
    dx = 12.0 / 64
    xp = np.array([3.0, 12.0 - dx / 4, 7.0])   # x1 keeps the x-shift of x0
    yp = np.array([11.99, 4.0])                 # y0 keeps the previous y-shift
    _exit_parity(msf, ctx, O, c, xp, yp)
    spec_g = c.gpu_specimen(msf, ctx)
    ex = ctx.exit_waves(spec_g, c.optics(msf), xp, yp)
    # raster order is x outer, y inner: (x1, y0) inherits both shifts of
    # (x0, y1), and (x1, y1) keeps the x-shift of x0; all three waves agree
    np.testing.assert_array_equal(ex[1], ex[2])
    np.testing.assert_array_equal(ex[1], ex[3])


def test_empty_specimen_is_free_propagation(msf, ctx, O):
    """No atoms: one slice of zperiod, t = 1, exit wave = M P psi0."""
    c = Case(nx=64, ny=64, cell=(10.0, 10.0, 5.0), voltage=300e3,
             q=np.zeros((0, 3)), Z=np.zeros(0, np.uint32), min_slice=1.0, alpha_max=0.02)
    spec_g = c.gpu_specimen(msf, ctx)
    assert spec_g.n_slices == 1
    lb, th, na = spec_g.slice_info()
    assert th[0] == 5.0 and na[0] == 0
    _exit_parity(msf, ctx, O, c, np.array([1.0, 4.2]), np.array([3.3]))


def test_error_paths(msf, ctx):
    q, Z = random_atoms(10, (8.0, 8.0, 4.0), seed=1)
    c = Case(nx=64, ny=64, cell=(8.0, 8.0, 4.0), voltage=300e3, q=q, Z=Z)
    spec = c.gpu_specimen(msf, ctx)
    # inner angle beyond the band limit (tem.cpp:721-732)
    bad = msf.Detector(1, 10.0, 20.0, 0)
    with pytest.raises(msf.MsfError) as e:
        ctx.adfstem(spec, c.optics(msf), np.array([1.0]), np.array([1.0]), detector=bad)
    assert e.value.code == -1
    # negative detector angle (tem.cpp:710-720)
    with pytest.raises(msf.MsfError) as e:
        ctx.adfstem(spec, c.optics(msf), np.array([1.0]), np.array([1.0]),
                    detector=msf.Detector(1, -0.01, 0.2, 0))
    assert e.value.code == -1
    # a position below the domain: the reference's size_t index would wrap
    with pytest.raises(msf.MsfError) as e:
        ctx.adfstem(spec, c.optics(msf), np.array([-1.0]), np.array([1.0]),
                    detector=c.detector(msf))
    assert e.value.code == -5
    # atomic number outside the parameter table (tem_io.cpp:4768-4781)
    with pytest.raises(msf.MsfError) as e:
        ctx.specimen(c.grid(msf), q, np.full(10, 104, np.uint32), 1.0)
    assert e.value.code == -1
    # a transform length that is not compiled in
    g = msf.make_grid(100, 100, 8.0, 8.0, 4.0, 300e3)
    with pytest.raises(msf.MsfError) as e:
        ctx.specimen(g, q, Z, 1.0)
    assert e.value.code == -2
    # the context still works after the failures
    img, _ = ctx.adfstem(spec, c.optics(msf), np.array([1.0]), np.array([1.0]),
                         detector=c.detector(msf))
    assert np.isfinite(img).all()


def test_graph_replay_matches_stream_launches(msf, ctx):
    """The captured job graph gives bit-identical results on every replay,
    with eager launches and new positions of the same count in between."""
    q, Z = random_atoms(60, (12.0, 12.0, 8.0), seed=17, species=(14, 29))
    c = Case(nx=64, ny=64, cell=(12.0, 12.0, 8.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02, inner=0.02, outer=0.2, fem=True, dr=0.7)
    spec = c.gpu_specimen(msf, ctx)
    rasters = [(np.linspace(0.0, 11.0, 6), np.linspace(0.5, 11.5, 5)),
               (np.linspace(1.3, 10.0, 6), np.linspace(0.1, 9.5, 5))]

    def run(xp, yp, graph, chains=1):
        ctx.configure(batch=8, graph=graph, chains=chains)
        out = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                          fem=c.femcfg(msf))
        return out, ctx.stats()["graph"]

    try:
        for rep in range(3):
            for xp, yp in rasters:
                (img_e, v_e), st_e = run(xp, yp, graph=False)
                assert st_e == 0
                (img_g, v_g), st_g = run(xp, yp, graph=True)
                assert st_g == (1 if rep == 0 and xp is rasters[0][0] else 2)
                np.testing.assert_array_equal(img_g, img_e)
                np.testing.assert_array_equal(v_g, v_e)
        # two batch chains in flight (a graph with two branches), replayed
        for rep in range(3):
            for xp, yp in rasters:
                (img_e, v_e), _ = run(xp, yp, graph=False)
                (img_2, v_2), st_2 = run(xp, yp, graph=True, chains=2)
                assert st_2 == (1 if rep == 0 and xp is rasters[0][0] else 2)
                np.testing.assert_array_equal(img_2, img_e)
                np.testing.assert_array_equal(v_2, v_e)
                (img_3, v_3), st_3 = run(xp, yp, graph=False, chains=2)
                assert st_3 == 0
                np.testing.assert_array_equal(img_3, img_e)
                np.testing.assert_array_equal(v_3, v_e)
    finally:
        ctx.configure()


@pytest.mark.parametrize("n,depth", [(256, 6.0), (384, 6.0), (512, 6.0), (512, 2.0)])
def test_fused_pass_pairs_match_one_chain(msf, ctx, n, depth):
    """Batches in pairs, the row pass of one and the column pass of the other
    in one launch (chains 3), give results bit-identical to one chain: odd
    and even batch counts, a smaller last batch, batches below one row
    workgroup's positions (passes launched one after the other), one slice,
    graph replay and stream launches."""
    L = 0.05 * n
    q, Z = random_atoms(int(0.02 * L * L * depth), (L, L, depth), seed=n + int(depth),
                        species=(14, 29))
    c = Case(nx=n, ny=n, cell=(L, L, depth), voltage=300e3, q=q, Z=Z, min_slice=2.0,
             alpha_max=0.02, inner=0.04, outer=0.12, fem=True, dr=0.5)
    spec = c.gpu_specimen(msf, ctx)
    xp = np.linspace(0.3, L - 0.7, 8)
    yp = np.linspace(0.1, L - 0.4, 6)

    def run(**cfg):
        ctx.configure(autotune=-1, **cfg)
        out = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                          fem=c.femcfg(msf))
        return out, ctx.stats()

    try:
        for b in (16, 12, 20, 5):
            (img_1, v_1), _ = run(batch=b, chains=1, graph=False)
            for graph in (False, True, True):
                (img_3, v_3), st = run(batch=b, chains=3, graph=graph)
                assert st["chains"] == 3 and st["batch"] == b
                # positions per row workgroup: 16 at 256 (alternate plan), else 8
                assert (st["fused_launches"] > 0) == (b >= (16 if n == 256 else 8))
                np.testing.assert_array_equal(img_3, img_1)
                np.testing.assert_array_equal(v_3, v_1)
        # the raster in one batch: one chain
        (img_a, v_a), st = run(batch=48, chains=3)
        assert st["chains"] == 1
        np.testing.assert_array_equal(img_a, img_1)
        np.testing.assert_array_equal(v_a, v_1)
    finally:
        ctx.configure()


def test_exit_waves_for_explicit_shifts(msf, ctx):
    """msf_exit_waves_shifted on the first and last raster positions equals
    those positions of a whole-raster msf_exit_waves, bit for bit."""
    q, Z = random_atoms(40, (12.0, 12.0, 6.0), seed=23, species=(14, 29))
    c = Case(nx=64, ny=64, cell=(12.0, 12.0, 6.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02)
    xp = np.array([0.5, 4.0, 11.99])
    yp = np.array([2.0, 11.97])
    spec = c.gpu_specimen(msf, ctx)
    full = ctx.exit_waves(spec, c.optics(msf), xp, yp)
    sh = msf.probe_shifts(c.grid(msf), xp, yp)
    sel = [0, xp.size * yp.size - 1]
    part = ctx.exit_waves_shifted(spec, c.optics(msf), sh[sel])
    np.testing.assert_array_equal(part, full[sel])


def test_specimen_may_outlive_its_context(msf):
    """Destroying a context detaches its specimens: freeing one afterwards
    only releases its own memory, and queries on it fail cleanly."""
    q, Z = random_atoms(10, (8.0, 8.0, 4.0), seed=2)
    c = Case(nx=64, ny=64, cell=(8.0, 8.0, 4.0), voltage=300e3, q=q, Z=Z)
    own = msf.Context(0)
    spec = c.gpu_specimen(msf, own)
    img, _ = own.adfstem(spec, c.optics(msf), np.array([1.0, 2.0]), np.array([1.0]),
                         detector=msf.Detector(1, 0.02, 0.2, 0))
    assert np.isfinite(img).all()
    own.close()
    with pytest.raises(msf.MsfError):
        spec.transmission(0)
    spec.free()


@pytest.mark.parametrize("mode", [0, 1])
def test_single_step_entry_points_match_oracle(msf, ctx, O, mode):
    """msf_probe_k, msf_specimen_propagator and msf_slice_step_batch (FP64 on
    the device) against orc_initial_wave, the oracle's propagators and
    orc_slice_step, for every slice of a small specimen."""
    q, Z = random_atoms(30, (9.0, 9.0, 6.0), seed=29, species=(14, 40))
    c = Case(nx=64, ny=96, cell=(9.0, 12.0, 6.0), voltage=200e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02, cs3=4e5, defocus=15.0, transmit_mode=mode)
    job = c.oracle_job(O)
    spec_o = O.Specimen(job, c.q, c.Z.astype(np.int32))
    spec_g = c.gpu_specimen(msf, ctx)
    kx, ky, xx, yy = O.domains(64, 96, 9.0, 12.0)
    _, bwt = O.bwcutoffs(64, 96, 9.0, 12.0)
    probe = O.probe(job)
    shifts = [(0, 0), (7, 40), (63, 95)]
    w_g = np.stack([ctx.probe_k(c.grid(msf), c.optics(msf), sx, sy) for sx, sy in shifts])
    w_o = np.stack([O.initial_wave(probe, sx, sy) for sx, sy in shifts])
    assert rel_l2(w_g, w_o) < 1e-11
    ctx.configure(transmit_mode=mode)
    try:
        for n in range(spec_g.n_slices):
            px, py = spec_g.propagator(n)
            np.testing.assert_allclose(px, spec_o.px[n], rtol=0, atol=1e-15)
            np.testing.assert_allclose(py, spec_o.py[n], rtol=0, atol=1e-15)
            w_g = ctx.slice_step(spec_g, n, w_o)   # same input on both sides
            w_o = np.stack([O.slice_step(w, spec_o.t_stack[n], spec_o.px[n], spec_o.py[n], kx,
                                         ky, bwt, mode) for w in w_o])
            assert rel_l2(w_g, w_o) < 1e-10, n
    finally:
        ctx.configure()


def test_reduction_entry_points_match_oracle(msf, ctx, O):
    """msf_adf_reduce and msf_fem_accumulate (FP64 on the device) on random
    k-space waves against orc_adf, orc_fem_radial and the V_omega sums of the
    oracle, for both ADF threshold modes."""
    rng = np.random.default_rng(17)
    nx, ny, Lx, Ly = 64, 96, 9.0, 12.0
    g = msf.make_grid(nx, ny, Lx, Ly, 6.0, 200e3)
    w = rng.normal(size=(5, nx, ny)) + 1j * rng.normal(size=(5, nx, ny))
    kx, ky, _, _ = O.domains(nx, ny, Lx, Ly)
    lam = msf.wavelength(200e3)[0]
    for mode in (msf.ADF_REFERENCE, msf.ADF_PHYSICAL):
        det = msf.Detector(1, 0.03, 0.15, mode)
        k1, k2 = 0.03 / lam, 0.15 / lam
        if mode == msf.ADF_PHYSICAL:
            k1, k2 = k1 * k1, k2 * k2
        adf_g = ctx.adf_reduce(g, det, w)
        adf_o = np.array([O.adf(wi, kx, ky, k1, k2) for wi in w])
        assert (adf_o > 0).all()
        np.testing.assert_allclose(adf_g, adf_o, rtol=1e-13)
    fem = msf.Fem(1, 0.5)
    alpha = 0.01
    r_g, c_g, v_g = ctx.fem_accumulate(g, fem, alpha, w)
    _, bwt = O.bwcutoffs(nx, ny, Lx, Ly)
    nb, dk = O.fem_nbins(kx, ky, 0.5, bwt)
    ro = [O.fem_radial(wi, kx, ky, dk, nb, lam, alpha) for wi in w]
    r_o = np.stack([r for r, _ in ro])
    assert r_g.shape == (5, nb)
    np.testing.assert_array_equal(c_g, ro[0][1])
    np.testing.assert_allclose(r_g, r_o, rtol=1e-13, atol=0)
    assert (r_o[:, 0] == 0).all() and (r_o[:, -1] > 0).all()   # direct beam zeroed
    with np.errstate(divide="ignore", invalid="ignore"):
        s1 = (r_o / c_g).sum(0)
        s2 = (r_o ** 2 / (c_g * c_g)).sum(0)
    v_o = O.fem_finalize(s1, s2, 5)
    assert rel_l2(v_g, v_o) < 1e-9
    with pytest.raises(msf.MsfError):
        ctx.adf_reduce(g, msf.Detector(1, -0.1, 0.15, 0), w)


def test_autotuned_launch_shape_gives_identical_results(msf, ctx):
    """The autotuner times batch depths, column tiles and one or two chains
    on the job itself once per shape; whatever it picks, ADF and V_omega are
    bit-identical to the fixed rules and to explicit settings."""
    q, Z = random_atoms(80, (12.0, 12.0, 6.0), seed=23, species=(14, 29))
    c = Case(nx=128, ny=128, cell=(12.0, 12.0, 6.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.7)
    spec = c.gpu_specimen(msf, ctx)
    xp = np.linspace(0.0, 11.5, 12)
    yp = np.linspace(0.2, 11.0, 10)
    run = lambda: ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                              fem=c.femcfg(msf))
    try:
        msf.tuning_clear()                                # the table is process-wide
        ctx.configure()
        img0, v0 = run()
        assert ctx.stats()["autotuned"] == 0              # below the automatic size
        ctx.configure(autotune=True)
        img1, v1 = run()
        st = ctx.stats()
        print("tuned:", st["batch"], st["col_tile"], st["chains"], flush=True)
        assert st["autotuned"] == 1 and st["col_tile"] in (4, 8, 16)
        img2, v2 = run()
        assert ctx.stats()["autotuned"] == 2                 # kept per shape
        assert (ctx.stats()["batch"], ctx.stats()["col_tile"]) == (st["batch"], st["col_tile"])
        outs = [(img1, v1), (img2, v2)]
        for tile in (4, 8, 16):
            ctx.configure(col_tile=tile)
            outs.append(run())
            assert ctx.stats()["col_tile"] == tile
        for img, v in outs:
            np.testing.assert_array_equal(img, img0)
            np.testing.assert_array_equal(v, v0)
    finally:
        ctx.configure()


def test_pixel_groups_kept_across_specimens_on_one_grid(msf, ctx):
    """The detector and FEM pixel groups depend on the grid, the aperture and
    the detector/FEM settings only, so runs on a new specimen of the same
    grid reuse them: realizations run one after another on one context give
    the same results bit for bit as each run alone on a fresh context, also
    when the aperture or the detector changes in between."""
    runs = []
    for seed in (31, 32, 33):
        q, Z = random_atoms(60, (12.0, 12.0, 6.0), seed=seed, species=(14, 29))
        runs.append(Case(nx=128, ny=128, cell=(12.0, 12.0, 6.0), voltage=300e3, q=q, Z=Z,
                         min_slice=1.5, alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.7))
    runs.append(Case(**{**runs[1].__dict__, "alpha_max": 0.015}))
    runs.append(Case(**{**runs[0].__dict__, "inner": 0.05}))
    xp = np.linspace(0.0, 11.0, 6)
    yp = np.linspace(0.5, 10.5, 5)

    def run(context, c):
        spec = c.gpu_specimen(msf, context)
        out = context.adfstem_fem(spec, c.optics(msf), xp, yp, c.femcfg(msf),
                                  detector=c.detector(msf))
        spec.free()
        return out

    chained = [run(ctx, c) for c in runs]
    for c, (img, v) in zip(runs, chained):
        fresh = msf.Context(0)
        try:
            img0, v0 = run(fresh, c)
        finally:
            fresh.close()
        np.testing.assert_array_equal(img, img0)
        for m in msf.FEM_MODES:
            np.testing.assert_array_equal(v[m], v0[m])


def test_failed_run_after_a_group_rebuild_leaves_the_cache_consistent(msf, ctx):
    """A run that rebuilds the pixel groups for detector B and then fails
    (here: its batch buffer cannot be allocated) must not let a later run
    with detector A reuse B's tables: that run equals a fresh context's bit
    for bit."""
    q, Z = random_atoms(60, (20.0, 20.0, 6.0), seed=41, species=(14, 29))
    a = Case(nx=256, ny=256, cell=(20.0, 20.0, 6.0), voltage=300e3, q=q, Z=Z, min_slice=1.5,
             alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.7)
    b = Case(**{**a.__dict__, "inner": 0.05, "dr": 0.9})
    xp = np.linspace(0.0, 19.0, 4)
    yp = np.linspace(0.5, 18.5, 3)
    spec = a.gpu_specimen(msf, ctx)
    try:
        ctx.adfstem(spec, a.optics(msf), xp, yp, detector=a.detector(msf), fem=a.femcfg(msf))
        # 1000 x 1000 positions in one batch: ~360 GB of wave buffer at 256^2
        big = np.linspace(0.0, 19.0, 1000)
        ctx.configure(batch=1000 * 1000)
        print("group rebuild for detector B, then an allocation failure", flush=True)
        with pytest.raises(msf.MsfError) as e:
            ctx.adfstem(spec, b.optics(msf), big, big, detector=b.detector(msf),
                        fem=b.femcfg(msf))
        assert e.value.code == -4                      # MSF_ERR_MEMORY
        ctx.configure()
        img, v = ctx.adfstem(spec, a.optics(msf), xp, yp, detector=a.detector(msf),
                             fem=a.femcfg(msf))
    finally:
        ctx.configure()
        spec.free()
    fresh = msf.Context(0)
    try:
        s0 = a.gpu_specimen(msf, fresh)
        img0, v0 = fresh.adfstem(s0, a.optics(msf), xp, yp, detector=a.detector(msf),
                                 fem=a.femcfg(msf))
        s0.free()
    finally:
        fresh.close()
    np.testing.assert_array_equal(img, img0)
    np.testing.assert_array_equal(v, v0)
