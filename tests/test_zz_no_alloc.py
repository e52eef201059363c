"""msf_run_stem_noalloc: the raster run with every device buffer in the
library's static scratch array.  Its ADF image and V_omega equal
msf_run_stem's bit for bit, on two calls in a row, the second with a
smaller input (a smaller grid, fewer slices and positions), so that the
second call runs in an array still holding the first call's contents; a run
that does not fit the array is rejected with MSF_ERR_INVALID."""
import numpy as np
import pytest

from helpers import Case, random_atoms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def _cases():
    q1, Z1 = random_atoms(220, (20.0, 20.0, 14.0), seed=71, species=(14, 29, 40))
    big = Case(nx=256, ny=256, cell=(20.0, 20.0, 14.0), voltage=200e3, q=q1, Z=Z1,
               min_slice=1.5, alpha_max=0.02, inner=0.03, outer=0.2, fem=True, dr=0.5)
    q2, Z2 = random_atoms(40, (10.0, 10.0, 5.0), seed=72, species=(14,))
    small = Case(nx=128, ny=128, cell=(10.0, 10.0, 5.0), voltage=300e3, q=q2, Z=Z2,
                 min_slice=1.5, alpha_max=0.015, inner=0.04, outer=0.18, fem=True, dr=0.7)
    return [(big, np.linspace(0.3, 19.0, 12), np.linspace(0.1, 18.7, 11)),
            (small, np.linspace(0.5, 9.0, 4), np.linspace(1.0, 8.5, 3))]


def test_noalloc_equals_run_stem_on_two_calls(msf, ctx):
    cases = _cases()
    specs = [c.gpu_specimen(msf, ctx) for c, _, _ in cases]
    try:
        print("noalloc: reference runs through msf_run_stem", flush=True)
        ref = [ctx.adfstem(s, c.optics(msf), xp, yp, detector=c.detector(msf), fem=c.femcfg(msf))
               for s, (c, xp, yp) in zip(specs, cases)]
        got = []
        for s, (c, xp, yp) in zip(specs, cases):
            print(f"noalloc: {c.nx}^2, {s.n_slices} slices, {xp.size * yp.size} positions",
                  flush=True)
            got.append(ctx.adfstem_noalloc(s, c.optics(msf), xp, yp, detector=c.detector(msf),
                                           fem=c.femcfg(msf)))
        for (img, v), (img0, v0) in zip(got, ref):
            assert np.isfinite(img).all() and (img > 0).all()
            np.testing.assert_array_equal(img, img0)
            np.testing.assert_array_equal(v, v0)
        # the same small call once more on another context: nothing is kept
        other = msf.Context(0)
        try:
            c, xp, yp = cases[1]
            s2 = c.gpu_specimen(msf, other)
            img2, v2 = other.adfstem_noalloc(s2, c.optics(msf), xp, yp, detector=c.detector(msf),
                                             fem=c.femcfg(msf))
            s2.free()
        finally:
            other.close()
        np.testing.assert_array_equal(img2, ref[1][0])
        np.testing.assert_array_equal(v2, ref[1][1])
    finally:
        for s in specs:
            s.free()


def test_noalloc_rejects_a_run_beyond_the_array(msf, ctx):
    # 1024^2 with 20 slices: the vacuum waves alone take 20 x 2 x 1024^2 x 8 B = 320 MiB
    q = np.zeros((20, 3))
    q[:, 0] = 10.0
    q[:, 1] = 10.0
    q[:, 2] = 1.0 + 2.0 * np.arange(20)
    c = Case(nx=1024, ny=1024, cell=(40.0, 40.0, 42.0), voltage=300e3, q=q,
             Z=np.full(20, 14, np.uint32), min_slice=1.0, alpha_max=0.02, inner=0.04, outer=0.2)
    spec = c.gpu_specimen(msf, ctx)
    try:
        assert spec.n_slices == 20
        with pytest.raises(msf.MsfError) as e:
            ctx.adfstem_noalloc(spec, c.optics(msf), np.array([5.0]), np.array([5.0]),
                                detector=c.detector(msf))
        assert e.value.code == -1 and "scratch" in str(e.value)
        # the context still runs ordinary calls
        img, _ = ctx.adfstem(spec, c.optics(msf), np.array([5.0]), np.array([5.0]),
                             detector=c.detector(msf))
        assert np.isfinite(img).all()
    finally:
        spec.free()
