"""The headline workloads at their full sizes.

- 256^2, 400 Si in a 20 A cube, 16 x 16 positions over the central 10 x 10 A:
  the whole ADF image against the FP64 oracle (1e-4) and exit intensities of
  three positions (1e-5).
- 512^2, 3500 Cu/Zr in a 40 A cube, FEM: an 8 x 8 subset of the 32 x 32
  raster (every fourth position), ADF and V_omega against the oracle (1e-4).
- 1024^2, 32000 Si in 80 x 80 x 100 A, 128 x 128 positions at 0.5 A: the
  whole job, then size-independent checks, since the oracle's shift-sum
  set-up (atoms x N^2 per slice) is out of reach at this size: finite,
  positive ADF and finite V_omega on populated bins; a 1024-position subset
  run with another batch depth and one chain reproduces its rows bit for
  bit; the FP32 slice loop against the device's FP64 stage-by-stage path
  (msf_probe_k + msf_slice_step_batch per slice, checked against the oracle
  at small sizes) for two positions: exit intensities 1e-5 and ADF 1e-4.
"""
import os
import sys

import numpy as np
import pytest

from helpers import Case, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ms-stem-fem_amd"))
import specimens as SP  # noqa: E402


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def test_si_256_full_raster_matches_oracle(msf, ctx, O):
    job, q, Z, xp, yp = SP.config_si_256(seed=1001)
    c = Case(nx=256, ny=256, cell=(20.0, 20.0, 20.0), voltage=300e3, q=q, Z=Z, min_slice=2.0,
             cs3=0.0, defocus=0.0, alpha_max=0.010, inner=0.040, outer=0.200)
    print("si256: full raster", flush=True)
    _, (img_o, _, _) = c.run_oracle(O, xp, yp)
    spec = c.gpu_specimen(msf, ctx)
    img_g, _ = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf))
    err = rel_l2(img_g, img_o)
    print("si256 ADF", err, flush=True)
    assert img_g.shape == (16, 16) and err < 1e-4
    xs, ys = xp[[0, 7, 15]], yp[[3]]
    _, (_, _, ex_o) = c.run_oracle(O, xs, ys, exit_waves=True)
    ex_g = ctx.exit_waves(spec, c.optics(msf), xs, ys)
    errs = [rel_l2(np.abs(ex_g[p]) ** 2, np.abs(ex_o[p]) ** 2) for p in range(3)]
    print("si256 exit", errs, flush=True)
    assert max(errs) < 1e-5


def test_cuzr_512_fem_subset_matches_oracle(msf, ctx, O):
    job, q, Z, xp, yp = SP.config_cuzr_512(seed=2001)
    lam = msf.wavelength(job["voltage"])[0]
    df = msf.scherzer_defocus(lam, job["cs3"])
    c = Case(nx=512, ny=512, cell=(40.0, 40.0, 40.0), voltage=200e3, q=q, Z=Z, min_slice=2.0,
             cs3=job["cs3"], defocus=df, alpha_max=0.002, inner=0.040, outer=0.200, fem=True,
             dr=job["dr"])
    xs, ys = xp[::4], yp[::4]
    print("cuzr512: 8 x 8 positions", flush=True)
    _, (img_o, v_o, _) = c.run_oracle(O, xs, ys)
    spec = c.gpu_specimen(msf, ctx)
    img_g, v_g = ctx.adfstem(spec, c.optics(msf), xs, ys, detector=c.detector(msf),
                             fem=c.femcfg(msf))
    e_adf, e_v = rel_l2(img_g, img_o), rel_l2(v_g, v_o)
    print("cuzr512 ADF", e_adf, "V", e_v, flush=True)
    assert e_adf < 1e-4 and e_v < 1e-4
    # V over 0 - 1.0 1/A, the range the workload reports
    nb = int(round(1.0 / 0.005))
    assert rel_l2(v_g[:nb], v_o[:nb]) < 1e-4


def test_si_1024_large_job(msf, ctx):
    print("si1024: generating 32000 atoms", flush=True)
    q, Z = SP.amorphous(32000, (80.0, 80.0, 100.0), 2.0, 4001)
    c = Case(nx=1024, ny=1024, cell=(80.0, 80.0, 100.0), voltage=300e3, q=q, Z=Z,
             min_slice=2.0, cs3=0.0, defocus=0.0, alpha_max=0.020, inner=0.060, outer=0.200,
             fem=True, dr=0.1414213562)
    grid = c.grid(msf)
    spec = c.gpu_specimen(msf, ctx)
    S = spec.n_slices
    assert 45 <= S <= 55
    xp = 8.0 + 0.5 * np.arange(128)
    yp = 8.0 + 0.5 * np.arange(128)
    print(f"si1024: {S} slices, 16384 positions", flush=True)
    img, v = ctx.adfstem(spec, c.optics(msf), xp, yp, detector=c.detector(msf),
                         fem=c.femcfg(msf))
    st = ctx.stats()
    print("si1024 job", st["slice_loop_ms"], "ms, batch", st["batch"], "chains", st["chains"],
          flush=True)
    assert img.shape == (128, 128) and np.isfinite(img).all() and (img > 0).all()
    lam = msf.wavelength(300e3)[0]
    edges = msf.fem_bins(grid, c.dr)
    live = lam * (edges + (edges[1] - edges[0])) > c.alpha_max   # outside the direct beam
    assert np.isfinite(v[live][:-1]).all() and (v[live][:-1] >= 0).all()
    # a subset with another launch shape reproduces its rows bit for bit
    try:
        ctx.configure(batch=37, chains=1)
        img_s, _ = ctx.adfstem(spec, c.optics(msf), xp[:8], yp, detector=c.detector(msf))
    finally:
        ctx.configure()
    np.testing.assert_array_equal(img_s, img[:8])
    # FP32 slice loop against the FP64 stage-by-stage device path
    sh = msf.probe_shifts(grid, xp, yp)[[0, 8255]]
    w = np.stack([ctx.probe_k(grid, c.optics(msf), int(a), int(b)) for a, b in sh])
    for n in range(S):
        w = ctx.slice_step(spec, n, w)
    ex = ctx.exit_waves_shifted(spec, c.optics(msf), sh)
    errs = [rel_l2(np.abs(ex[k]) ** 2, np.abs(w[k]) ** 2) for k in range(2)]
    adf64 = ctx.adf_reduce(grid, c.detector(msf), w)
    adf32 = img.reshape(-1)[[0, 8255]]
    print("si1024 exit", errs, "ADF", np.abs(adf32 / adf64 - 1), flush=True)
    assert max(errs) < 1e-5
    np.testing.assert_allclose(adf32, adf64, rtol=1e-4)
