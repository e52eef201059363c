"""End-to-end run of the ms-stem-fem command line on a small xyz model: the
float64 ADF sidecar and the V_omega text file match the FP64 oracle
(1e-4 relative L2), the 16-bit TIFFs are the min-max quantisation of the
sidecar, and the diffraction images are the log-scaled exit waves."""
import os
import struct
import subprocess

import numpy as np
import pytest

from helpers import Case, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ms-stem-fem_amd", "ms-stem-fem")


def _tiff(path):
    b = open(path, "rb").read()
    ifd = struct.unpack_from("<I", b, 4)[0]
    tags = {}
    for k in range(struct.unpack_from("<H", b, ifd)[0]):
        tid, typ, cnt, val = struct.unpack_from("<HHII", b, ifd + 2 + 12 * k)
        tags[tid] = (cnt, val & 0xFFFF if typ == 3 and cnt == 1 else val)
    w, h = tags[256][1], tags[257][1]
    n = tags[273][0]
    offs = [tags[273][1]] if n == 1 else struct.unpack_from(f"<{n}I", b, tags[273][1])
    cnts = [tags[279][1]] if n == 1 else struct.unpack_from(f"<{n}I", b, tags[279][1])
    raw = b"".join(b[o:o + c] for o, c in zip(offs, cnts))
    return np.frombuffer(raw, "<u2").reshape(h, w)[::-1]   # rows are stored bottom-up


def test_cli_adfstem_fem_run(tmp_path, O):
    rng = np.random.default_rng(31)
    n_at = 45
    q = rng.random((n_at, 3)) * [10.0, 10.0, 6.0]
    names = rng.choice(["Si", "Cu"], n_at)
    lines = [f"{n_at}", "10 10 6"] + [f"{e} {x:.6f} {y:.6f} {z:.6f}" for e, (x, y, z) in zip(names, q)]
    (tmp_path / "m.xyz").write_text("\n".join(lines) + "\n")
    args = [CLI, "-m", "128", "128", "300000", "-a", "m.xyz", "--output_prefix", "run",
            "--adfstemuncorrfem", "--cs3", "1.0e6", "--defocus", "25", "--alphamax", "0.02",
            "--rasterspacing", "2.5", "--minslice", "1.5", "--dr", "0.5", "--images",
            "--V_omega", "--Vbar_r", "--V_re", "--Omega_vimage", "--V_gt17", "--V_rbar"]
    print("cli:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    # the same job through the oracle: positions 0, 2.5, 5, 7.5 on both axes
    qq = np.array([[float(v) for v in l.split()[1:]] for l in lines[2:]])
    Z = np.array([14 if e == "Si" else 29 for e in names], np.uint32)
    c = Case(nx=128, ny=128, cell=(10.0, 10.0, 6.0), voltage=300e3, q=qq, Z=Z, min_slice=1.5,
             cs3=1.0e6, defocus=25.0, alpha_max=0.02, inner=0.040, outer=0.200, fem=True, dr=0.5)
    xp = np.arange(4) * 2.5
    _, (img_o, v_o, ex_o) = c.run_oracle(O, xp, xp, exit_waves=True)
    job = c.oracle_job(O)
    _, modes_o = O.adfstem_fem(job, O.Specimen(job, c.q, c.Z.astype(np.int32)), xp, xp)

    pre = tmp_path / "run_adfstem_uncorr"
    raw = open(f"{pre}_stem_image_output.f64", "rb").read()
    head, body = raw.split(b"\n", 1)
    assert head == b"msf-f64 4 4"
    img = np.frombuffer(body, "<f8").reshape(4, 4)
    assert rel_l2(img, img_o) < 1e-4

    txt = np.loadtxt(f"{pre}_fem_V_omega.txt")
    assert txt.shape == (v_o.size, 2)
    np.testing.assert_allclose(txt[:, 0], np.arange(v_o.size) * txt[1, 0], rtol=1e-9)
    assert rel_l2(txt[:, 1], v_o) < 1e-4

    for mode, suffix in (("vbar_r", "Vbar_r"), ("v_re", "V_re"), ("omega_vimage", "Omega_vimage"),
                         ("v_gt17", "V_gt17"), ("v_rbar", "V_rbar")):
        t = np.loadtxt(f"{pre}_fem_{suffix}.txt")
        np.testing.assert_array_equal(t[:, 0], txt[:, 0])
        assert rel_l2(t[:, 1], modes_o[mode]) < 1e-4, mode

    lo, hi = img.min(), img.max()
    lin = _tiff(f"{pre}_stem_image_output.tif")
    np.testing.assert_array_equal(lin, np.floor((img - lo) * (65535.0 / (hi - lo)) + 0.5))
    lg = _tiff(f"{pre}_stem_image_output_logscale.tif")
    t = np.log(img - lo + 1) / np.log(hi - lo + 1)
    np.testing.assert_array_equal(lg, np.where(t > 0, np.floor(t * 65535.0 + 0.5), 0))

    for s, tag, k in (("1e-05", "first", 0), ("1", "first", 0), ("1e+20", "first", 0),
                      ("1", "final", 15)):
        d = _tiff(f"{pre}_{tag}_diffraction_scalefactor_{s}.tif").astype(float)
        m = np.log(1.0 + float(s) * np.abs(ex_o[k]) ** 2)
        m = np.fft.fftshift(m)
        want = np.floor((m - m.min()) * (65535.0 / (m.max() - m.min())) + 0.5)
        bad = np.abs(d - want) > 2
        assert bad.mean() < 1e-3, (s, tag, bad.sum())

    # per-pixel averages over the raster and the 2-D variance
    # (adfstem.cpp:3063-3097, :3493-3520): log(1 + s v^2) of the real arrays,
    # quadrants swapped; the variance's 0/0 pixels are left out of the range
    # and written as 0
    _, r_o = O.adfstem_fem(job, O.Specimen(job, c.q, c.Z.astype(np.int32)), xp, xp,
                           modes=("omega_vimage",), exit_waves=True)
    a = np.abs(r_o["exit_waves"]).mean(axis=0)
    sq = (np.abs(r_o["exit_waves"]) ** 2).mean(axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        var = sq / (a * a) - 1.0
    kx, ky, _, _ = O.domains(128, 128, 10.0, 10.0)
    var[O.wavelength(300e3)[0] ** 2 * (kx[:, None] ** 2 + ky[None, :] ** 2) <= 0.02 ** 2] = 0.0
    for name, v, s in (("diffracted_wave_avg", a, "1e+13"), ("diffracted_wave_sqr_avg", sq, "1e+13"),
                       ("variance_2-D", var, "1"), ("variance_2-D", var, "1e-10")):
        d = _tiff(f"{pre}_{name}_diffraction_scalefactor_{s}.tif").astype(float)
        m = np.fft.fftshift(np.log(1.0 + float(s) * v * v))
        lo, hi = np.nanmin(m), np.nanmax(m)
        want = np.where(np.isnan(m), 0.0, np.floor((m - lo) * (65535.0 / (hi - lo)) + 0.5))
        np.testing.assert_array_equal(d[np.isnan(m)], 0.0)
        off = np.abs(d - want)
        print(name, s, "max off", off.max(), "frac>2", (off > 2).mean(), flush=True)
        assert (off > 2).mean() < 1e-3, (name, s)


def test_cli_correlograph_run(tmp_path, O):
    """--correlograph_variance: the average and sums-of-squares text files
    against the oracle (1e-4 relative L2) and the images derived from them
    (adfstem.cpp:3617-3745)."""
    rng = np.random.default_rng(31)
    n_at = 45
    q = rng.random((n_at, 3)) * [10.0, 10.0, 6.0]
    names = rng.choice(["Si", "Cu"], n_at)
    lines = [f"{n_at}", "10 10 6"] + [f"{e} {x:.6f} {y:.6f} {z:.6f}" for e, (x, y, z) in zip(names, q)]
    (tmp_path / "m.xyz").write_text("\n".join(lines) + "\n")
    # alphamax 0.01 at 300 kV leaves the rings beyond 0.51 1/A outside the
    # zeroed direct beam, so the variance over k in [0.2, 1.0] is populated
    args = [CLI, "-m", "128", "128", "300000", "-a", "m.xyz", "--output_prefix", "run",
            "--adfstemuncorrfem", "--cs3", "1.0e6", "--defocus", "25", "--alphamax", "0.01",
            "--rasterspacing", "2.5", "--minslice", "1.5", "--dr", "0.5",
            "--correlograph_variance", "--correlograph_everytxt", "--correlograph_everyimage"]
    print("cli:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert not (tmp_path / "run_adfstem_uncorr_fem_V_omega.txt").exists()  # not the default here

    qq = np.array([[float(v) for v in l.split()[1:]] for l in lines[2:]])
    Z = np.array([14 if e == "Si" else 29 for e in names], np.uint32)
    c = Case(nx=128, ny=128, cell=(10.0, 10.0, 6.0), voltage=300e3, q=qq, Z=Z, min_slice=1.5,
             cs3=1.0e6, defocus=25.0, alpha_max=0.01, fem=True, dr=0.5)
    xp = np.arange(4) * 2.5
    job = c.oracle_job(O)
    _, mo = O.adfstem_fem(job, O.Specimen(job, c.q, c.Z.astype(np.int32)), xp, xp,
                          modes=("v_omega",), correlograph=True, correlograph_each=True)

    pre = tmp_path / "run_adfstem_uncorr"
    nphi, nk = mo["correlograph_avg"].shape
    avg = np.loadtxt(f"{pre}_correlograph_avg.txt")
    sqr = np.loadtxt(f"{pre}_correlograph_variance.txt")
    assert avg.shape == sqr.shape == (nk + 1, nphi + 1)
    assert avg[0, 0] == 0 and np.array_equal(avg[:, 0], sqr[:, 0])
    np.testing.assert_allclose(avg[0, 1:], np.arange(nphi) * 2 * 3.14159265369 / nphi, rtol=1e-9)
    dk = avg[2, 0]
    np.testing.assert_allclose(avg[1:, 0], np.arange(nk) * dk, rtol=1e-9)
    a, s = avg[1:, 1:].T, sqr[1:, 1:].T       # [phi, k]
    assert rel_l2(a, mo["correlograph_avg"]) < 1e-4
    assert rel_l2(s, mo["correlograph_sqr_sum"]) < 1e-4

    # images: phi down the rows, k across; the variance over the k bins from
    # the one holding 0.2 up to the one holding 1.0 (excluded)
    ca = _tiff(f"{pre}_correlograph_avg.tif").astype(float)
    want = np.floor((a - a.min()) * (65535.0 / (a.max() - a.min())) + 0.5)
    assert ca.shape == (nphi, nk) and np.abs(ca - want).max() <= 1
    kb = np.arange(nk + 1) * dk
    s0 = [i for i in range(nk) if kb[i] <= 0.2 < kb[i + 1]][0]
    s1 = [i for i in range(nk) if kb[i] <= 1.0 < kb[i + 1]][0]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = s[:, s0:s1] / (16 * a[:, s0:s1] ** 2)
    var = np.where((s[:, s0:s1] != 0) & (a[:, s0:s1] != 0) & (t >= 1), t - 1, 0.0)
    assert var.max() > 0.1
    cv = _tiff(f"{pre}_correlograph_variance.tif").astype(float)
    want = np.floor((var - var.min()) * (65535.0 / (var.max() - var.min())) + 0.5)
    assert cv.shape == var.shape and np.abs(cv - want).max() <= 2
    cl = _tiff(f"{pre}_correlograph_variance_logscale.tif").astype(float)
    tl = np.log(var - var.min() + 1) / np.log(var.max() - var.min() + 1)
    assert np.abs(cl - np.where(tl > 0, np.floor(tl * 65535.0 + 0.5), 0)).max() <= 2

    # every position: <prefix>_<x index>_<y index>.txt and .tif (adfstem.cpp:2655-2685)
    for ix, iy in ((0, 0), (1, 3), (3, 2)):
        e = np.loadtxt(f"{pre}_{ix}_{iy}.txt")
        np.testing.assert_array_equal(e[0], avg[0])
        assert rel_l2(e[1:, 1:].T, mo["correlograph_each"][ix * 4 + iy]) < 1e-4
        im = _tiff(f"{pre}_{ix}_{iy}.tif")
        assert im.shape == (nphi, nk)
    assert not (tmp_path / "run_adfstem_uncorr_4_0.txt").exists()


def test_cli_mtf_run(tmp_path, O):
    """--mtf_file / --mtf_sampling_rate: V_omega of the MTF-filtered waves and
    the final diffraction image (taken after the MTF) against the oracle; the
    ADF image and first diffraction are taken before the MTF."""
    rng = np.random.default_rng(31)
    n_at = 45
    q = rng.random((n_at, 3)) * [10.0, 10.0, 6.0]
    names = rng.choice(["Si", "Cu"], n_at)
    lines = [f"{n_at}", "10 10 6"] + [f"{e} {x:.6f} {y:.6f} {z:.6f}" for e, (x, y, z) in zip(names, q)]
    (tmp_path / "m.xyz").write_text("\n".join(lines) + "\n")
    dom = np.linspace(0.0, 1.0, 6)
    tab = 1.0 / (1.0 + (dom / 0.5) ** 2)
    (tmp_path / "mtf.txt").write_text(", ".join(f"{v:.17g}" for v in dom) + "\n" +
                                      ", ".join(f"{v:.17g}" for v in tab) + "\n")
    args = [CLI, "-m", "128", "128", "300000", "-a", "m.xyz", "--output_prefix", "run",
            "--adfstemuncorrfem", "--cs3", "1.0e6", "--defocus", "25", "--alphamax", "0.01",
            "--rasterspacing", "2.5", "--minslice", "1.5", "--dr", "0.5", "--images",
            "--mtf_file", "mtf.txt", "--mtf_sampling_rate", "9"]
    print("cli:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    qq = np.array([[float(v) for v in l.split()[1:]] for l in lines[2:]])
    Z = np.array([14 if e == "Si" else 29 for e in names], np.uint32)
    c = Case(nx=128, ny=128, cell=(10.0, 10.0, 6.0), voltage=300e3, q=qq, Z=Z, min_slice=1.5,
             cs3=1.0e6, defocus=25.0, alpha_max=0.01, inner=0.040, outer=0.200, fem=True, dr=0.5,
             mtf=tab, mtf_domain=dom, mtf_rate=9.0)
    xp = np.arange(4) * 2.5
    _, (img_o, v_o, ex_o) = c.run_oracle(O, xp, xp, exit_waves=True)
    pre = tmp_path / "run_adfstem_uncorr"
    raw = open(f"{pre}_stem_image_output.f64", "rb").read()
    img = np.frombuffer(raw.split(b"\n", 1)[1], "<f8").reshape(4, 4)
    assert rel_l2(img, img_o) < 1e-4
    txt = np.loadtxt(f"{pre}_fem_V_omega.txt")
    assert rel_l2(txt[:, 1], v_o) < 1e-4
    d = _tiff(f"{pre}_final_diffraction_scalefactor_1.tif").astype(float)
    m = np.fft.fftshift(np.log(1.0 + np.abs(ex_o[15]) ** 2))
    want = np.floor((m - m.min()) * (65535.0 / (m.max() - m.min())) + 0.5)
    bad = np.abs(d - want) > 2
    assert bad.mean() < 1e-3, bad.sum()
