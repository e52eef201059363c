"""libmsf.so loads without a GPU, exports every entry point of
include/msf.h, and its host-side rules agree with the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="msf.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(msf_\w+)\(", txt, re.M)))


@pytest.fixture(scope="module")
def msf():
    import msf as m
    m.lib()
    return m


def test_library_exports_every_declared_symbol(msf):
    names = _declared()
    assert len(names) >= 18
    lib = C.CDLL(msf.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(msf.EXPORTS) == names
    io = _declared("msf_io.h")
    assert len(io) == 12
    for n in io:
        assert hasattr(lib, n), n
    assert sorted(msf.IO_EXPORTS) == io


def test_wavelength_and_scherzer_bitwise(msf, O):
    for v in (60e3, 100e3, 200e3, 300e3):
        assert msf.wavelength(v) == O.wavelength(v)
    lam = O.wavelength(200e3)[0]
    assert msf.scherzer_defocus(lam, 1.2e7) == O.scherzer_defocus_uncorrected(lam, 1.2e7)
    assert msf.scherzer_alphamax(lam, 1.2e7) == O.scherzer_alphamax_uncorrected(lam, 1.2e7)
    L = O.lib()
    assert msf.scherzer_defocus(lam, 5e7, 1) == L.orc_scherzer_defocus_correctedtoCs5(lam, 5e7)
    assert msf.scherzer_alphamax(lam, 5e7, 1) == L.orc_scherzer_alphamax_correctedtoCs5(lam, 5e7)
    assert msf.scherzer_cs3(lam, 5e7) == L.orc_scherzer_cs3_correctedtoCs5(lam, 5e7)
    # Kirkland's table value: C_S3 = -2.289 (lambda C_S5^2)^(1/3)
    assert abs(msf.scherzer_cs3(lam, 5e7) / (-2.289 * (lam * 5e7 ** 2) ** (1 / 3)) - 1) < 1e-15


@pytest.mark.parametrize("n,L,dr", [(512, 40.0, 0.1414213562), (256, 20.0, 1 / np.sqrt(2)),
                                    (384, 37.0, 0.3), (128, 12.0, 0.5)])
def test_fem_bins_match_oracle(msf, O, n, L, dr):
    g = msf.make_grid(n, n, L, L, L, 200e3)
    edges = msf.fem_bins(g, dr)
    kx, ky, _, _ = O.domains(n, n, L, L)
    _, bwt = O.bwcutoffs(n, n, L, L)
    nb, dk = O.fem_nbins(kx, ky, dr, bwt)
    assert edges.size == nb
    np.testing.assert_array_equal(edges, np.arange(nb, dtype=np.uint32) * dk)


def test_supported_lengths(msf):
    for n in (16, 32, 64, 96, 128, 192, 256, 384, 512, 768, 1024):
        assert msf.supported_length(n)
    for n in (100, 2048, 0):
        assert not msf.supported_length(n)


def test_errors_are_reported_not_crashing(msf):
    with pytest.raises(msf.MsfError):
        msf.wavelength(-1.0)
    with pytest.raises(msf.MsfError):
        msf.fem_bins(msf.make_grid(64, 64, 8.0, 8.0, 8.0, 200e3), -1.0)


def test_probe_shifts_match_oracle(msf, O):
    """msf_probe_shifts (host rule, no GPU) against the oracle's raster loop,
    including positions where no index qualifies and the shift persists."""
    g = msf.make_grid(64, 48, 12.0, 9.0, 5.0, 300e3)
    xx, yy = O.domains(64, 48, 12.0, 9.0)[2:4]
    x_p = np.array([0.0, 3.3, 11.99, 11.95, 6.0, 12.0 - 12.0 / 128])
    y_p = np.array([8.99, 0.05, 4.4, 9.0])
    got = msf.probe_shifts(g, x_p, y_p)
    want = np.stack(O.probe_shifts(xx, yy, x_p, y_p), axis=1)
    np.testing.assert_array_equal(got, want)
    with pytest.raises(msf.MsfError) as e:
        msf.probe_shifts(g, np.array([-0.5]), np.array([1.0]))
    assert e.value.code == -5


@pytest.mark.parametrize("n, L", [(512, 40.0), (384, 37.3), (1024, 80.0), (96, 7.5)])
def test_probe_shift_search_equals_the_linear_scan(msf, O, n, L):
    """The library bisects for the first grid index with x[i] - q >= dx/2;
    the oracle scans for it as the reference does (adfstem.cpp:1590-1606).
    Random positions, positions exactly on the grid points and exactly half a
    pixel from them (where the test flips), and positions past the last
    point (the shift persists from the previous position)."""
    g = msf.make_grid(n, n, L, L, 5.0, 300e3)
    xx, yy = O.domains(n, n, L, L)[2:4]
    rng = np.random.default_rng(n)
    dx = xx[1] - xx[0]
    x_p = np.concatenate([rng.random(200) * L, xx[::7], xx[::5] + dx / 2, xx[::11] - dx / 2,
                          np.nextafter(xx[::9] + dx / 2, 0), [L - 1e-9, L - dx / 4]])
    x_p = x_p[x_p >= 0]
    y_p = np.array([0.0, yy[-1], yy[n // 3] + (yy[1] - yy[0]) / 2])
    got = msf.probe_shifts(g, x_p, y_p)
    want = np.stack(O.probe_shifts(xx, yy, x_p, y_p), axis=1)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n, L, dr", [(128, 12.0, 0.5), (256, 20.0, 0.70710678118654746),
                                      (64, 40.0, 1.0), (512, 51.2, 0.25)])
def test_correlograph_bins_match_oracle(msf, O, n, L, dr):
    """n_phi = (# of boundaries ii*dk in [0.2, 1.0]) - 1 (adfstem.cpp:969-985),
    counted here independently; boundaries ii*dk and ii*2PI/n_phi with the
    reference's PI."""
    g = msf.make_grid(n, n, L, L, 5.0, 200e3)
    nb = msf.fem_bins(g, dr).size
    nphi, k, phi = msf.correlograph_bins(g, dr)
    dk = k[1]
    assert k.size == nb + 1
    np.testing.assert_array_equal(k, np.array([ii * dk for ii in range(nb + 1)]))
    want = sum(1 for ii in range(nb + 1) if 0.2 <= ii * dk <= 1.0) - 1
    assert nphi == max(want, 0) == O.lib().orc_correlograph_nphi(dk, nb)
    if nphi:
        np.testing.assert_array_equal(phi, np.arange(nphi + 1) * (2 * 3.14159265369 / nphi))


def test_headers_compile_as_c99(tmp_path):
    """include/msf.h and include/msf_io.h are plain C: a C99 translation unit
    that includes both and touches the configuration and statistics structs
    compiles without warnings under -pedantic."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "h.c"
    src.write_text('#include "msf.h"\n#include "msf_io.h"\n'
                   "int main(void) { msf_config c = {0}; msf_stats s = {0}; msf_fem_outputs o = {0};\n"
                   "  c.fft_plan = 2; c.col_tile = 4;\n"
                   "  return (int)(c.fft_plan + s.row_fft_plan + s.col_fft_plan) + (o.variance_2d != 0) - 2; }\n")
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o",
                        str(tmp_path / "h.o")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


def test_tuning_file_round_trip(msf, tmp_path):
    """msf_tuning_load merges a file's entries, msf_tuning_save writes them
    back, malformed files are rejected; host only."""
    msf.tuning_clear()
    p = tmp_path / "tune.txt"
    p.write_text("# comment\n"
                 "gfx950-256cu 512 512 352 256 av- plan0 = 64 8 2 0\n"
                 "gfx950-256cu 1024 1024 688 256 av- plan3 = 17 4 2 3\n")
    assert msf.tuning_load(p) == 2
    out = tmp_path / "saved.txt"
    msf.tuning_save(out)
    lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert sorted(lines) == ["gfx950-256cu 1024 1024 688 256 av- plan3 = 17 4 2 3",
                             "gfx950-256cu 512 512 352 256 av- plan0 = 64 8 2 0"]
    bad = tmp_path / "bad.txt"
    bad.write_text("gfx950-256cu 512 512 352 256 av- plan0 = 64 5 2 0\n")   # no 5-column tile
    with pytest.raises(msf.MsfError):
        msf.tuning_load(bad)
    with pytest.raises(msf.MsfError):
        msf.tuning_load(tmp_path / "missing.txt")
    msf.tuning_clear()
    msf.tuning_save(out)
    assert [l for l in out.read_text().splitlines() if not l.startswith("#")] == []
