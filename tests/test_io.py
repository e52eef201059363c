"""Model readers and result writers of the command line (include/msf_io.h),
host code only.  Each expectation restates the reference reader or writer
it covers (tem_io.cpp, write_tif.hpp, io_txt.cpp) on hand-built inputs."""
import struct

import numpy as np
import pytest

import msf


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


# ------------------------------------------------------------ elements
def test_element_symbols_and_numbers():
    assert msf.element_z("Si") == 14
    assert msf.element_z("si") == 14
    assert msf.element_z("SI") == 14
    assert msf.element_z("14") == 14
    assert msf.element_z("h") == 1 and msf.element_z("lr") == 103
    assert msf.element_z("Cu") == 29 and msf.element_z("Zr") == 40
    for bad in ("", "Xx", "014", "14.0", "104", "0", "Uuo"):
        assert msf.element_z(bad) == 0, bad
    zs = [msf.element_z(str(z)) for z in range(1, 104)]
    assert zs == list(range(1, 104))


def test_guess_z_windows_in_reference_order():
    assert msf.guess_z(28.0855) == 14
    assert msf.guess_z(63.546) == 29
    assert msf.guess_z(91.224) == 40
    assert msf.guess_z(39.948) == 18
    assert msf.guess_z(40.0) == 20        # the Ar window is open at 40.0
    assert msf.guess_z(36.0) == 18        # exact-mass aliases
    assert msf.guess_z(56.0) == 28
    # windows are tested in order: Ni's 58.693 falls in Co's [58, 59] first
    assert msf.guess_z(58.693) == 27
    # Cr's window ends at 51.99999, Mn's starts at 54.0
    with pytest.raises(msf.MsfError):
        msf.guess_z(52.5)
    # the Tm window is empty ([168.9, 168.0]); only its alias 138.0 matches
    with pytest.raises(msf.MsfError):
        msf.guess_z(168.93)
    assert msf.guess_z(138.0) == 69
    with pytest.raises(msf.MsfError):
        msf.guess_z(5.0)


# ------------------------------------------------------------ xyz
def test_xyz_reader_shifts_by_negative_minimum(tmp_path):
    p = _write(tmp_path, "m.xyz", "3\n10 12 14\nSi 1.0 2.0 3.0\ncu -1.5 4.0 0.5\n29 2.0 -0.25 9.0\n")
    q, Z, per = msf.read_model(p)
    assert per == (10.0, 12.0, 14.0)
    np.testing.assert_array_equal(Z, [14, 29, 29])
    # extrema start at 0, so only negative coordinates move the origin
    np.testing.assert_array_equal(q, [[2.5, 2.25, 3.0], [0.0, 4.25, 0.5], [3.5, 0.0, 9.0]])


def test_xyz_reader_all_positive_keeps_coordinates(tmp_path):
    p = _write(tmp_path, "m.xyz", "2\n10 10 10 extra words\nO 1 2 3\nO 4 5 6 0.9 more\n")
    q, Z, _ = msf.read_model(p)
    np.testing.assert_array_equal(q, [[1, 2, 3], [4, 5, 6]])
    np.testing.assert_array_equal(Z, [8, 8])


def test_xyz_reader_ignores_lines_beyond_the_declared_count(tmp_path):
    p = _write(tmp_path, "m.xyz", "1\n10 10 10\nO 1 2 3\nnot an atom line\n")
    q, Z, _ = msf.read_model(p)
    assert Z.size == 1


@pytest.mark.parametrize("text, why", [
    ("0\n10 10 10\n", "declared population 0"),
    ("2\n10 10 10\nSi 1 1 1\n", "fewer atoms than declared"),
    ("1\n10 10 10\nSi 1 1 1", "last atom line without a newline (good() fails at EOF)"),
    ("1\n10 0 10\nSi 1 1 1\n", "non-positive period"),
    ("1\n10 10 10\nXx 1 1 1\n", "unknown element"),
    ("2\n10 10 10\n\nSi 1 1 1\n", "blank line read as an atom"),
    ("2\n5 5 5\nSi 0 0 0\nSi 5.5 1 1\n", "positions exceed the period"),
])
def test_xyz_reader_rejects(tmp_path, text, why):
    p = _write(tmp_path, "bad.xyz", text)
    with pytest.raises(msf.MsfError):
        msf.read_model(p)


def test_model_extension_rule(tmp_path):
    """Only a name whose text after the last '.' is exactly "xyz" is read as
    xyz; anything else goes to the LAMMPS data reader (tem.cpp:270-330)."""
    text = "1\n10 10 10\nSi 1 1 1\n"
    ok = _write(tmp_path, "a.xyz", text)
    assert msf.read_model(ok)[1].tolist() == [14]
    for name in ("a.XYZ", "a.xyz.txt"):
        with pytest.raises(msf.MsfError):
            msf.read_model(_write(tmp_path, name, text))


# ------------------------------------------------------------ LAMMPS data
LAMMPS = """LAMMPS data file written by a test

4 atoms
2 atom types

-1.0 9.0 xlo xhi
0.0 8.0 ylo yhi
2.0 12.0 zlo zhi

Masses

1 28.0855
2 63.546

Atoms # atomic

1 1 0.0 1.0 3.0
2 2 -1.0 0.0 2.0
3 1 9.0 8.0 12.0
4 2 4.5 4.0 7.0
"""


def test_lammps_reader_shifts_by_box_lower_bounds(tmp_path):
    q, Z, per = msf.read_model(_write(tmp_path, "m.data", LAMMPS))
    assert per == (10.0, 8.0, 10.0)
    np.testing.assert_array_equal(Z, [14, 29, 14, 29])
    np.testing.assert_array_equal(q, [[1, 1, 1], [0, 0, 0], [10, 8, 10], [5.5, 4, 5]])


@pytest.mark.parametrize("old, new", [
    ("4 atoms", "4 atom"),                 # descriptor must be "atoms"
    ("2 atom types", "2 types"),
    ("-1.0 9.0 xlo xhi", "9.0 -1.0 xlo xhi"),   # upper <= lower
    ("Masses", "Mass"),
    ("2 63.546", "2 5.0"),                 # mass without an element window
    ("4 2 4.5 4.0 7.0", "4 3 4.5 4.0 7.0"),      # atom type without a mass
    ("4 2 4.5 4.0 7.0", "4 2 9.5 4.0 7.0"),      # outside the box
    ("4 2 4.5 4.0 7.0", "5 2 4.5 4.0 7.0"),      # numbered beyond the count
    ("2 63.546", "2 28.0855"),             # two types map to one element
])
def test_lammps_reader_rejects(tmp_path, old, new):
    assert old in LAMMPS
    with pytest.raises(msf.MsfError):
        msf.read_model(_write(tmp_path, "bad.data", LAMMPS.replace(old, new, 1)))


@pytest.mark.parametrize("text", [
    "-1\n10 10 10\nSi 1 2 3\n",              # wraps to 2^64 - 1 as a size_t
    "18446744073709551615\n10 10 10\nSi 1 2 3\n",
    "99999999999999999999\n10 10 10\nSi 1 2 3\n",  # beyond size_t: the extraction fails
])
def test_xyz_reader_malformed_count_is_an_error_not_an_abort(tmp_path, text):
    # the reference reads atom lines until they run out (tem_io.cpp:4747-4815);
    # no count may reach the allocator
    with pytest.raises(msf.MsfError):
        msf.read_model(_write(tmp_path, "huge.xyz", text))


@pytest.mark.parametrize("old, new", [
    ("4 atoms", "-1 atoms"),
    ("2 atom types", "-1 atom types"),
    ("2 atom types", "18446744073709551615 atom types"),
])
def test_lammps_reader_malformed_counts_are_errors(tmp_path, old, new):
    assert old in LAMMPS
    with pytest.raises(msf.MsfError):
        msf.read_model(_write(tmp_path, "huge.data", LAMMPS.replace(old, new, 1)))


# ------------------------------------------------------------ TIFF
def _read_tiff16(path):
    b = open(path, "rb").read()
    assert b[:4] == b"II*\x00"
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    tags = {}
    for k in range(n):
        tid, typ, cnt, val = struct.unpack_from("<HHII", b, ifd + 2 + 12 * k)
        if typ == 3 and cnt == 1:
            val &= 0xFFFF
        tags[tid] = (typ, cnt, val)
    w, h = tags[256][2], tags[257][2]
    assert tags[258][2] == 16 and tags[259][2] == 1 and tags[277][2] == 1
    nstrip = tags[273][1]
    offs = [tags[273][2]] if nstrip == 1 else list(struct.unpack_from(f"<{nstrip}I", b, tags[273][2]))
    cnts = [tags[279][2]] if nstrip == 1 else list(struct.unpack_from(f"<{nstrip}I", b, tags[279][2]))
    raw = b"".join(b[o:o + c] for o, c in zip(offs, cnts))
    img = np.frombuffer(raw, "<u2").reshape(h, w)
    xr = struct.unpack_from("<II", b, tags[282][2])
    return img, tags, xr[0] / xr[1]


def test_tiff16_linear_quantisation_and_row_order(tmp_path):
    rng = np.random.default_rng(4)
    data = rng.random((5, 7)) * 3e-9 + 1e-10
    p = str(tmp_path / "img.tif")
    msf.write_tiff16(p, data, resolution_unit=3, x_resolution=8e7, y_resolution=8e7)
    img, tags, xres = _read_tiff16(p)
    lo, hi = data.min(), data.max()
    expect = np.floor((data - lo) * (65535.0 / (hi - lo)) + 0.5).astype(np.uint16)
    # rows are written bottom-up (write_tif.hpp:317-321)
    np.testing.assert_array_equal(img, expect[::-1])
    assert img.max() == 65535 and img.min() == 0
    assert tags[278][2] == 7                  # RowsPerStrip = width
    assert tags[296][2] == 3 and tags[274][2] == 1 and tags[262][2] == 1
    assert xres == 8e7


def test_tiff16_log_scale_and_tall_image(tmp_path):
    data = np.linspace(0.0, 5.0, 40).reshape(20, 2)   # several strips of 2 rows
    p = str(tmp_path / "log.tif")
    msf.write_tiff16(p, data, logscale=True)
    img, tags, _ = _read_tiff16(p)
    t = np.log(data - 0.0 + 1) / np.log(5.0 - 0.0 + 1)
    expect = np.where(t > 0, np.floor(t * 65535.0 + 0.5), 0).astype(np.uint16)
    np.testing.assert_array_equal(img, expect[::-1])
    assert tags[273][1] == 10


def test_tiff16_constant_image_is_zero(tmp_path):
    p = str(tmp_path / "c.tif")
    msf.write_tiff16(p, np.full((3, 3), 2.5))
    img, _, _ = _read_tiff16(p)
    assert not img.any()


# ------------------------------------------------------------ txt, sidecar
def test_variance_txt_format(tmp_path):
    p = str(tmp_path / "v.txt")
    edges = np.arange(4) * 0.005
    vals = np.array([0.0, 1.25e-3, np.nan, 3.0])
    msf.write_variance_txt(p, vals, edges)
    lines = open(p).read().splitlines()
    assert lines == ["0.0000000000e+00 0.0000000000e+00", "5.0000000000e-03 1.2500000000e-03",
                     "1.0000000000e-02 nan", "1.5000000000e-02 3.0000000000e+00"]


def test_f64_sidecar_roundtrip(tmp_path):
    p = str(tmp_path / "a.f64")
    a = np.arange(12.0).reshape(3, 4) / 7
    msf.write_f64(p, a)
    raw = open(p, "rb").read()
    head, body = raw.split(b"\n", 1)
    assert head == b"msf-f64 3 4"
    np.testing.assert_array_equal(np.frombuffer(body, "<f8").reshape(3, 4), a)


def test_correlograph_txt_format(tmp_path):
    """io_txt.cpp:96-165: "0 " and the phi boundaries, then per k bin its
    boundary and the values data[ii + jj * n_k]; every value followed by a
    space."""
    p = str(tmp_path / "c.txt")
    data = np.array([[1.0, 2.0, 3.0], [-4.5, 0.0, 6.25e-7]])  # [n_phi = 2, n_k = 3]
    msf.write_correlograph_txt(p, data, np.arange(4) * 0.25, np.arange(3) * np.pi)
    text = open(p).read()
    assert text == ("0 0.0000000000e+00 3.1415926536e+00 \n"
                    "0.0000000000e+00 1.0000000000e+00 -4.5000000000e+00 \n"
                    "2.5000000000e-01 2.0000000000e+00 0.0000000000e+00 \n"
                    "5.0000000000e-01 3.0000000000e+00 6.2500000000e-07 \n")
    # no phi sectors: the corner and one boundary per k bin
    msf.write_correlograph_txt(p, np.zeros((0, 2)), [0.0, 0.5, 1.0], [])
    assert open(p).read() == "0 \n0.0000000000e+00 \n5.0000000000e-01 \n"


def test_mtf_reader(tmp_path):
    """read_mtf_file (tem_io.cpp:5722-5800): domain on the first line, values
    on the second, commas as blanks; trailing blanks repeat a line's last
    value (the stream extraction that finds only blanks leaves its variable
    as it was)."""
    p = tmp_path / "m.txt"
    p.write_text("0, 0.5, 1.0\n1.0, 0.6, 0.2\n")
    d, m = msf.read_mtf(str(p))
    np.testing.assert_array_equal(d, [0.0, 0.5, 1.0])
    np.testing.assert_array_equal(m, [1.0, 0.6, 0.2])
    p.write_text("0 0.5 \n1 0.6 \n")
    d, m = msf.read_mtf(str(p))
    np.testing.assert_array_equal(d, [0.0, 0.5, 0.5])
    np.testing.assert_array_equal(m, [1.0, 0.6, 0.6])
    p.write_text("0 0.25 1\n1e0\t0.5 2.5e-1\nignored line\n")
    d, m = msf.read_mtf(str(p))
    np.testing.assert_array_equal(m, [1.0, 0.5, 0.25])


@pytest.mark.parametrize("text", ["0 0.5\n1 0.6",          # second line lacks its newline
                                  "0 0.5 1\n1 0.6\n",       # unequal lengths
                                  "0 x\n1 0.6\n",           # the reference would loop forever
                                  ""])
def test_mtf_reader_rejects(tmp_path, text):
    p = tmp_path / "m.txt"
    p.write_text(text)
    with pytest.raises(msf.MsfError):
        msf.read_mtf(str(p))
    with pytest.raises(msf.MsfError):
        msf.read_mtf(str(tmp_path / "missing.txt"))
