"""libmsf.so loads without a GPU, exports every entry point of
include/msf.h, and its host-side rules agree with the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "msf.h")).read()
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


def test_wavelength_and_scherzer_bitwise(msf, O):
    for v in (60e3, 100e3, 200e3, 300e3):
        assert msf.wavelength(v) == O.wavelength(v)
    lam = O.wavelength(200e3)[0]
    assert msf.scherzer_defocus(lam, 1.2e7) == O.scherzer_defocus_uncorrected(lam, 1.2e7)
    assert msf.scherzer_alphamax(lam, 1.2e7) == O.scherzer_alphamax_uncorrected(lam, 1.2e7)
    L = O.lib()
    assert msf.scherzer_defocus(lam, 5e7, 1) == L.orc_scherzer_defocus_correctedtoCs5(lam, 5e7)
    assert msf.scherzer_alphamax(lam, 5e7, 1) == L.orc_scherzer_alphamax_correctedtoCs5(lam, 5e7)


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
    for n in (64, 128, 256, 384, 512, 1024):
        assert msf.supported_length(n)
    for n in (100, 2048, 0):
        assert not msf.supported_length(n)


def test_errors_are_reported_not_crashing(msf):
    with pytest.raises(msf.MsfError):
        msf.wavelength(-1.0)
    with pytest.raises(msf.MsfError):
        msf.fem_bins(msf.make_grid(64, 64, 8.0, 8.0, 8.0, 200e3), -1.0)
