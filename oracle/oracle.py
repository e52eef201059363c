"""ctypes binding of the FP64 CPU restatement (liboracle) of ms-stem-fem's
multislice ADF-STEM/FEM path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product library never uses
it.  Arrays are numpy; complex arrays are numpy complex128 of shape (nx, ny)
in the reference's row-major [i*Ny + j] layout.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmsf_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class Job(C.Structure):
    """orc_job: the inputs tem.cpp hands to adfstem()."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("xperiod", C.c_double), ("yperiod", C.c_double), ("zperiod", C.c_double),
        ("xmin", C.c_double), ("ymin", C.c_double), ("zmin", C.c_double),
        ("voltage", C.c_double), ("min_slice", C.c_double),
        ("corrected", C.c_int32),
        ("cs3", C.c_double), ("cs5", C.c_double), ("defocus", C.c_double),
        ("alpha_max", C.c_double),
        ("adf", C.c_int32), ("det_inner", C.c_double), ("det_outer", C.c_double),
        ("fem", C.c_int32), ("dr", C.c_double),
        ("transmit_mode", C.c_int32), ("adf_mode", C.c_int32),
        ("mtf", C.c_void_p), ("mtf_domain", C.c_void_p), ("mtf_n", C.c_int32),
        ("mtf_rate", C.c_double),
    ]


class _Spec(C.Structure):
    _fields_ = [
        ("n_slices", C.c_int32),
        ("t_stack", C.POINTER(C.c_double)), ("px", C.POINTER(C.c_double)),
        ("py", C.POINTER(C.c_double)), ("lower_bound", C.POINTER(C.c_double)),
        ("thickness", C.POINTER(C.c_double)),
    ]


def build():
    """Compile liboracle with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_wavelength.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_bwcutoffs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_domains.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, _dp, _dp, _dp, _dp]
        for f in ("orc_scherzer_defocus_uncorrected", "orc_scherzer_alphamax_uncorrected",
                  "orc_scherzer_defocus_correctedtoCs5", "orc_scherzer_alphamax_correctedtoCs5",
                  "orc_scherzer_cs3_correctedtoCs5"):
            getattr(L, f).argtypes = [C.c_double, C.c_double]
            getattr(L, f).restype = C.c_double
        L.orc_raster.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_fft1d.argtypes = [_cp, C.c_int, C.c_int]
        L.orc_fft2d.argtypes = [_cp, C.c_int, C.c_int, C.c_int]
        L.orc_fe.argtypes = [C.c_int, C.c_double]
        L.orc_fe.restype = C.c_double
        L.orc_template.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_double, _cp]
        L.orc_autoslice.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                    _i32, _dp, _dp, C.POINTER(C.c_int32)]
        L.orc_propagator.argtypes = [C.c_double, C.c_double, _dp, C.c_int, C.c_double, _cp]
        L.orc_shift_index.argtypes = [_dp, C.c_int, C.c_double, C.c_int64]
        L.orc_transmission.argtypes = [_dp, _i32, _i32, C.c_int, _i32, C.c_int, _cp,
                                       _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, _cp]
        L.orc_probe.argtypes = [C.POINTER(Job), _dp, _dp, _dp, _dp, _cp]
        L.orc_probe_shifts.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                       _i64, _i64]
        L.orc_propagate_position.argtypes = [_cp, C.c_int64, C.c_int64, C.c_int, _cp, _cp, _cp,
                                             _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                             _cp]
        L.orc_adf.argtypes = [_cp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_adf.restype = C.c_double
        L.orc_fem_nbins.argtypes = [_dp, _dp, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_fem_radial.argtypes = [_cp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                     C.c_double, C.c_double, _dp, _i32]
        L.orc_fem_finalize.argtypes = [_dp, _dp, C.c_int, C.c_int64, _dp]
        L.orc_specimen_build.argtypes = [C.POINTER(Job), _dp, _i32, C.c_int, C.POINTER(_Spec)]
        L.orc_specimen_free.argtypes = [C.POINTER(_Spec)]
        L.orc_adfstem.argtypes = [C.POINTER(Job), C.POINTER(_Spec), _dp, C.c_int, _dp, C.c_int,
                                  C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_initial_wave.argtypes = [_cp, C.c_int64, C.c_int64, C.c_int, C.c_int, _cp]
        L.orc_slice_step.argtypes = [_cp, _cp, _cp, _cp, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                     C.c_int]
        L.orc_adfstem_fem.argtypes = [C.POINTER(Job), C.POINTER(_Spec), _dp, C.c_int, _dp,
                                      C.c_int, C.c_int] + [C.c_void_p] * 11
        L.orc_correlograph_nphi.argtypes = [C.c_double, C.c_int]
        L.orc_mtf_2d.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, _dp, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, _dp]
        L.orc_apply_mtf.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double]
        L.orc_correlograph.argtypes = [_dp, C.c_long, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                       C.c_int, C.c_double, C.c_double, _dp, _i32, _dp, _dp,
                                       C.c_void_p]
        L.orc_fem_modes.argtypes = [_dp, C.c_long, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                    C.c_int, C.c_double, C.c_double, _dp, _i32] + [C.c_void_p] * 5
        _lib = L
    return _lib


# ---------------------------------------------------------------- scalars
def wavelength(voltage):
    lam, gam = C.c_double(), C.c_double()
    lib().orc_wavelength(voltage, C.byref(lam), C.byref(gam))
    return lam.value, gam.value


def bwcutoffs(nx, ny, xperiod, yperiod):
    a, b = C.c_double(), C.c_double()
    lib().orc_bwcutoffs(nx, ny, xperiod, yperiod, C.byref(a), C.byref(b))
    return a.value, b.value


def domains(nx, ny, xperiod, yperiod, xmin=0.0, ymin=0.0):
    kx, ky, xx, yy = (np.zeros(nx), np.zeros(ny), np.zeros(nx), np.zeros(ny))
    lib().orc_domains(nx, ny, xperiod, yperiod, xmin, ymin, kx, ky, xx, yy)
    return kx, ky, xx, yy


def raster(n, period, qmin=0.0):
    q = np.zeros(n)
    lib().orc_raster(n, period, qmin, q)
    return q


def scherzer_defocus_uncorrected(lam, cs3):
    return lib().orc_scherzer_defocus_uncorrected(lam, cs3)


def scherzer_alphamax_uncorrected(lam, cs3):
    return lib().orc_scherzer_alphamax_uncorrected(lam, cs3)


# ---------------------------------------------------------------- FFT
def fft1d(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    assert lib().orc_fft1d(a, a.size, sign) == 0
    return a


def fft2d(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    assert lib().orc_fft2d(a, a.shape[0], a.shape[1], sign) == 0
    return a


# ---------------------------------------------------------------- physics
def fe(Z, ksq):
    return lib().orc_fe(int(Z), float(ksq))


def template(Z, lam, gam, ab_inv, cutoff, kx, ky, xmin=0.0, ymin=0.0):
    out = np.zeros((kx.size, ky.size), dtype=np.complex128)
    assert lib().orc_template(int(Z), lam, gam, ab_inv, cutoff, kx, kx.size, ky, ky.size,
                              xmin, ymin, out) == 0
    return out


def autoslice(q, zmin, zperiod, min_thickness):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    n = q.size // 3
    m = max(n, 1)
    sl, lb, th = np.zeros(m, np.int32), np.zeros(m), np.zeros(m)
    ns = C.c_int32()
    lib().orc_autoslice(q, n, zmin, zperiod, min_thickness, sl, lb, th, C.byref(ns))
    k = ns.value
    return sl[:n], lb[:k].copy(), th[:k].copy()


def propagator(lam, thickness, k, bwcutoff_t):
    p = np.zeros(k.size, dtype=np.complex128)
    lib().orc_propagator(lam, thickness, np.ascontiguousarray(k, np.float64), k.size,
                         bwcutoff_t, p)
    return p


def shift_index(grid, q, previous=0):
    return lib().orc_shift_index(np.ascontiguousarray(grid, np.float64), grid.size, q, previous)


def transmission(q, Z, atoms, zlist, templates, kx, ky, xx, yy, bwcutoff_t):
    nx, ny = kx.size, ky.size
    t = np.zeros((nx, ny), dtype=np.complex128)
    rc = lib().orc_transmission(np.ascontiguousarray(q, np.float64).reshape(-1),
                                np.ascontiguousarray(Z, np.int32),
                                np.ascontiguousarray(atoms, np.int32), len(atoms),
                                np.ascontiguousarray(zlist, np.int32), len(zlist),
                                np.ascontiguousarray(templates, np.complex128),
                                kx, ky, xx, yy, nx, ny, bwcutoff_t, t)
    assert rc == 0, rc
    return t


def make_job(mtf=None, mtf_domain=None, mtf_rate=0.0, **kw):
    """orc_job; mtf / mtf_domain (equal lengths, Nyquist units) and mtf_rate
    select the detector MTF.  The arrays are kept alive on the job."""
    j = Job()
    defaults = dict(zmin=0.0, xmin=0.0, ymin=0.0, min_slice=1.0, corrected=0, cs3=0.0,
                    cs5=0.0, defocus=0.0, alpha_max=0.01, adf=1, det_inner=0.04,
                    det_outer=0.2, fem=0, dr=1 / np.sqrt(2), transmit_mode=0, adf_mode=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(j, k, v)
    if mtf is not None:
        m = np.ascontiguousarray(mtf, np.float64)
        d = np.ascontiguousarray(mtf_domain, np.float64)
        assert m.size == d.size
        j._mtf_keep = (m, d)
        j.mtf, j.mtf_domain, j.mtf_n, j.mtf_rate = m.ctypes.data, d.ctypes.data, m.size, mtf_rate
    return j


def mtf_2d(job, mtf, mtf_domain, rate):
    """orc_mtf_2d on the job's grid: complex [nx, ny]."""
    kx, ky, _, _ = domains(job.nx, job.ny, job.xperiod, job.yperiod, job.xmin, job.ymin)
    _, bwt = bwcutoffs(job.nx, job.ny, job.xperiod, job.yperiod)
    m = np.ascontiguousarray(mtf, np.float64)
    d = np.ascontiguousarray(mtf_domain, np.float64)
    out = np.zeros((job.nx, job.ny), np.complex128)
    assert lib().orc_mtf_2d(m, d, m.size, rate, kx, ky, job.nx, job.ny, job.xperiod,
                            job.yperiod, bwt, out.view(np.float64)) == 0
    return out


def apply_mtf(job, psi, mtf2d):
    """orc_apply_mtf on a copy of one k-space wave [nx, ny]."""
    kx, ky, _, _ = domains(job.nx, job.ny, job.xperiod, job.yperiod, job.xmin, job.ymin)
    _, bwt = bwcutoffs(job.nx, job.ny, job.xperiod, job.yperiod)
    lam = wavelength(job.voltage)[0]
    w = np.ascontiguousarray(psi, np.complex128).copy()
    assert lib().orc_apply_mtf(w.view(np.float64), np.ascontiguousarray(mtf2d).view(np.float64),
                               kx, ky, job.nx, job.ny, lam, job.alpha_max, bwt) == 0
    return w


def probe(job):
    kx, ky, xx, yy = domains(job.nx, job.ny, job.xperiod, job.yperiod, job.xmin, job.ymin)
    out = np.zeros((job.nx, job.ny), dtype=np.complex128)
    assert lib().orc_probe(C.byref(job), kx, ky, xx, yy, out) == 0
    return out


def probe_shifts(xx, yy, x_p, y_p):
    x_p = np.ascontiguousarray(x_p, np.float64)
    y_p = np.ascontiguousarray(y_p, np.float64)
    n = x_p.size * y_p.size
    sx, sy = np.zeros(n, np.int64), np.zeros(n, np.int64)
    lib().orc_probe_shifts(xx, xx.size, yy, yy.size, x_p, x_p.size, y_p, y_p.size, sx, sy)
    return sx, sy


def propagate_position(probe_x, sx, sy, t_stack, px, py, kx, ky, bwcutoff_t, transmit_mode=0):
    nx, ny = probe_x.shape
    psi = np.zeros((nx, ny), dtype=np.complex128)
    rc = lib().orc_propagate_position(np.ascontiguousarray(probe_x), int(sx), int(sy),
                                      t_stack.shape[0], np.ascontiguousarray(t_stack),
                                      np.ascontiguousarray(px), np.ascontiguousarray(py),
                                      kx, ky, nx, ny, bwcutoff_t, transmit_mode, psi)
    assert rc == 0, rc
    return psi


def adf(psi, kx, ky, k1, k2):
    return lib().orc_adf(np.ascontiguousarray(psi), kx, ky, kx.size, ky.size, k1, k2)


def fem_nbins(kx, ky, dr, bwcutoff_t):
    dk = C.c_double()
    nb = lib().orc_fem_nbins(kx, ky, dr, bwcutoff_t, C.byref(dk))
    return nb, dk.value


def fem_radial(psi, kx, ky, delta_k, nbins, lam, alpha_max):
    r, c = np.zeros(max(nbins, 1)), np.zeros(max(nbins, 1), np.int32)
    lib().orc_fem_radial(np.ascontiguousarray(psi), kx, ky, kx.size, ky.size, delta_k, nbins,
                         lam, alpha_max, r, c)
    return r[:nbins], c[:nbins]


def fem_finalize(sum_r, sum_r2, n_positions):
    sum_r = np.ascontiguousarray(sum_r, np.float64)
    v = np.zeros(sum_r.size)
    lib().orc_fem_finalize(sum_r, np.ascontiguousarray(sum_r2, np.float64), sum_r.size,
                           n_positions, v)
    return v


class Specimen:
    """orc_specimen: slices (lower bound, thickness), propagators and the
    FP64 transmission stack of tem.cpp:944-1276."""

    def __init__(self, job, q, Z):
        q = np.ascontiguousarray(q, np.float64).reshape(-1)
        Z = np.ascontiguousarray(Z, np.int32)
        self._s = _Spec()
        rc = lib().orc_specimen_build(C.byref(job), q, Z, Z.size, C.byref(self._s))
        assert rc == 0, rc
        ns, nx, ny = self._s.n_slices, job.nx, job.ny
        self.n_slices = ns
        self.t_stack = np.ctypeslib.as_array(self._s.t_stack, (ns * nx * ny * 2,)).view(
            np.complex128).reshape(ns, nx, ny).copy()
        self.px = np.ctypeslib.as_array(self._s.px, (ns * nx * 2,)).view(
            np.complex128).reshape(ns, nx).copy()
        self.py = np.ctypeslib.as_array(self._s.py, (ns * ny * 2,)).view(
            np.complex128).reshape(ns, ny).copy()
        self.lower_bound = np.ctypeslib.as_array(self._s.lower_bound, (ns,)).copy()
        self.thickness = np.ctypeslib.as_array(self._s.thickness, (ns,)).copy()

    def __del__(self):
        try:
            lib().orc_specimen_free(C.byref(self._s))
        except Exception:
            pass


def adfstem(job, spec, x_p, y_p, n_threads=0, exit_waves=False):
    """Whole raster (adfstem.cpp:1556-2866): returns (adf image [nxp, nyp] or
    None, V_omega [nbins] or None, exit psi_k [P, nx, ny] or None)."""
    x_p = np.ascontiguousarray(x_p, np.float64)
    y_p = np.ascontiguousarray(y_p, np.float64)
    P = x_p.size * y_p.size
    img = np.zeros(P) if job.adf else None
    v = None
    if job.fem:
        kx, ky, _, _ = domains(job.nx, job.ny, job.xperiod, job.yperiod)
        _, bwt = bwcutoffs(job.nx, job.ny, job.xperiod, job.yperiod)
        nb, _ = fem_nbins(kx, ky, job.dr, bwt)
        v = np.zeros(max(nb, 1))
    ex = np.zeros((P, job.nx, job.ny), np.complex128) if exit_waves else None
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    rc = lib().orc_adfstem(C.byref(job), C.byref(spec._s), x_p, x_p.size, y_p, y_p.size,
                           n_threads, ptr(img), ptr(v), ptr(ex))
    assert rc == 0, rc
    if img is not None:
        img = img.reshape(x_p.size, y_p.size)
    if v is not None:
        v = v[:nb]
    return img, v, ex


FEM_MODES = ("v_omega", "vbar_r", "v_re", "omega_vimage", "v_gt17", "v_rbar")


def adfstem_fem(job, spec, x_p, y_p, modes=FEM_MODES, n_threads=0, correlograph=False,
                correlograph_each=False, exit_waves=False):
    """Whole raster with the FEM modes of adfstem.cpp:2489-3605: returns
    (adf image or None, {mode: values [nbins]}); with exit_waves the dict
    also holds "exit_waves" [P, nx, ny], each position's wave after the
    detector MTF, the one the FEM statistics see."""
    x_p = np.ascontiguousarray(x_p, np.float64)
    y_p = np.ascontiguousarray(y_p, np.float64)
    P = x_p.size * y_p.size
    kx, ky, _, _ = domains(job.nx, job.ny, job.xperiod, job.yperiod)
    _, bwt = bwcutoffs(job.nx, job.ny, job.xperiod, job.yperiod)
    nb, _ = fem_nbins(kx, ky, job.dr, bwt)
    img = np.zeros(P) if job.adf else None
    outs = {m: np.zeros(max(nb, 1)) for m in modes}
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    corr = sqr = each = None
    if correlograph:
        _, dk = fem_nbins(kx, ky, job.dr, bwt)
        nphi = lib().orc_correlograph_nphi(dk, nb)
        corr = np.zeros(max(nb * nphi, 1))
        sqr = np.zeros(max(nb * nphi, 1))
        if correlograph_each:
            each = np.zeros(max(P * nb * nphi, 1))
    ex = np.zeros((P, job.nx, job.ny), np.complex128) if exit_waves else None
    rc = lib().orc_adfstem_fem(C.byref(job), C.byref(spec._s), x_p, x_p.size, y_p, y_p.size,
                               n_threads, ptr(img), *[ptr(outs.get(m)) for m in FEM_MODES],
                               ptr(corr), ptr(sqr), ptr(each), ptr(ex))
    assert rc == 0, rc
    if img is not None:
        img = img.reshape(x_p.size, y_p.size)
    res = {m: v[:nb] for m, v in outs.items()}
    if correlograph:
        # stored [ii + jj * nbins]: k bin fastest
        res["correlograph_avg"] = corr[:nb * nphi].reshape(nphi, nb)
        res["correlograph_sqr_sum"] = sqr[:nb * nphi].reshape(nphi, nb)
        if correlograph_each:
            res["correlograph_each"] = each[:P * nb * nphi].reshape(P, nphi, nb)
    if exit_waves:
        res["exit_waves"] = ex
    return img, res


def fem_modes(mag2, kx, ky, delta_k, nbins, lam, alpha_max, radial1, counts1):
    """orc_fem_modes on given per-position intensities mag2 [P, nx, ny]."""
    mag2 = np.ascontiguousarray(mag2, np.float64)
    P, nx, ny = mag2.shape
    outs = {m: np.zeros(max(nbins, 1)) for m in FEM_MODES[1:]}
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib().orc_fem_modes(mag2.reshape(-1), P, kx, ky, nx, ny, delta_k, nbins, lam, alpha_max,
                             np.ascontiguousarray(radial1, np.float64).reshape(-1),
                             np.ascontiguousarray(counts1, np.int32).reshape(-1),
                             *[ptr(outs[m]) for m in FEM_MODES[1:]])
    assert rc == 0
    return {m: v[:nbins] for m, v in outs.items()}


def initial_wave(probe_x, sx, sy):
    """orc_initial_wave: shifted probe, FFT / sqrt(NxNy)."""
    nx, ny = probe_x.shape
    out = np.zeros((nx, ny), np.complex128)
    assert lib().orc_initial_wave(np.ascontiguousarray(probe_x, np.complex128), sx, sy, nx, ny,
                                  out) == 0
    return out


def slice_step(psi_k, t, px, py, kx, ky, bwcutoff_t, transmit_mode=0):
    """orc_slice_step on a copy of psi_k [nx, ny]."""
    w = np.ascontiguousarray(psi_k, np.complex128).copy()
    nx, ny = w.shape
    assert lib().orc_slice_step(w, np.ascontiguousarray(t, np.complex128),
                                np.ascontiguousarray(px, np.complex128),
                                np.ascontiguousarray(py, np.complex128), kx, ky, nx, ny,
                                bwcutoff_t, transmit_mode) == 0
    return w
