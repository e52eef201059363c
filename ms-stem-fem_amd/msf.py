"""ctypes binding of libmsf.so (include/msf.h), the MI355X multislice
ADF-STEM / FEM path.

Mirrors the reference's call surface for this path: Specimen replaces the
set-up of tem.cpp:944-1276 (scatterer_pap_LUT, assign_atoms_to_slices_z_auto,
slice::update_propagator, slice::update_transmission_function), Context.
adfstem() replaces TEM_NS::adfstem (adfstem.hpp:66-121) for the default ADF
and V_omega modes.  Errors raise MsfError carrying the library's message, the
way the reference prints to cerr and returns EXIT_FAILURE.

There is no CPU fallback: loading fails loudly when libmsf.so is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsf.so")

TRANSMIT_REFERENCE, TRANSMIT_EXACT = 0, 1
PENDING = 1
ADF_REFERENCE, ADF_PHYSICAL = 0, 1

_ERRORS = {-1: "invalid argument", -2: "unsupported size", -3: "device error",
           -4: "out of device memory", -5: "position outside the domain"}


class MsfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"msf error {code} ({_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32),
                ("xperiod", C.c_double), ("yperiod", C.c_double), ("zperiod", C.c_double),
                ("xmin", C.c_double), ("ymin", C.c_double), ("zmin", C.c_double),
                ("voltage", C.c_double)]


class Optics(C.Structure):
    _fields_ = [("corrected", C.c_int32), ("cs3", C.c_double), ("cs5", C.c_double),
                ("defocus", C.c_double), ("alpha_max", C.c_double)]


class Detector(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("inner", C.c_double), ("outer", C.c_double),
                ("threshold_mode", C.c_int32)]


class Fem(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("dr", C.c_double), ("mtf", C.c_void_p),
                ("mtf_domain", C.c_void_p), ("mtf_n", C.c_uint32),
                ("mtf_sampling_rate", C.c_double)]


def make_fem(dr, mtf=None, mtf_domain=None, sampling_rate=0.0, enabled=1):
    """msf_fem with an optional detector MTF table (values and spatial
    frequencies in Nyquist units, as in the MTF file); the arrays are kept
    alive on the returned struct."""
    f = Fem(enabled, dr)
    if mtf is not None:
        m = np.ascontiguousarray(mtf, np.float64)
        d = np.ascontiguousarray(mtf_domain, np.float64)
        if m.size != d.size:
            raise ValueError("mtf and mtf_domain differ in length")
        f._keep = (m, d)
        f.mtf, f.mtf_domain, f.mtf_n, f.mtf_sampling_rate = (m.ctypes.data, d.ctypes.data,
                                                              m.size, sampling_rate)
    return f


class MtfTable(C.Structure):
    _fields_ = [("n", C.c_uint32), ("domain", C.POINTER(C.c_double)),
                ("mtf", C.POINTER(C.c_double))]


class Model(C.Structure):
    _fields_ = [("n_atoms", C.c_uint32), ("q", C.POINTER(C.c_double)),
                ("Z", C.POINTER(C.c_uint32)), ("xperiod", C.c_double), ("yperiod", C.c_double),
                ("zperiod", C.c_double), ("xmin", C.c_double), ("ymin", C.c_double),
                ("zmin", C.c_double)]


FEM_MODES = ("v_omega", "vbar_r", "v_re", "omega_vimage", "v_gt17", "v_rbar")


class FemOutputs(C.Structure):
    _fields_ = [(m, C.c_void_p) for m in FEM_MODES] + [("correlograph_avg", C.c_void_p),
                                                       ("correlograph_sqr_sum", C.c_void_p),
                                                       ("correlograph_each", C.c_void_p),
                                                       ("diffraction_mean_abs", C.c_void_p),
                                                       ("diffraction_mean_sqr", C.c_void_p),
                                                       ("variance_2d", C.c_void_p)]


class Job(C.Structure):
    _fields_ = [("grid", Grid), ("q_xyz", C.c_void_p), ("Z", C.c_void_p), ("n_atoms", C.c_uint32),
                ("min_slice", C.c_double), ("optics", Optics), ("x_p", C.c_void_p),
                ("n_xp", C.c_uint32), ("y_p", C.c_void_p), ("n_yp", C.c_uint32),
                ("detector", Detector), ("fem", Fem), ("adf_image", C.c_void_p),
                ("fem_out", FemOutputs)]


class RunJob(C.Structure):
    _fields_ = [("specimen", C.c_void_p), ("optics", Optics), ("x_p", C.c_void_p),
                ("n_xp", C.c_uint32), ("y_p", C.c_void_p), ("n_yp", C.c_uint32),
                ("detector", Detector), ("fem", Fem), ("adf_image", C.c_void_p),
                ("fem_out", FemOutputs)]


class Ticket:
    """A submitted job: its id and the output arrays the worker fills."""

    def __init__(self, job_id, img, outs, keep):
        self.id, self.img, self.outs, self._keep = job_id, img, outs, keep


class JobInfo(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("footprint", C.c_uint64),
                ("slot", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("running", C.c_uint32), ("max_running", C.c_uint32),
                ("queued", C.c_uint64), ("jobs_done", C.c_uint64), ("mem_limit", C.c_uint64),
                ("admitted_bytes", C.c_uint64), ("peak_admitted_bytes", C.c_uint64),
                ("specimen_builds", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("groups", C.c_uint64), ("grouped_jobs", C.c_uint64),
                ("group_launches", C.c_uint64)]


class Config(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("transmit_mode", C.c_int32),
                ("profile_kernels", C.c_int32), ("col_tile", C.c_int32),
                ("no_graph", C.c_int32), ("chains", C.c_int32), ("autotune", C.c_int32),
                ("fft_plan", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("setup_ms", C.c_double), ("slice_loop_ms", C.c_double),
                ("total_ms", C.c_double), ("probe_slices", C.c_uint64),
                ("batch", C.c_uint32), ("n_batches", C.c_uint32),
                ("row_pass_us", C.c_double), ("col_pass_us", C.c_double),
                ("first_row_pass_us", C.c_double), ("last_col_pass_us", C.c_double),
                ("reduce_us", C.c_double), ("row_pass_count", C.c_uint64),
                ("col_pass_count", C.c_uint64), ("row_pass_bytes", C.c_double),
                ("col_pass_bytes", C.c_double), ("graph", C.c_int32), ("chains", C.c_int32),
                ("col_tile", C.c_int32), ("autotuned", C.c_int32), ("row_fft_plan", C.c_int32),
                ("col_fft_plan", C.c_int32), ("fused_launches", C.c_uint64),
                ("group_runs", C.c_uint32), ("group_launches", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# device bytes of the static array the allocation-free run works in
# (MSF_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 256 << 20

IO_EXPORTS = ["msf_io_last_error", "msf_element_z", "msf_guess_z", "msf_read_xyz",
              "msf_read_lammps_data", "msf_model_free", "msf_write_tiff16",
              "msf_write_variance_txt", "msf_write_f64", "msf_write_correlograph_txt",
              "msf_read_mtf", "msf_mtf_table_free"]

EXPORTS = ["msf_create", "msf_destroy", "msf_last_error", "msf_set_config", "msf_get_stats",
           "msf_supported_length", "msf_wavelength", "msf_scherzer_defocus",
           "msf_scherzer_alphamax", "msf_specimen_build", "msf_specimen_free",
           "msf_specimen_slices", "msf_specimen_slice_info", "msf_specimen_transmission",
           "msf_probe_realspace", "msf_fem_bins", "msf_run_stem", "msf_exit_waves",
           "msf_scherzer_cs3", "msf_probe_shifts", "msf_exit_waves_shifted", "msf_run_stem_fem",
           "msf_submit", "msf_poll", "msf_wait", "msf_specimen_propagator", "msf_probe_k",
           "msf_slice_step_batch", "msf_correlograph_bins", "msf_adf_reduce",
           "msf_fem_accumulate", "msf_mtf_2d", "msf_apply_mtf", "msf_wait_info",
           "msf_serve_configure", "msf_serve_get_stats", "msf_job_footprint",
           "msf_run_stem_noalloc", "msf_copy_rate", "msf_submit_run", "msf_tuning_save",
           "msf_tuning_load", "msf_tuning_clear", "msf_run_group",
           "msf_serve_group"]

_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    """Load libmsf.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C ms-stem-fem_amd/csrc); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.msf_create.argtypes = [C.POINTER(vp), C.c_int]
        L.msf_destroy.argtypes = [vp]
        L.msf_last_error.argtypes = [vp]
        L.msf_last_error.restype = C.c_char_p
        L.msf_set_config.argtypes = [vp, C.POINTER(Config)]
        L.msf_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.msf_supported_length.argtypes = [C.c_uint32]
        L.msf_wavelength.argtypes = [C.c_double, _dp, _dp]
        L.msf_scherzer_defocus.argtypes = [C.c_int32, C.c_double, C.c_double]
        L.msf_scherzer_defocus.restype = C.c_double
        L.msf_scherzer_alphamax.argtypes = [C.c_int32, C.c_double, C.c_double]
        L.msf_scherzer_alphamax.restype = C.c_double
        L.msf_specimen_build.argtypes = [vp, C.POINTER(Grid), _dp, C.POINTER(C.c_uint32),
                                         C.c_uint32, C.c_double, C.POINTER(vp)]
        L.msf_specimen_free.argtypes = [vp]
        L.msf_specimen_slices.argtypes = [vp]
        L.msf_specimen_slices.restype = C.c_uint32
        L.msf_specimen_slice_info.argtypes = [vp, _dp, _dp, C.POINTER(C.c_uint32)]
        L.msf_specimen_transmission.argtypes = [vp, C.c_uint32, _dp]
        L.msf_probe_realspace.argtypes = [vp, C.POINTER(Grid), C.POINTER(Optics), _dp]
        L.msf_fem_bins.argtypes = [C.POINTER(Grid), C.c_double, C.POINTER(C.c_uint32), _dp]
        L.msf_correlograph_bins.argtypes = [C.POINTER(Grid), C.c_double, C.POINTER(C.c_uint32),
                                            _dp, _dp]
        L.msf_run_stem.argtypes = [vp, vp, C.POINTER(Optics), C.POINTER(Detector),
                                   C.POINTER(Fem), _dp, C.c_uint32, _dp, C.c_uint32, _dp, _dp]
        L.msf_copy_rate.argtypes = [vp, C.c_uint64, C.c_uint32, _dp]
        L.msf_tuning_save.argtypes = [C.c_char_p]
        L.msf_tuning_load.argtypes = [C.c_char_p, C.POINTER(C.c_uint32)]
        L.msf_tuning_clear.argtypes = []
        L.msf_run_stem_noalloc.argtypes = [vp, vp, C.POINTER(Optics), C.POINTER(Detector),
                                           C.POINTER(Fem), _dp, C.c_uint32, _dp, C.c_uint32, _dp,
                                           _dp]
        L.msf_exit_waves.argtypes = [vp, vp, C.POINTER(Optics), _dp, C.c_uint32, _dp,
                                     C.c_uint32, C.POINTER(C.c_float)]
        L.msf_probe_shifts.argtypes = [C.POINTER(Grid), _dp, C.c_uint32, _dp, C.c_uint32,
                                       C.POINTER(C.c_int32)]
        L.msf_run_stem_fem.argtypes = [vp, vp, C.POINTER(Optics), C.POINTER(Detector),
                                       C.POINTER(Fem), _dp, C.c_uint32, _dp, C.c_uint32, _dp,
                                       C.POINTER(FemOutputs)]
        L.msf_exit_waves_shifted.argtypes = [vp, vp, C.POINTER(Optics), C.POINTER(C.c_int32),
                                             C.c_uint32, C.POINTER(C.c_float)]
        L.msf_specimen_propagator.argtypes = [vp, C.c_uint32, _dp, _dp]
        L.msf_probe_k.argtypes = [vp, C.POINTER(Grid), C.POINTER(Optics), C.c_int32, C.c_int32,
                                  _dp]
        L.msf_slice_step_batch.argtypes = [vp, vp, C.c_uint32, _dp, C.c_uint32]
        L.msf_adf_reduce.argtypes = [vp, C.POINTER(Grid), C.POINTER(Detector), _dp, C.c_uint32,
                                     _dp]
        L.msf_fem_accumulate.argtypes = [vp, C.POINTER(Grid), C.POINTER(Fem), C.c_double, _dp,
                                         C.c_uint32, _dp, C.POINTER(C.c_int32), _dp]
        L.msf_mtf_2d.argtypes = [vp, C.POINTER(Grid), C.POINTER(Fem), _dp]
        L.msf_apply_mtf.argtypes = [vp, C.POINTER(Grid), C.POINTER(Fem), C.c_double, _dp,
                                    C.c_uint32]
        L.msf_read_mtf.argtypes = [C.c_char_p, C.POINTER(MtfTable)]
        L.msf_mtf_table_free.argtypes = [C.POINTER(MtfTable)]
        L.msf_submit.argtypes = [vp, C.POINTER(Job), C.POINTER(C.c_uint64)]
        L.msf_submit_run.argtypes = [vp, C.POINTER(RunJob), C.POINTER(C.c_uint64)]
        L.msf_run_group.argtypes = [vp, C.POINTER(RunJob), C.c_uint32, C.POINTER(C.c_int32)]
        L.msf_poll.argtypes = [vp, C.c_uint64]
        L.msf_wait.argtypes = [vp, C.c_uint64]
        L.msf_wait_info.argtypes = [vp, C.c_uint64, C.POINTER(JobInfo)]
        L.msf_serve_configure.argtypes = [vp, C.c_uint32, C.c_uint64]
        L.msf_serve_group.argtypes = [vp, C.c_uint32]
        L.msf_serve_get_stats.argtypes = [vp, C.POINTER(ServeStats)]
        L.msf_job_footprint.argtypes = [C.POINTER(Job)]
        L.msf_job_footprint.restype = C.c_uint64
        L.msf_scherzer_cs3.argtypes = [C.c_double, C.c_double]
        L.msf_scherzer_cs3.restype = C.c_double
        # host-side readers and writers (include/msf_io.h)
        L.msf_io_last_error.restype = C.c_char_p
        L.msf_element_z.argtypes = [C.c_char_p]
        L.msf_element_z.restype = C.c_uint32
        L.msf_guess_z.argtypes = [C.c_double, C.POINTER(C.c_uint32)]
        L.msf_read_xyz.argtypes = [C.c_char_p, C.POINTER(Model)]
        L.msf_read_lammps_data.argtypes = [C.c_char_p, C.POINTER(Model)]
        L.msf_model_free.argtypes = [C.POINTER(Model)]
        L.msf_model_free.restype = None
        L.msf_write_tiff16.argtypes = [C.c_char_p, _dp, C.c_uint32, C.c_uint32, C.c_double,
                                       C.c_double, C.c_int, C.c_uint32, C.c_double, C.c_double]
        L.msf_write_variance_txt.argtypes = [C.c_char_p, _dp, _dp, C.c_uint32]
        L.msf_write_f64.argtypes = [C.c_char_p, _dp, C.c_uint32, C.c_uint32]
        L.msf_write_correlograph_txt.argtypes = [C.c_char_p, _dp, _dp, C.c_uint32, _dp,
                                                 C.c_uint32]
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def wavelength(voltage):
    lam, gam = C.c_double(), C.c_double()
    rc = lib().msf_wavelength(voltage, C.byref(lam), C.byref(gam))
    if rc:
        raise MsfError(rc, "msf_wavelength")
    return lam.value, gam.value


def scherzer_defocus(lam, cs, corrected=0):
    return lib().msf_scherzer_defocus(corrected, lam, cs)


def scherzer_alphamax(lam, cs, corrected=0):
    return lib().msf_scherzer_alphamax(corrected, lam, cs)


def scherzer_cs3(lam, cs5):
    return lib().msf_scherzer_cs3(lam, cs5)


def _io_check(rc, what):
    if rc:
        raise MsfError(rc, f"{what}: {lib().msf_io_last_error().decode()}")


def probe_shifts(grid, x_p, y_p):
    """(sx, sy) of every raster position, shape [n_xp * n_yp, 2] (host only)."""
    x_p = np.ascontiguousarray(x_p, np.float64)
    y_p = np.ascontiguousarray(y_p, np.float64)
    out = np.zeros((x_p.size * y_p.size, 2), np.int32)
    rc = lib().msf_probe_shifts(C.byref(grid), _p(x_p), x_p.size, _p(y_p), y_p.size,
                                _p(out, C.c_int32))
    if rc:
        raise MsfError(rc, "msf_probe_shifts")
    return out


def element_z(name):
    return int(lib().msf_element_z(name.encode()))


def guess_z(mass):
    z = C.c_uint32()
    _io_check(lib().msf_guess_z(mass, C.byref(z)), "msf_guess_z")
    return z.value


def read_model(path):
    """Reads an xyz file (extension "xyz") or a LAMMPS data file (any other
    extension), as the reference's main does (tem.cpp:270-330).  Returns
    (q [n, 3], Z [n], (xperiod, yperiod, zperiod))."""
    m = Model()
    xyz = path[path.rfind(".") + 1:] == "xyz"   # find_last_of(".") + 1, npos + 1 == 0
    fn = lib().msf_read_xyz if xyz else lib().msf_read_lammps_data
    _io_check(fn(os.fsencode(path), C.byref(m)), "read " + path)
    try:
        n = m.n_atoms
        q = np.ctypeslib.as_array(m.q, shape=(3 * n,)).reshape(n, 3).copy() if n else np.zeros((0, 3))
        Z = np.ctypeslib.as_array(m.Z, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
        return q, Z, (m.xperiod, m.yperiod, m.zperiod)
    finally:
        lib().msf_model_free(C.byref(m))


def read_mtf(path):
    """msf_read_mtf: (domain [n], mtf [n]) as read from an MTF file."""
    t = MtfTable()
    _io_check(lib().msf_read_mtf(os.fsencode(path), C.byref(t)), "read " + path)
    try:
        return (np.ctypeslib.as_array(t.domain, shape=(t.n,)).copy(),
                np.ctypeslib.as_array(t.mtf, shape=(t.n,)).copy())
    finally:
        lib().msf_mtf_table_free(C.byref(t))


def write_tiff16(path, data, vmin=None, vmax=None, logscale=False, resolution_unit=3,
                 x_resolution=1.0, y_resolution=1.0):
    d = np.ascontiguousarray(data, np.float64)
    vmin = float(d.min()) if vmin is None else vmin
    vmax = float(d.max()) if vmax is None else vmax
    _io_check(lib().msf_write_tiff16(os.fsencode(path), _p(d), d.shape[0], d.shape[1], vmin, vmax,
                                     1 if logscale else 0, resolution_unit, x_resolution,
                                     y_resolution), "msf_write_tiff16")


def write_variance_txt(path, values, lower_edges):
    v = np.ascontiguousarray(values, np.float64)
    e = np.ascontiguousarray(lower_edges, np.float64)
    _io_check(lib().msf_write_variance_txt(os.fsencode(path), _p(v), _p(e), v.size),
              "msf_write_variance_txt")


def write_f64(path, data):
    d = np.ascontiguousarray(data, np.float64)
    d2 = d.reshape(d.shape[0], -1) if d.ndim > 1 else d.reshape(1, -1)
    _io_check(lib().msf_write_f64(os.fsencode(path), _p(d2), d2.shape[0], d2.shape[1]),
              "msf_write_f64")


def tuning_save(path):
    """msf_tuning_save: the process-wide tuning table as a text file."""
    if lib().msf_tuning_save(str(path).encode()):
        raise MsfError(-1, f"msf_tuning_save({path})")


def tuning_load(path):
    """msf_tuning_load: merges a tuning file; returns its entry count."""
    n = C.c_uint32()
    if lib().msf_tuning_load(str(path).encode(), C.byref(n)):
        raise MsfError(-1, f"msf_tuning_load({path})")
    return n.value


def tuning_clear():
    lib().msf_tuning_clear()


def supported_length(n):
    return bool(lib().msf_supported_length(n))


def make_grid(nx, ny, xperiod, yperiod, zperiod, voltage, xmin=0.0, ymin=0.0, zmin=0.0):
    return Grid(nx, ny, xperiod, yperiod, zperiod, xmin, ymin, zmin, voltage)


def fem_bins(grid, dr):
    n = C.c_uint32()
    rc = lib().msf_fem_bins(C.byref(grid), dr, C.byref(n), None)
    if rc:
        raise MsfError(rc, "msf_fem_bins")
    edges = np.zeros(max(n.value, 1))
    lib().msf_fem_bins(C.byref(grid), dr, C.byref(n), _p(edges))
    return edges[:n.value]


def job_footprint(grid, q, Z, min_slice, x_p, y_p, fem=None):
    """msf_job_footprint: the serving engine's admission estimate in bytes."""
    q = np.ascontiguousarray(q, np.float64).reshape(-1)
    Z = np.ascontiguousarray(Z, np.uint32)
    x_p = np.ascontiguousarray(x_p, np.float64)
    y_p = np.ascontiguousarray(y_p, np.float64)
    job = Job(grid, q.ctypes.data, Z.ctypes.data, Z.size, min_slice, Optics(), x_p.ctypes.data,
              x_p.size, y_p.ctypes.data, y_p.size, Detector(), fem if fem is not None else Fem(),
              None, FemOutputs())
    return lib().msf_job_footprint(C.byref(job))


def correlograph_bins(grid, dr):
    """msf_correlograph_bins: (n_phi, k boundaries [n_bins + 1], phi boundaries [n_phi + 1])."""
    n = C.c_uint32()
    nb = fem_bins(grid, dr).size
    k = np.zeros(nb + 1)
    rc = lib().msf_correlograph_bins(C.byref(grid), dr, C.byref(n), _p(k), None)
    if rc:
        raise MsfError(rc, "msf_correlograph_bins")
    phi = np.zeros(n.value + 1)
    lib().msf_correlograph_bins(C.byref(grid), dr, C.byref(n), _p(k), _p(phi))
    return n.value, k, phi


def write_correlograph_txt(path, data, k_edges, phi_edges):
    """msf_write_correlograph_txt; data [n_phi, n_k] (k fastest)."""
    data = np.ascontiguousarray(data, np.float64)
    n_phi, n_k = data.shape
    k = np.ascontiguousarray(k_edges, np.float64)
    phi = np.ascontiguousarray(phi_edges, np.float64)
    assert k.size >= n_k and phi.size >= n_phi
    rc = lib().msf_write_correlograph_txt(str(path).encode(), _p(data), _p(k), n_k, _p(phi), n_phi)
    if rc:
        raise MsfError(rc, lib().msf_io_last_error().decode())


def _fem_outputs(spec, fem, P, modes=FEM_MODES, correlograph=False, correlograph_each=False,
                 pixel_images=False):
    """Output arrays of msf_run_stem_fem for a specimen and FEM settings:
    (msf_fem_outputs, the arrays, a function giving the result dict)."""
    nb = fem_bins(spec.grid, fem.dr).size if fem.enabled else 0
    outs = {m: np.zeros(max(nb, 1)) for m in modes}
    ptrs = [outs[m].ctypes.data if m in outs else None for m in FEM_MODES]
    nphi = 0
    if correlograph:
        nphi = correlograph_bins(spec.grid, fem.dr)[0]
        for m in ("correlograph_avg", "correlograph_sqr_sum"):
            outs[m] = np.zeros(max(nb * nphi, 1))
            ptrs.append(outs[m].ctypes.data)
        if correlograph_each:
            outs["correlograph_each"] = np.zeros(max(P * nb * nphi, 1))
            ptrs.append(outs["correlograph_each"].ctypes.data)
    fo = FemOutputs(*ptrs)
    img_names = ("diffraction_mean_abs", "diffraction_mean_sqr", "variance_2d")
    if pixel_images:
        for m in img_names:
            outs[m] = np.zeros((spec.grid.nx, spec.grid.ny))
            setattr(fo, m, outs[m].ctypes.data)

    def finish():
        res = {m: v[:nb] for m, v in outs.items() if m in FEM_MODES}
        if pixel_images:
            res.update({m: outs[m] for m in img_names})
        if correlograph:
            for m in ("correlograph_avg", "correlograph_sqr_sum"):
                res[m] = outs[m][:nb * nphi].reshape(nphi, nb)
            if correlograph_each:
                res["correlograph_each"] = outs["correlograph_each"][:P * nb * nphi].reshape(
                    P, nphi, nb)
        return res
    return fo, outs, finish


class Context:
    """One device context (msf_create); one per host thread."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().msf_create(C.byref(self._h), device)
        if rc:
            raise MsfError(rc, f"msf_create(device={device})")

    def close(self):
        if self._h:
            lib().msf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise MsfError(rc, f"{what}: {lib().msf_last_error(self._h).decode()}")

    def configure(self, batch=0, transmit_mode=TRANSMIT_REFERENCE, profile_kernels=False,
                  col_tile=0, graph=True, chains=0, autotune=0, fft_plan=0):
        """fft_plan: 1 primary radix plans, 2 the alternate plans of lengths
        256, 512 and 1024 (16 values per lane, one fewer stage), 0 automatic
        (the alternate at 256 and 1024, or the autotuner's choice).
        autotune: 0 (or False) automatic, for runs of >= 16384 probe-slices,
        launch shapes only; 1 (or True) every run, radix plans too with
        fft_plan 0; -1 the fixed rules."""
        cfg = Config(batch, transmit_mode, 1 if profile_kernels else 0, col_tile,
                     0 if graph else 1, chains, int(autotune), fft_plan)
        self._check(lib().msf_set_config(self._h, C.byref(cfg)), "msf_set_config")

    def copy_rate(self, nbytes=1 << 30, reps=5):
        """msf_copy_rate: GB/s (read + write) of the library's streaming copy
        kernel over nbytes, best of reps."""
        g = C.c_double()
        self._check(lib().msf_copy_rate(self._h, nbytes, reps, C.byref(g)), "msf_copy_rate")
        return g.value

    def stats(self):
        s = Stats()
        self._check(lib().msf_get_stats(self._h, C.byref(s)), "msf_get_stats")
        return s.as_dict()

    def specimen(self, grid, q, Z, min_slice):
        return Specimen(self, grid, q, Z, min_slice)

    def probe_realspace(self, grid, optics):
        out = np.zeros((grid.nx, grid.ny), np.complex128)
        self._check(lib().msf_probe_realspace(self._h, C.byref(grid), C.byref(optics),
                                              _p(out.view(np.float64))), "msf_probe_realspace")
        return out

    def adfstem(self, spec, optics, x_p, y_p, detector=None, fem=None):
        """Returns (adf image [n_xp, n_yp] or None, V_omega [n_bins] or None)."""
        x_p = np.ascontiguousarray(x_p, np.float64)
        y_p = np.ascontiguousarray(y_p, np.float64)
        det = detector if detector is not None else Detector(0, 0.0, 0.0, 0)
        fm = fem if fem is not None else Fem(0, 0.0)
        img = np.zeros(x_p.size * y_p.size) if det.enabled else None
        v = None
        if fm.enabled:
            v = np.zeros(max(fem_bins(spec.grid, fm.dr).size, 1))
        self._check(lib().msf_run_stem(self._h, spec._h, C.byref(optics), C.byref(det),
                                       C.byref(fm), _p(x_p), x_p.size, _p(y_p), y_p.size,
                                       _p(img), _p(v)), "msf_run_stem")
        if img is not None:
            img = img.reshape(x_p.size, y_p.size)
        if v is not None:
            v = v[:fem_bins(spec.grid, fm.dr).size]
        return img, v

    def adfstem_noalloc(self, spec, optics, x_p, y_p, detector=None, fem=None):
        """msf_run_stem_noalloc: adfstem() with every device buffer in the
        library's static scratch array (no device allocation)."""
        x_p = np.ascontiguousarray(x_p, np.float64)
        y_p = np.ascontiguousarray(y_p, np.float64)
        det = detector if detector is not None else Detector(0, 0.0, 0.0, 0)
        fm = fem if fem is not None else Fem(0, 0.0)
        img = np.zeros(x_p.size * y_p.size) if det.enabled else None
        nb = fem_bins(spec.grid, fm.dr).size if fm.enabled else 0
        v = np.zeros(max(nb, 1)) if fm.enabled else None
        self._check(lib().msf_run_stem_noalloc(self._h, spec._h, C.byref(optics), C.byref(det),
                                               C.byref(fm), _p(x_p), x_p.size, _p(y_p),
                                               y_p.size, _p(img), _p(v)), "msf_run_stem_noalloc")
        return (img.reshape(x_p.size, y_p.size) if img is not None else None,
                v[:nb] if v is not None else None)

    def adfstem_fem(self, spec, optics, x_p, y_p, fem, modes=FEM_MODES, detector=None,
                    correlograph=False, correlograph_each=False, pixel_images=False):
        """msf_run_stem_fem: returns (adf image or None, {mode: values [n_bins]});
        with correlograph, also "correlograph_avg" and "correlograph_sqr_sum"
        shaped [n_phi, n_bins]; with pixel_images, "diffraction_mean_abs",
        "diffraction_mean_sqr" and "variance_2d" shaped [nx, ny]."""
        x_p = np.ascontiguousarray(x_p, np.float64)
        y_p = np.ascontiguousarray(y_p, np.float64)
        det = detector if detector is not None else Detector(0, 0.0, 0.0, 0)
        img = np.zeros(x_p.size * y_p.size) if det.enabled else None
        fo, outs, finish = _fem_outputs(spec, fem, x_p.size * y_p.size, modes, correlograph,
                                        correlograph_each, pixel_images)
        self._check(lib().msf_run_stem_fem(self._h, spec._h, C.byref(optics), C.byref(det),
                                           C.byref(fem), _p(x_p), x_p.size, _p(y_p), y_p.size,
                                           _p(img), C.byref(fo)), "msf_run_stem_fem")
        if img is not None:
            img = img.reshape(x_p.size, y_p.size)
        return img, finish()

    def submit(self, grid, q, Z, min_slice, optics, x_p, y_p, detector=None, fem=None,
               modes=("v_omega",)):
        """msf_submit: queue a whole job (specimen build + run) on the
        context's serving engine; returns a Ticket for poll() / wait()."""
        q = np.ascontiguousarray(q, np.float64).reshape(-1)
        Z = np.ascontiguousarray(Z, np.uint32)
        x_p = np.ascontiguousarray(x_p, np.float64)
        y_p = np.ascontiguousarray(y_p, np.float64)
        det = detector if detector is not None else Detector(0, 0.0, 0.0, 0)
        fm = fem if fem is not None else Fem(0, 0.0)
        img = np.zeros(x_p.size * y_p.size) if det.enabled else None
        outs, nb = {}, 0
        if fm.enabled:
            nb = fem_bins(grid, fm.dr).size
            outs = {m: np.zeros(max(nb, 1)) for m in modes}
        fo = FemOutputs(*[outs[m].ctypes.data if m in outs else None for m in FEM_MODES])
        job = Job(grid, q.ctypes.data, Z.ctypes.data, Z.size, min_slice, optics, x_p.ctypes.data,
                  x_p.size, y_p.ctypes.data, y_p.size, det, fm,
                  img.ctypes.data if img is not None else None, fo)
        jid = C.c_uint64()
        self._check(lib().msf_submit(self._h, C.byref(job), C.byref(jid)), "msf_submit")
        t = Ticket(jid.value, img, outs, (q, Z, x_p, y_p, nb))
        t.footprint = lib().msf_job_footprint(C.byref(job))
        return t

    def run_group(self, runs, modes=("v_omega",)):
        """msf_run_group: runs on resident specimens of one grid size with
        shared slice-pass launches.  `runs`: dicts with spec, optics, x_p,
        y_p and optionally detector, fem, modes, correlograph,
        correlograph_each, pixel_images (as adfstem_fem).  Returns (results,
        status): per run (adf image [n_xp, n_yp] or None, {output: values})
        as adfstem_fem, and its status code."""
        keep, jobs, res = [], [], []
        for r in runs:
            x_p = np.ascontiguousarray(r["x_p"], np.float64)
            y_p = np.ascontiguousarray(r["y_p"], np.float64)
            det = r.get("detector") or Detector(0, 0.0, 0.0, 0)
            fm = r.get("fem") or Fem(0, 0.0)
            img = np.zeros(x_p.size * y_p.size) if det.enabled else None
            fo, outs, finish = _fem_outputs(r["spec"], fm, x_p.size * y_p.size,
                                            r.get("modes", modes) if fm.enabled else (),
                                            r.get("correlograph", False),
                                            r.get("correlograph_each", False),
                                            r.get("pixel_images", False))
            jobs.append(RunJob(r["spec"]._h, r["optics"], x_p.ctypes.data, x_p.size,
                               y_p.ctypes.data, y_p.size, det, fm,
                               img.ctypes.data if img is not None else None, fo))
            keep.append((x_p, y_p, outs))
            res.append((img, finish, x_p.size, y_p.size))
        arr = (RunJob * max(len(jobs), 1))(*jobs)
        status = (C.c_int32 * max(len(jobs), 1))()
        rc = lib().msf_run_group(self._h, arr, len(jobs), status)
        out = [((img.reshape(nxp, nyp) if img is not None else None), finish())
               for img, finish, nxp, nyp in res]
        st = [int(status[k]) for k in range(len(jobs))]
        if rc and all(x != 0 for x in st):
            self._check(rc, "msf_run_group")
        return out, st

    def submit_run(self, spec, optics, x_p, y_p, detector=None, fem=None, modes=("v_omega",)):
        """msf_submit_run: queue a run on a resident specimen of this context
        on the serving engine; returns a Ticket for poll() / wait()."""
        x_p = np.ascontiguousarray(x_p, np.float64)
        y_p = np.ascontiguousarray(y_p, np.float64)
        det = detector if detector is not None else Detector(0, 0.0, 0.0, 0)
        fm = fem if fem is not None else Fem(0, 0.0)
        img = np.zeros(x_p.size * y_p.size) if det.enabled else None
        outs, nb = {}, 0
        if fm.enabled:
            nb = fem_bins(spec.grid, fm.dr).size
            outs = {m: np.zeros(max(nb, 1)) for m in modes}
        fo = FemOutputs(*[outs[m].ctypes.data if m in outs else None for m in FEM_MODES])
        job = RunJob(spec._h, optics, x_p.ctypes.data, x_p.size, y_p.ctypes.data, y_p.size, det,
                     fm, img.ctypes.data if img is not None else None, fo)
        jid = C.c_uint64()
        self._check(lib().msf_submit_run(self._h, C.byref(job), C.byref(jid)), "msf_submit_run")
        return Ticket(jid.value, img, outs, (spec, None, x_p, y_p, nb))

    def poll(self, ticket):
        return lib().msf_poll(self._h, ticket.id)

    def wait(self, ticket, info=False):
        """Returns (adf image [n_xp, n_yp] or None, {mode: values}), and with
        info the job's timings (msf_wait_info) as a third element."""
        ji = JobInfo()
        self._check(lib().msf_wait_info(self._h, ticket.id, C.byref(ji)), "msf_wait")
        img = ticket.img
        nx, ny, nb = ticket._keep[2].size, ticket._keep[3].size, ticket._keep[4]
        res = ((img.reshape(nx, ny) if img is not None else None),
               {m: v[:nb] for m, v in ticket.outs.items()})
        if info:
            return res + ({k: getattr(ji, k) for k, _ in JobInfo._fields_},)
        return res

    def serve_configure(self, slots=0, mem_limit=0):
        """msf_serve_configure (before the first submit)."""
        self._check(lib().msf_serve_configure(self._h, slots, mem_limit), "msf_serve_configure")

    def serve_group(self, max_jobs):
        """msf_serve_group (before the first submit): run up to max_jobs queued
        jobs of one grid size as a group with shared slice-pass launches."""
        self._check(lib().msf_serve_group(self._h, max_jobs), "msf_serve_group")

    def serve_stats(self):
        st = ServeStats()
        self._check(lib().msf_serve_get_stats(self._h, C.byref(st)), "msf_serve_get_stats")
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def probe_k(self, grid, optics, sx, sy):
        out = np.zeros((grid.nx, grid.ny), np.complex128)
        self._check(lib().msf_probe_k(self._h, C.byref(grid), C.byref(optics), sx, sy,
                                      _p(out.view(np.float64))), "msf_probe_k")
        return out

    def adf_reduce(self, grid, detector, psi_k):
        """msf_adf_reduce on k-space waves [count, nx, ny] (complex): ADF per wave."""
        w = np.ascontiguousarray(psi_k, np.complex128).reshape(-1, grid.nx, grid.ny)
        out = np.zeros(w.shape[0])
        self._check(lib().msf_adf_reduce(self._h, C.byref(grid), C.byref(detector),
                                         _p(w.view(np.float64)), w.shape[0], _p(out)),
                    "msf_adf_reduce")
        return out

    def fem_accumulate(self, grid, fem, alpha_max, psi_k):
        """msf_fem_accumulate on k-space waves [count, nx, ny] taken as one
        raster: (ring sums [count, n_bins], counts [n_bins], V_omega [n_bins])."""
        w = np.ascontiguousarray(psi_k, np.complex128).reshape(-1, grid.nx, grid.ny)
        nb = fem_bins(grid, fem.dr).size
        radial = np.zeros(max(w.shape[0] * nb, 1))
        counts = np.zeros(max(nb, 1), np.int32)
        v = np.zeros(max(nb, 1))
        self._check(lib().msf_fem_accumulate(self._h, C.byref(grid), C.byref(fem), alpha_max,
                                             _p(w.view(np.float64)), w.shape[0], _p(radial),
                                             counts.ctypes.data_as(C.POINTER(C.c_int32)), _p(v)),
                    "msf_fem_accumulate")
        return radial[:w.shape[0] * nb].reshape(w.shape[0], nb), counts[:nb], v[:nb]

    def mtf_2d(self, grid, fem):
        """msf_mtf_2d: the 2-D detector MTF [nx, ny] (complex)."""
        out = np.zeros((grid.nx, grid.ny), np.complex128)
        self._check(lib().msf_mtf_2d(self._h, C.byref(grid), C.byref(fem),
                                     _p(out.view(np.float64))), "msf_mtf_2d")
        return out

    def apply_mtf(self, grid, fem, alpha_max, psi_k):
        """msf_apply_mtf on a copy of k-space waves [count, nx, ny] (complex)."""
        w = np.ascontiguousarray(psi_k, np.complex128).reshape(-1, grid.nx, grid.ny).copy()
        self._check(lib().msf_apply_mtf(self._h, C.byref(grid), C.byref(fem), alpha_max,
                                        _p(w.view(np.float64)), w.shape[0]), "msf_apply_mtf")
        return w.reshape(np.shape(psi_k))

    def slice_step(self, spec, slice_index, psi_k):
        """msf_slice_step_batch on a copy of psi_k [count, nx, ny] (complex)."""
        w = np.ascontiguousarray(psi_k, np.complex128).copy()
        count = w.shape[0] if w.ndim == 3 else 1
        self._check(lib().msf_slice_step_batch(self._h, spec._h, slice_index,
                                               _p(w.view(np.float64)), count),
                    "msf_slice_step_batch")
        return w

    def exit_waves_shifted(self, spec, optics, shifts):
        sh = np.ascontiguousarray(shifts, np.int32).reshape(-1, 2)
        out = np.zeros((sh.shape[0], spec.grid.nx, spec.grid.ny), np.complex64)
        self._check(lib().msf_exit_waves_shifted(self._h, spec._h, C.byref(optics),
                                                 _p(sh, C.c_int32), sh.shape[0],
                                                 _p(out.view(np.float32), C.c_float)),
                    "msf_exit_waves_shifted")
        return out

    def exit_waves(self, spec, optics, x_p, y_p):
        x_p = np.ascontiguousarray(x_p, np.float64)
        y_p = np.ascontiguousarray(y_p, np.float64)
        out = np.zeros((x_p.size * y_p.size, spec.grid.nx, spec.grid.ny), np.complex64)
        self._check(lib().msf_exit_waves(self._h, spec._h, C.byref(optics), _p(x_p), x_p.size,
                                         _p(y_p), y_p.size, _p(out.view(np.float32), C.c_float)),
                    "msf_exit_waves")
        return out


class Specimen:
    """Device-resident slices of one specimen (msf_specimen_build)."""

    def __init__(self, ctx, grid, q, Z, min_slice):
        self.ctx = ctx
        self.grid = grid
        q = np.ascontiguousarray(q, np.float64).reshape(-1)
        Z = np.ascontiguousarray(Z, np.uint32)
        self._h = C.c_void_p()
        ctx._check(lib().msf_specimen_build(ctx._h, C.byref(grid), _p(q), _p(Z, C.c_uint32),
                                            Z.size, min_slice, C.byref(self._h)),
                   "msf_specimen_build")
        self.n_slices = lib().msf_specimen_slices(self._h)

    def free(self):
        if self._h:
            lib().msf_specimen_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def slice_info(self):
        n = self.n_slices
        lb, th, na = np.zeros(n), np.zeros(n), np.zeros(n, np.uint32)
        rc = lib().msf_specimen_slice_info(self._h, _p(lb), _p(th), _p(na, C.c_uint32))
        if rc:
            raise MsfError(rc, "msf_specimen_slice_info")
        return lb, th, na

    def propagator(self, s):
        px = np.zeros(self.grid.nx, np.complex128)
        py = np.zeros(self.grid.ny, np.complex128)
        rc = lib().msf_specimen_propagator(self._h, s, _p(px.view(np.float64)),
                                           _p(py.view(np.float64)))
        if rc:
            raise MsfError(rc, "msf_specimen_propagator")
        return px, py

    def transmission(self, s):
        out = np.zeros((self.grid.nx, self.grid.ny), np.complex128)
        rc = lib().msf_specimen_transmission(self._h, s, _p(out.view(np.float64)))
        if rc:
            raise MsfError(rc, "msf_specimen_transmission")
        return out
