"""Shared set-up for the parity tests: one job description run through the
FP64 oracle (checker) and through libmsf.so (product)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    m = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), m), "NaN/inf pattern differs"
    d = np.linalg.norm((a - b)[m])
    n = np.linalg.norm(b[m])
    return d / n if n > 0 else d


def random_atoms(n, cell, seed, species=(14,)):
    """Uniform atoms without a distance constraint (small parity cases)."""
    rng = np.random.default_rng(seed)
    q = rng.random((n, 3)) * np.asarray(cell, float)
    Z = rng.choice(np.asarray(species, np.uint32), size=n)
    return q, Z.astype(np.uint32)


class Case:
    def __init__(self, nx, ny, cell, voltage, q, Z, min_slice=1.0, cs3=0.0, cs5=0.0,
                 defocus=0.0, alpha_max=0.01, corrected=0, adf=True, inner=0.04, outer=0.2,
                 fem=False, dr=1 / np.sqrt(2), transmit_mode=0, adf_mode=0, mtf=None,
                 mtf_domain=None, mtf_rate=0.0):
        self.__dict__.update(locals())
        del self.__dict__["self"]

    def oracle_job(self, O):
        return O.make_job(nx=self.nx, ny=self.ny, xperiod=self.cell[0], yperiod=self.cell[1],
                          zperiod=self.cell[2], voltage=self.voltage, min_slice=self.min_slice,
                          corrected=self.corrected, cs3=self.cs3, cs5=self.cs5,
                          defocus=self.defocus, alpha_max=self.alpha_max,
                          adf=1 if self.adf else 0, det_inner=self.inner, det_outer=self.outer,
                          fem=1 if self.fem else 0, dr=self.dr,
                          transmit_mode=self.transmit_mode, adf_mode=self.adf_mode,
                          mtf=self.mtf, mtf_domain=self.mtf_domain, mtf_rate=self.mtf_rate)

    def grid(self, msf):
        return msf.make_grid(self.nx, self.ny, self.cell[0], self.cell[1], self.cell[2],
                             self.voltage)

    def optics(self, msf):
        return msf.Optics(self.corrected, self.cs3, self.cs5, self.defocus, self.alpha_max)

    def detector(self, msf):
        return msf.Detector(1 if self.adf else 0, self.inner, self.outer, self.adf_mode)

    def femcfg(self, msf):
        return msf.make_fem(self.dr, self.mtf, self.mtf_domain, self.mtf_rate,
                            enabled=1 if self.fem else 0)

    def run_oracle(self, O, x_p, y_p, exit_waves=False, threads=0):
        job = self.oracle_job(O)
        spec = O.Specimen(job, self.q, self.Z.astype(np.int32))
        return spec, O.adfstem(job, spec, x_p, y_p, n_threads=threads, exit_waves=exit_waves)

    def gpu_specimen(self, msf, ctx):
        return ctx.specimen(self.grid(msf), self.q, self.Z, self.min_slice)
