"""Seeded amorphous specimens and rasters for the benchmark workloads.

Positions come from random sequential addition with a minimum-image distance:
candidates are uniform in the cell, drawn from xoshiro256** seeded through
splitmix64 (u = (x >> 11) * 2^-53), and accepted when no accepted atom lies
closer than `min_dist` under periodic boundaries.  For each accepted atom of a
two-species specimen one more draw picks the species (u < p -> first).  Output
follows the xyz reader's conventions (tem_io.cpp:4627-4956): positions in
[0, period), domain minima 0.
"""
import numpy as np

_M64 = (1 << 64) - 1


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & _M64


class Xoshiro256ss:
    def __init__(self, seed):
        s = seed & _M64
        st = []
        for _ in range(4):  # splitmix64
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            st.append(z ^ (z >> 31))
        self.s = st

    def next(self):
        s = self.s
        r = (_rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return r

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def amorphous(n_atoms, cell, min_dist, seed, species=(14,), p_first=0.5, max_tries=200):
    """Random-sequential-addition specimen.  Returns (q [n, 3] float64,
    Z [n] uint32).  Raises if the density cannot be reached."""
    a, b, c = (float(v) for v in cell)
    rng = Xoshiro256ss(seed)
    ncell = [max(1, int(v // min_dist)) for v in (a, b, c)]
    cw = [v / n for v, n in zip((a, b, c), ncell)]
    grid = {}
    q = np.zeros((n_atoms, 3))
    Z = np.zeros(n_atoms, np.uint32)
    d2 = min_dist * min_dist
    period = np.array([a, b, c])
    count = 0
    tries = 0
    limit = max_tries * max(n_atoms, 1)
    while count < n_atoms:
        tries += 1
        if tries > limit:
            raise RuntimeError(f"could not place {n_atoms} atoms at min distance {min_dist}")
        x = np.array([rng.uniform() * a, rng.uniform() * b, rng.uniform() * c])
        ci = [min(int(x[k] // cw[k]), ncell[k] - 1) for k in range(3)]
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((ci[0] + dx) % ncell[0], (ci[1] + dy) % ncell[1], (ci[2] + dz) % ncell[2])
                    for j in grid.get(key, ()):
                        d = x - q[j]
                        d -= period * np.round(d / period)
                        if d @ d < d2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        q[count] = x
        if len(species) == 1:
            Z[count] = species[0]
        else:
            Z[count] = species[0] if rng.uniform() < p_first else species[1]
        grid.setdefault(tuple(ci), []).append(count)
        count += 1
    return q, Z


_ELEMENTS = {1: "H", 6: "C", 8: "O", 13: "Al", 14: "Si", 29: "Cu", 40: "Zr", 79: "Au"}


def write_xyz(path, q, Z, cell):
    """xyz file as read by read_position_xyz_file (tem_io.cpp:4627-4956):
    population, periods, then `element x y z` lines."""
    with open(path, "w") as f:
        f.write(f"{len(Z)}\n{cell[0]:.10f} {cell[1]:.10f} {cell[2]:.10f}\n")
        for z, (x, y, w) in zip(Z, q):
            f.write(f"{_ELEMENTS[int(z)]} {x:.10f} {y:.10f} {w:.10f}\n")


def raster(n, period, qmin=0.0):
    """assign_stem_raster_points (func.cpp:108-127): qmin + i * period / n."""
    d = period / n
    return np.array([qmin + (i * d) for i in range(n)])


# ---------------------------------------------------------------- workloads
def config_si_256(seed=1001, n_pos=16):
    """400 Si in a 20 A cube, 256^2, 300 kV, aperture 10 mrad, minslice 2.0,
    16 x 16 positions over the central 10 x 10 A, ADF 40-200 mrad."""
    q, Z = amorphous(400, (20.0, 20.0, 20.0), 2.0, seed)
    job = dict(nx=256, ny=256, xperiod=20.0, yperiod=20.0, zperiod=20.0, voltage=300e3,
               min_slice=2.0, cs3=0.0, defocus=0.0, alpha_max=0.010, det_inner=0.040,
               det_outer=0.200)
    xp = 5.0 + 0.625 * np.arange(n_pos)
    return job, q, Z, xp, xp.copy()


def config_cuzr_512(seed=2001):
    """3500 Cu/Zr (coin flip) in a 40 A cube, min distance 2.2 A, 512^2,
    200 kV, Cs3 = 1.2 mm, Scherzer defocus, aperture 2 mrad, minslice 2.0,
    32 x 32 raster at 1.25 A, FEM V_omega with dr = 0.1414213562."""
    q, Z = amorphous(3500, (40.0, 40.0, 40.0), 2.2, seed, species=(29, 40))
    job = dict(nx=512, ny=512, xperiod=40.0, yperiod=40.0, zperiod=40.0, voltage=200e3,
               min_slice=2.0, cs3=1.2e7, defocus=None, alpha_max=0.002, det_inner=0.040,
               det_outer=0.200, dr=0.1414213562)
    xp = raster(32, 40.0)
    return job, q, Z, xp, xp.copy()


def uneven_jobs(n_jobs=48, seed=3000, grids=None):
    """Batched uneven-size STEM jobs: job j draws from numpy's default_rng(seed
    + j) N in {256, 384, 512}, a cubic cell side U[20, 50] A and a raster side
    U{16..48} (positions at spacing L/side over the cell); amorphous Si at
    0.05 A^-3 with a 2.0 A minimum separation (seed + j); minslice 2.0;
    ADF 50-180 mrad.  100 kV and a 20 mrad aperture (the workload names
    neither): at 300 kV the 50 mrad inner angle lies beyond the 2/3 band
    limit lambda * N / (3 L) of the coarser grids (N 256 with L > 33.6 A),
    which the reference rejects (tem.cpp:721-732).  With `grids`, only the
    jobs whose N is listed (the others are drawn but not built)."""
    jobs = []
    for j in range(n_jobs):
        rng = np.random.default_rng(seed + j)
        N = int(rng.choice([256, 384, 512]))
        L = float(rng.uniform(20.0, 50.0))
        side = int(rng.integers(16, 49))
        if grids is not None and N not in grids:
            continue
        q, Z = amorphous(int(round(0.05 * L ** 3)), (L, L, L), 2.0, seed + j)
        jobs.append(dict(nx=N, ny=N, cell=(L, L, L), q=q, Z=Z, x_p=raster(side, L),
                         y_p=raster(side, L), voltage=100e3, alpha_max=0.020, det_inner=0.050,
                         det_outer=0.180, min_slice=2.0))
    return jobs


def serving_mix(n_jobs, seed=5000, rate=4.0):
    """Seeded job stream in the manner of a serving workload: Poisson
    arrivals at `rate` jobs/s; N in {256, 512, 1024} with weights
    0.5/0.35/0.15 at 0.078125 A per pixel; slices U{10..50} from a cell depth
    of 2 A per slice (minslice 2.0); raster side U{16..64} over the cell;
    amorphous Si (0.05 A^-3) or Cu/Zr (p = 0.5); 300 kV, 20 mrad, ADF
    60-200 mrad and FEM V_omega.  Atom positions are uniform without the
    minimum separation: the device work depends only on grid, slices and
    raster.  Returns a list of dicts (arrival time t in s)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(n_jobs):
        t += float(rng.exponential(1.0 / rate))
        N = int(rng.choice([256, 512, 1024], p=[0.5, 0.35, 0.15]))
        L = N * 0.078125
        S = int(rng.integers(10, 51))
        depth = 2.0 * S
        side = int(rng.integers(16, 65))
        cuzr = bool(rng.random() < 0.5)
        n_at = int(round((0.05 if not cuzr else 3500 / 40.0 ** 3) * L * L * depth))
        q = rng.random((n_at, 3)) * [L, L, depth]
        Z = (np.where(rng.random(n_at) < 0.5, 29, 40) if cuzr else np.full(n_at, 14)).astype(np.uint32)
        jobs.append(dict(t=t, nx=N, ny=N, cell=(L, L, depth), q=q, Z=Z, slices=S,
                         x_p=raster(side, L), y_p=raster(side, L), voltage=300e3,
                         alpha_max=0.020, det_inner=0.060, det_outer=0.200, dr=0.1414213562,
                         min_slice=2.0))
    return jobs
