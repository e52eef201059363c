"""Benchmark: probe-slice propagations/s of the multislice FEM job at 512^2.

Workload (one "step" = one whole job on resident inputs): 3,500 Cu/Zr atoms
(seeded coin flip, 2.2 A minimum separation, seed 2001) in a 40 A cube, 512^2
grid, 200 kV, Cs3 = 1.2 mm at Scherzer defocus, 2 mrad aperture, minslice 2.0
(auto-slicer), 32 x 32 raster at 1.25 A, FEM V_omega (dr = 0.1414213562).
The specimen's transmission functions are built before the timed region and
stay resident in HBM; each step runs probe set-up, the slice loop over every
position and the FP64 FEM statistics.

value  = positions x slices x steps / wall time of the timed steps (batched).
plain  = the same kernels one position at a time (batch 1) on a sample.
roofline: the dominant slice-pass kernel, compulsory HBM bytes per launch
          (read + write of the band-limited wave) over its average duration,
          measured with HIP events on the library's stream in a profiled step.
cpu_baseline: the FP64 CPU restatement (oracle, OpenMP over positions) on a
          bounded sample of positions of the same job; cpu_baseline.parity
          compares the GPU path with it on a sample of the job (relative L2 of
          V_omega and of exit-wave intensities).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ms-stem-fem_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "probe-slice propagations/s at 512² (batched FP32 vs plain); p99 job latency"


def algorithmic_bytes(nx, ny, batch):
    """Bytes one slice-pass launch moves by the unit-of-work accounting: each
    pass reads and writes the FP32 complex wave once, 16 nx ny bytes per
    position (32 nx ny per probe-slice propagation, two passes)."""
    return 16.0 * nx * ny * batch


def _tune(args, default=0):
    """msf_config.autotune of a workload: --autotune if given, else the
    workload's default (0: the library's automatic tuning)."""
    return default if args.autotune is None else args.autotune


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    m = np.isfinite(b)
    if not np.array_equal(np.isfinite(a), m):
        return float("inf")
    return float(np.linalg.norm((a - b)[m]) / max(np.linalg.norm(b[m]), 1e-300))


def cpu_baseline(job, q, Z, xp, yp, threads, target_s, gpu=None):
    """Oracle (FP64, own FFT) on a bounded sample of the raster.  With
    gpu = (ctx, spec, optics, fem) it also checks the GPU path against the
    oracle on that sample: V_omega of the positions (xp[0], yp[:]) and the
    exit-wave intensities of four of them, relative L2."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    oj = O.make_job(nx=job["nx"], ny=job["ny"], xperiod=job["xperiod"], yperiod=job["yperiod"],
                    zperiod=job["zperiod"], voltage=job["voltage"], min_slice=job["min_slice"],
                    cs3=job["cs3"], defocus=job["defocus"], alpha_max=job["alpha_max"], adf=0,
                    fem=1, dr=job["dr"])
    t0 = time.perf_counter()
    spec = O.Specimen(oj, q, Z.astype(np.int32))
    setup_s = time.perf_counter() - t0
    parity = None
    if gpu is not None:
        ctx, gspec, optics, fem = gpu
        _, v_o, _ = O.adfstem(oj, spec, xp[:1], yp, n_threads=threads)
        _, v_g = ctx.adfstem(gspec, optics, xp[:1], yp, fem=fem)
        _, _, ex_o = O.adfstem(oj, spec, xp[:1], yp[:4], n_threads=threads, exit_waves=True)
        ex_g = ctx.exit_waves(gspec, optics, xp[:1], yp[:4])
        parity = {"sample": f"positions (x_p[0], y_p[0:{yp.size}]) of the timed job; exit waves of the first 4",
                  "v_omega_rel_l2": rel_l2(v_g, v_o),
                  "exit_intensity_rel_l2_max": max(rel_l2(np.abs(ex_g[k]) ** 2, np.abs(ex_o[k]) ** 2)
                                                   for k in range(4)),
                  "tolerance": {"v_omega": 1e-4, "exit_intensity": 1e-5}}
        log(f"parity vs oracle: V {parity['v_omega_rel_l2']:.2e}, |psi|^2 "
            f"{parity['exit_intensity_rel_l2_max']:.2e}")
    # one core: two positions on one thread (also sizes the sample)
    t0 = time.perf_counter()
    O.adfstem(oj, spec, xp[:1], yp[:2], n_threads=1)
    one = (time.perf_counter() - t0) / 2
    n_pos = int(max(threads, min(xp.size * yp.size, round(target_s * threads / max(one, 1e-3)))))
    n_pos = max(threads, (n_pos // threads) * threads)
    ys = np.resize(yp, n_pos)
    t0 = time.perf_counter()
    for k in range(0, n_pos, yp.size):
        O.adfstem(oj, spec, xp[:1], ys[k:k + yp.size], n_threads=threads)
    dt = time.perf_counter() - t0
    return {"value": n_pos * spec.n_slices / dt, "unit": "probe-slice propagations/s",
            "cores": threads, "kind": "port", "parity": parity,
            "one_core": round(spec.n_slices / one, 2), "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)), "cpu_model": _cpu_model(),
            "sample": f"{n_pos} positions x {spec.n_slices} slices of the same job "
                      f"(FP64 restatement, OpenMP over positions, {dt:.1f} s; one_core: 2 "
                      f"positions on one thread; specimen set-up {setup_s:.1f} s not included)",
            "note": "the oracle's FFT is an unoptimised recursive mixed-radix FFT that rebuilds "
                    "its twiddles on every 2-D call; an FFTW-class CPU path would be faster"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_calibration(nbytes):
    """A plain copy of known size (the runtime's blit kernel), so that memory
    counters of the slice passes can be scaled against it."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 x {nbytes} B device-to-device")


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """Device-to-device hipMemcpyAsync of nbytes timed with HIP events: the
    streaming rate a plain copy reaches on this GPU (bytes read + written
    per second), reported beside the roofline as the achievable reference."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    best = None
    for _ in range(reps):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        best = ms.value if best is None else min(best, ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def serving_sample(msf, SP, n_jobs, rate, slots, seed=5000):
    """A bounded serving run: n_jobs of a seeded mixed stream submitted at
    their Poisson arrival times to a context with `slots` serving slots;
    per-job latency = queue + run time (msf_wait_info)."""
    jobs = SP.serving_mix(n_jobs, seed=seed, rate=rate)
    ctx = msf.Context(0)
    try:
        ctx.serve_configure(slots=slots)
        tickets = []
        work = 0
        t0 = time.perf_counter()
        for j in jobs:
            delay = j["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            grid = msf.make_grid(j["nx"], j["ny"], j["cell"][0], j["cell"][1], j["cell"][2],
                                 j["voltage"])
            optics = msf.Optics(0, 0.0, 0.0, 0.0, j["alpha_max"])
            det = msf.Detector(1, j["det_inner"], j["det_outer"], msf.ADF_REFERENCE)
            tickets.append(ctx.submit(grid, j["q"], j["Z"], j["min_slice"], optics, j["x_p"],
                                      j["y_p"], detector=det, fem=msf.Fem(1, j["dr"])))
            work += j["x_p"].size * j["y_p"].size * j["slices"] * (j["nx"] * j["ny"]) / 512 ** 2
            if len(tickets) % 100 == 0:
                log(f"serving: {len(tickets)}/{n_jobs} jobs submitted at t = {time.perf_counter() - t0:.1f} s")
        lat = []
        for t in tickets:
            _, _, info = ctx.wait(t, info=True)
            lat.append(info["queue_ms"] + info["run_ms"])
        span = time.perf_counter() - t0
        st = ctx.serve_stats()
    finally:
        ctx.close()
    lat = np.array(lat)
    return {"jobs": n_jobs, "arrival_rate_per_s": rate, "slots": slots,
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1),
                           "p99": round(float(np.percentile(lat, 99)), 1)},
            "throughput_jobs_per_s": round(n_jobs / span, 2),
            "throughput_512eq_props_per_s": round(work / span, 1),
            "makespan_s": round(span, 2), "max_running": st["max_running"],
            "peak_admitted_gib": round(st["peak_admitted_bytes"] / 2 ** 30, 2),
            "mix": "N 256/512/1024 (0.5/0.35/0.15), 10-50 slices, rasters 16-64 per side, "
                   "Si or Cu/Zr, Poisson arrivals; 512eq = probe-slices x (N/512)^2"}


def _base_line(value, unit, steps, warmup, ms, config, roofline, extra):
    out = {"metric": METRIC, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": config, "roofline": roofline, "cpu_baseline": None}
    out.update(extra)
    return out


def single_job(args, msf, desc, seed, grid, q, Z, min_slice, optics, xp, yp, det, fem):
    """One job timed over `steps` runs after `warmup` runs, plus a profiled
    run for the dominant slice pass (compulsory bytes over its HIP-event
    duration)."""
    ctx = msf.Context(0)
    spec = ctx.specimen(grid, q, Z, min_slice)
    S, P = spec.n_slices, xp.size * yp.size
    ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args))
    for _ in range(args.warmup):
        ctx.adfstem(spec, optics, xp, yp, detector=det, fem=fem)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.adfstem(spec, optics, xp, yp, detector=det, fem=fem)
        times.append(time.perf_counter() - t0)
    st = ctx.stats()
    ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args), profile_kernels=True)
    ctx.adfstem(spec, optics, xp, yp, detector=det, fem=fem)
    pst = ctx.stats()
    kern = {"row_pass": (pst["row_pass_us"], pst["row_pass_bytes"]),
            "col_pass": (pst["col_pass_us"], pst["col_pass_bytes"])}
    dom = max(kern, key=lambda k: kern[k][0])
    us, nbytes = kern[dom]
    alg = algorithmic_bytes(grid.nx, grid.ny, pst["batch"])
    achieved = alg / (us * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_launch": alg, "compulsory_bytes_per_launch": nbytes,
                "achieved_compulsory": round(nbytes / (us * 1e-6) / 1e9, 1),
                "avg_us": round(us, 2),
                "other_kernels_us": {k: round(v[0], 2) for k, v in kern.items() if k != dom}}
    value = P * S * len(times) / sum(times)
    config = {"workload": desc, "grid": int(grid.nx), "atoms": int(Z.size), "slices": int(S),
              "positions": int(P), "batch": int(st["batch"]), "chains": int(st["chains"]),
              "col_tile": int(st["col_tile"]), "graph": int(st["graph"]),
              "fft_plan": [int(st["row_fft_plan"]), int(st["col_fft_plan"])], "seed": seed}
    spec.free()
    ctx.close()
    return _base_line(value, "probe-slice propagations/s", args.steps, args.warmup,
                      1e3 * sum(times) / len(times), config, roofline,
                      {"slice_loop_ms": round(st["slice_loop_ms"], 3)})


def uneven_batch(args, msf, SP):
    """48 uneven jobs (N 256/384/512): first run one after another on one
    context with the autotuner and graph replay (an untimed pass tunes each
    shape, a second is timed), then submitted together to the serving engine
    (3 slots), once with every job alone and once with grouping (--group:
    queued jobs of one grid size run together, their slice passes sharing
    launches); each job builds its own specimen.  `value` is the submission
    with the engine's defaults (jobs alone); `grouped` the second."""
    grids = [args.grid] if args.grid else None
    log("generating 48 uneven jobs (seeds 3000-3047)" + (f", grid {args.grid} only" if grids else ""))
    jobs = SP.uneven_jobs(48, seed=3000, grids=grids)
    prep = []
    for j in jobs:
        grid = msf.make_grid(j["nx"], j["ny"], j["cell"][0], j["cell"][1], j["cell"][2], j["voltage"])
        prep.append((grid, msf.Optics(0, 0.0, 0.0, 0.0, j["alpha_max"]),
                     msf.Detector(1, j["det_inner"], j["det_outer"], msf.ADF_REFERENCE)))
    ctx = msf.Context(0)
    ctx.configure(autotune=_tune(args, 1), fft_plan=args.fft_plan)
    slices = []
    lat = []
    t_all = 0.0
    for rep in range(2):
        for j, (grid, optics, det) in zip(jobs, prep):
            t0 = time.perf_counter()
            spec = ctx.specimen(grid, j["q"], j["Z"], j["min_slice"])
            ctx.adfstem(spec, optics, j["x_p"], j["y_p"], detector=det)
            n_sl = spec.n_slices
            spec.free()
            dt = time.perf_counter() - t0
            if rep == 0:
                slices.append(n_sl)
            else:
                lat.append(dt)
                t_all += dt
        log(f"direct pass {rep}: done")
    ctx.close()
    work = sum(j["x_p"].size * j["y_p"].size * s for j, s in zip(jobs, slices))
    work512 = sum(j["x_p"].size * j["y_p"].size * s * (j["nx"] / 512) ** 2
                  for j, s in zip(jobs, slices))
    direct = {"props_per_s": round(work / t_all, 1), "props_512eq_per_s": round(work512 / t_all, 1),
              "makespan_s": round(t_all, 3),
              "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                             "p99": round(1e3 * float(np.percentile(lat, 99)), 2)}}
    log(f"direct: {direct['props_per_s']:.4g} props/s")
    # one batched submission through the serving engine, jobs alone and then
    # grouped (queued jobs of one grid size share slice-pass launches)
    def served(group):
        ctx = msf.Context(0)
        try:
            ctx.serve_configure(slots=3)
            ctx.serve_group(group)
            t0 = time.perf_counter()
            tickets = [ctx.submit(g, j["q"], j["Z"], j["min_slice"], o, j["x_p"], j["y_p"],
                                  detector=d) for j, (g, o, d) in zip(jobs, prep)]
            slat = []
            for t in tickets:
                _, _, info = ctx.wait(t, info=True)
                slat.append(info["queue_ms"] + info["run_ms"])
            return time.perf_counter() - t0, slat, ctx.serve_stats()
        finally:
            ctx.close()

    span1, slat1, sst1 = served(0)
    log(f"served alone: {work / span1:.4g} props/s")
    span, slat, sst = served(args.group)
    log(f"served grouped ({args.group}): {work / span:.4g} props/s, {sst['groups']} groups, "
        f"{sst['grouped_jobs']} jobs, {sst['group_launches']} launches")
    Ns = [j["nx"] for j in jobs]
    config = {"workload": "uneven_stem_48: 48 Si jobs, N 256/384/512, cubic 20-50 A at 0.05 A^-3, "
                          "minslice 2.0, rasters 16-48 per side, ADF 50-180 mrad, 100 kV, 20 mrad",
              "jobs": len(jobs), "grids": {str(n): Ns.count(n) for n in (256, 384, 512)},
              "atoms": int(sum(j["Z"].size for j in jobs)), "slices": int(sum(slices)),
              "positions": int(sum(j["x_p"].size * j["y_p"].size for j in jobs)),
              "probe_slices": int(work), "slots": 3, "seed": 3000}
    pct = lambda v, q: round(float(np.percentile(v, q)), 1)
    return _base_line(work / span1, "probe-slice propagations/s", 1, 0, 1e3 * span1, config, None,
                      {"props_512eq_per_s": round(work512 / span1, 1),
                       "latency_ms": {"p50": pct(slat1, 50), "p99": pct(slat1, 99)},
                       "max_running": sst1["max_running"],
                       "peak_admitted_gib": round(sst1["peak_admitted_bytes"] / 2 ** 30, 2),
                       "grouped": {"max_jobs": args.group, "props_per_s": round(work / span, 1),
                                   "makespan_s": round(span, 3),
                                   "latency_ms": {"p50": pct(slat, 50), "p99": pct(slat, 99)},
                                   "groups": sst["groups"], "grouped_jobs": sst["grouped_jobs"],
                                   "group_launches": sst["group_launches"],
                                   "max_running": sst["max_running"]},
                       "direct_autotuned_graph": direct})


def cuzr_realizations(args, msf, SP):
    """The 512^2 Cu/Zr FEM job on 4 seeded specimen realizations (seeds
    2001-2004), each building its own specimen; V_omega over the bins of
    0-1.0 1/A averaged over the realizations, outside the per-run
    computation (empty rings are NaN in every realization alike)."""
    seeds = (2001, 2002, 2003, 2004)
    ctx = msf.Context(0)
    ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args))
    fem = None
    runs = []
    for seed in seeds:
        log(f"generating specimen (seed {seed})")
        job, q, Z, xp, yp = SP.config_cuzr_512(seed=seed)
        lam, _ = msf.wavelength(job["voltage"])
        grid = msf.make_grid(job["nx"], job["ny"], job["xperiod"], job["yperiod"], job["zperiod"],
                             job["voltage"])
        optics = msf.Optics(0, job["cs3"], 0.0, msf.scherzer_defocus(lam, job["cs3"]),
                            job["alpha_max"])
        fem = msf.Fem(1, job["dr"])
        runs.append((grid, q, Z, job["min_slice"], optics, xp, yp))
    for _ in range(args.warmup):   # the first realization, untimed
        g, q, Z, ms, o, xp, yp = runs[0]
        spec = ctx.specimen(g, q, Z, ms)
        ctx.adfstem(spec, o, xp, yp, fem=fem)
        spec.free()
    vs, work, t_all = [], 0, 0.0
    for g, q, Z, ms, o, xp, yp in runs:
        t0 = time.perf_counter()
        spec = ctx.specimen(g, q, Z, ms)
        _, v = ctx.adfstem(spec, o, xp, yp, fem=fem)
        S = spec.n_slices
        spec.free()
        t_all += time.perf_counter() - t0
        work += xp.size * yp.size * S
        vs.append(v)
    edges = msf.fem_bins(runs[0][0], fem.dr)
    dk = float(edges[1] - edges[0])
    nb = int(np.sum(edges + dk <= 1.0 + 1e-9))          # bins within 0-1.0 1/A
    vavg = np.mean(np.stack(vs)[:, :nb], axis=0)
    if args.v_out:
        msf.write_variance_txt(args.v_out, vavg, edges[:nb])
    st = ctx.stats()
    ctx.close()
    config = {"workload": "cuzr_fem_512 x 4 realizations: 3500 Cu/Zr atoms, 40 A cube, 512x512, "
                          "200 kV, Cs3 1.2 mm Scherzer, 2 mrad, 32x32 raster, V_omega averaged over "
                          "seeds 2001-2004", "grid": 512, "realizations": len(seeds),
              "positions": int(runs[0][5].size * runs[0][6].size), "probe_slices": int(work),
              "batch": int(st["batch"]), "chains": int(st["chains"]), "seed": list(seeds)}
    return _base_line(work / t_all, "probe-slice propagations/s", len(seeds), args.warmup,
                      1e3 * t_all / len(seeds), config, None,
                      {"v_omega_mean": {"delta_k": dk, "n_bins": nb,
                                        "values": [None if not np.isfinite(x) else float(f"{x:.6e}")
                                                   for x in vavg]}})


def other_workload(args):
    import msf
    import specimens as SP
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    if args.workload == "uneven48":
        out = uneven_batch(args, msf, SP)
    elif args.workload == "cuzr512x4":
        out = cuzr_realizations(args, msf, SP)
    elif args.workload == "si256":
        job, q, Z, xp, yp = SP.config_si_256(seed=1001)
        grid = msf.make_grid(256, 256, 20.0, 20.0, 20.0, job["voltage"])
        out = single_job(args, msf, "si_stem_256: 400 Si atoms, 20 A cube, 256x256, 300 kV, Cs 0, "
                         "10 mrad, 16x16 positions over the central 10x10 A, ADF 40-200 mrad",
                         1001, grid, q, Z, job["min_slice"],
                         msf.Optics(0, 0.0, 0.0, 0.0, job["alpha_max"]), xp, yp,
                         msf.Detector(1, job["det_inner"], job["det_outer"], msf.ADF_REFERENCE), None)
    else:
        log("generating 32000 Si atoms (seed 4001)")
        q, Z = SP.amorphous(32000, (80.0, 80.0, 100.0), 2.0, 4001)
        grid = msf.make_grid(1024, 1024, 80.0, 80.0, 100.0, 300e3)
        xp = 8.0 + 0.5 * np.arange(128)
        out = single_job(args, msf, "si_fem_1024: 32000 Si atoms, 80x80x100 A, 1024x1024, 300 kV, "
                         "Cs 0, 20 mrad, 128x128 positions at 0.5 A, ADF 60-200 mrad, V_omega",
                         4001, grid, q, Z, 2.0, msf.Optics(0, 0.0, 0.0, 0.0, 0.020), xp, xp.copy(),
                         msf.Detector(1, 0.060, 0.200, msf.ADF_REFERENCE), msf.Fem(1, 0.1414213562))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cuzr512",
                    choices=["cuzr512", "cuzr512x4", "si256", "uneven48", "si1024"],
                    help="cuzr512: the headline 512^2 FEM job (default); cuzr512x4 (its 4 seeded "
                         "realizations with the averaged V_omega), si256, uneven48, si1024: the other "
                         "benchmark configurations, one JSON line each")
    ap.add_argument("--v-out", default="",
                    help="cuzr512x4: write the averaged V_omega as the reference's text file")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--autotune", type=int, default=None, choices=[-1, 0, 1],
                    help="msf_config.autotune (-1: the fixed rules, as for counter profiles)")
    ap.add_argument("--group", type=int, default=16,
                    help="uneven48: queued jobs of one grid size a served group may hold")
    ap.add_argument("--grid", type=int, default=0, choices=[0, 256, 384, 512],
                    help="uneven48: only the jobs of this grid (0: all 48)")
    ap.add_argument("--fft-plan", type=int, default=0,
                    help="radix plans: 1 primary, 2 alternate, 0 automatic per length")
    ap.add_argument("--plain-positions", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=100,
                    help="jobs of the bounded serving sample (0: skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run hipMemcpy device-to-device of this size first (counter scale)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the multislice path has no multi-GPU mode in this "
              "build (positions would shard across GPUs with no exchange); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload != "cuzr512":
        return other_workload(args)

    import msf
    import specimens as SP

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)

    log("generating specimen (seed 2001)")
    job, q, Z, xp, yp = SP.config_cuzr_512(seed=2001)
    lam, _ = msf.wavelength(job["voltage"])
    job["defocus"] = msf.scherzer_defocus(lam, job["cs3"])
    ctx = msf.Context(0)
    grid = msf.make_grid(job["nx"], job["ny"], job["xperiod"], job["yperiod"], job["zperiod"],
                         job["voltage"])
    t0 = time.perf_counter()
    spec = ctx.specimen(grid, q, Z, job["min_slice"])
    spec_ms = (time.perf_counter() - t0) * 1e3
    S = spec.n_slices
    P = xp.size * yp.size
    optics = msf.Optics(0, job["cs3"], 0.0, job["defocus"], job["alpha_max"])
    fem = msf.Fem(1, job["dr"])
    log(f"specimen: {S} slices, build {spec_ms:.1f} ms")

    ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args))
    for _ in range(args.warmup):
        ctx.adfstem(spec, optics, xp, yp, fem=fem)
    times = []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, v = ctx.adfstem(spec, optics, xp, yp, fem=fem)
        times.append(time.perf_counter() - t0)
    elapsed = time.perf_counter() - t_all
    st = ctx.stats()
    value = P * S * args.steps / elapsed
    log(f"batched: {value:.4g} props/s, batch {st['batch']}, loop {st['slice_loop_ms']:.2f} ms")

    # profiled step: per-kernel durations with HIP events on the library stream
    ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args), profile_kernels=True)
    ctx.adfstem(spec, optics, xp, yp, fem=fem)
    pst = ctx.stats()
    ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args))
    kern = {"row_pass": (pst["row_pass_us"], pst["row_pass_bytes"]),
            "col_pass": (pst["col_pass_us"], pst["col_pass_bytes"])}
    dom = max(kern, key=lambda k: kern[k][0])
    us, nbytes = kern[dom]
    alg = algorithmic_bytes(job["nx"], job["ny"], pst["batch"])
    achieved = alg / (us * 1e-6) / 1e9 if us > 0 else 0.0
    # HBM traffic of one full-batch launch from the committed counter
    # summary (FETCH_SIZE/WRITE_SIZE per position, copy-calibrated)
    traffic = None
    lds = None
    pmc = os.path.join(ROOT, "profiles", f"{dom}_cuzr512_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            prof = json.load(f)
        per_pos = prof.get("hbm_bytes_per_position")
        if per_pos:
            traffic = float(per_pos) * pst["batch"]   # the profiled launches' batch
        # LDS bytes of one full launch (SQ_INSTS_LDS_*_BANDWIDTH counters) over
        # this run's event-timed launch duration, against the LDS peak rates
        pl = prof.get("lds")
        if pl and us > 0:
            rd = pl["read_bytes_per_launch"] / (us * 1e-6) / 1e12
            wr = pl["write_bytes_per_launch"] / (us * 1e-6) / 1e12
            lds = {"read_tbs": round(rd, 2), "write_tbs": round(wr, 2),
                   "peak_read_tbs": pl["peak_read_tbs_published"],
                   "peak_write_tbs": pl["peak_write_tbs_published"],
                   "frac": round(rd / pl["peak_read_tbs_published"]
                                 + wr / pl["peak_write_tbs_published"], 4),
                   "busy_frac_counters": pl["busy_frac"]}
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "bytes_per_launch": alg, "compulsory_bytes_per_launch": nbytes,
                "achieved_compulsory": round(nbytes / (us * 1e-6) / 1e9, 1) if us > 0 else 0.0,
                "avg_us": round(us, 2),
                "other_kernels_us": {k: round(v[0], 2) for k, v in kern.items() if k != dom}
                | {"first_row_pass": round(pst["first_row_pass_us"], 2),
                   "last_col_pass": round(pst["last_col_pass_us"], 2),
                   "reduce": round(pst["reduce_us"], 2)}}

    roofline["lds"] = lds
    # achievable streaming rate: the library's own copy kernel over 1 GiB
    # (read + write), and the runtime's blit copy for comparison
    copy_gbs = round(ctx.copy_rate(1 << 30, 5), 1)
    roofline["copy_gbs"] = copy_gbs
    roofline["copy_kernel"] = "msf_copy_rate: 16-byte vector loads and stores, 1 GiB, best of 5"
    roofline["blit_copy_gbs"] = round(copy_bandwidth(), 1)
    # the band-limited wave's bytes actually moved, against the copy rate
    roofline["frac_of_copy"] = round(roofline["achieved_compulsory"] / copy_gbs, 4)
    log(f"copy kernel: {copy_gbs:.0f} GB/s (blit {roofline['blit_copy_gbs']:.0f}); "
        f"{dom}: {achieved:.0f} GB/s algorithmic, {roofline['achieved_compulsory']:.0f} GB/s "
        f"of compulsory traffic")

    plain = None
    if not args.no_plain:
        n = min(args.plain_positions, P)
        ctx.configure(batch=1, graph=False, chains=1, fft_plan=args.fft_plan)   # one position per launch, stream launches
        xs = np.resize(xp, n)
        ctx.adfstem(spec, optics, xs[:1], yp[:1], fem=fem)
        t0 = time.perf_counter()
        ctx.adfstem(spec, optics, xs, yp[:1], fem=fem)
        dt = time.perf_counter() - t0
        pls = ctx.stats()
        ctx.configure(batch=args.batch, fft_plan=args.fft_plan, autotune=_tune(args))
        plain = {"value": n * S / dt, "positions": n, "batch": 1,
                 "shape": f"stream launches, one position per launch: row pass one row per "
                          f"workgroup, column pass {pls['col_tile']} column(s) per workgroup",
                 "speedup_batched_vs_plain": round(value / (n * S / dt), 2)}
        log(f"plain: {plain['value']:.4g} props/s")

    serving = None
    if args.serving_jobs > 0:
        log(f"serving sample: {args.serving_jobs} jobs, Poisson 4/s, 3 slots")
        serving = serving_sample(msf, SP, args.serving_jobs, 4.0, 3)
        log(f"serving: p50 {serving['latency_ms']['p50']} ms, p99 {serving['latency_ms']['p99']} ms, "
            f"{serving['throughput_jobs_per_s']} jobs/s, max {serving['max_running']} running")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(job, q, Z, xp, yp, threads, args.cpu_seconds,
                           gpu=(ctx, spec, optics, fem))

    lat = np.array(times) * 1e3
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "probe-slice propagations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(lat.mean()), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cuzr_fem_512: 3500 Cu/Zr atoms, 40 A cube, 512x512, 200 kV, "
                               "Cs3 1.2 mm Scherzer, 2 mrad, 32x32 raster, V_omega",
                   "grid": 512, "atoms": int(Z.size), "slices": int(S), "positions": int(P),
                   "batch": int(st["batch"]), "chains": int(st["chains"]),
                   "col_tile": int(st["col_tile"]), "graph": int(st["graph"]),
                   "autotuned": int(st["autotuned"]),
                   "fft_plan": [int(st["row_fft_plan"]), int(st["col_fft_plan"])], "seed": 2001},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain,
        "serving": serving,
        "job_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3)},
        "slice_loop_ms": round(st["slice_loop_ms"], 3),
        "specimen_build_ms": round(spec_ms, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
