/*
 * msf.h -- C ABI of the MI355X multislice ADF-STEM / FEM path.
 *
 * Drop-in for the MPI+FFTW hot path of ms-stem-fem (reference files cited
 * relative to its repository root).  The reference has no FFI; its seam is
 * the C++ call TEM_NS::adfstem() (adfstem.hpp:66-121) made from tem.cpp
 * (:1526-1558 uncorrected, :1607-1639 Cs5-corrected), preceded by the
 * specimen set-up of tem.cpp:944-1276 (scatterer_pap_LUT, auto-slicing,
 * slice::update_propagator, slice::update_transmission_function).  The
 * entry points below replace those calls one for one; results that the
 * reference writes to files (STEM image TIFF, *_fem_V_omega.txt) are
 * returned in caller-owned host buffers instead.
 *
 * Conventions: every function returns MSF_OK (0) or a negative MSF_ERR_*
 * code; msf_last_error() gives the message of the last failure on a context.
 * All pointers are host pointers; the library owns every device buffer.
 * Complex arrays are interleaved (re, im) in the reference's row-major
 * [i*ny + j] layout, i = x row.  One context per host thread.
 */
#ifndef MSF_H
#define MSF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSF_OK 0
#define MSF_PENDING 1            /* msf_poll: the job is queued or running */
#define MSF_ERR_INVALID (-1)     /* bad argument (flags.fail paths of check_runtime_flags) */
#define MSF_ERR_UNSUPPORTED (-2) /* grid size without a compiled transform length */
#define MSF_ERR_DEVICE (-3)      /* HIP runtime error */
#define MSF_ERR_MEMORY (-4)      /* device allocation failed */
#define MSF_ERR_DOMAIN (-5)      /* position outside the domain the shift rule covers */

typedef struct msf_ctx msf_ctx;
typedef struct msf_specimen msf_specimen;

/* Simulation cell and sampling: tem.cpp:424-446 (lambda, gamma from the
 * voltage), :683-695 (band limits), :756-883 (domains).  xmin/ymin/zmin are
 * the domain minima; the xyz reader sets them to 0 (tem_io.cpp:4856-4858). */
typedef struct msf_grid {
    uint32_t nx, ny;                   /* samples_x, samples_y */
    double xperiod, yperiod, zperiod;  /* cell periods [A] */
    double xmin, ymin, zmin;           /* domain minima [A] */
    double voltage;                    /* microscope_voltage [V] */
} msf_grid;

/* Probe optics: probe_wavefunction.cpp:37-239, lenses.cpp:249-290. */
typedef struct msf_optics {
    int32_t corrected;                 /* 0: adfstemuncorr (Cs3); 1: adfstemcorr (Cs3 + Cs5) */
    double cs3, cs5;                   /* [A] */
    double defocus;                    /* [A] */
    double alpha_max;                  /* objective aperture [rad] */
} msf_optics;

/* Annular detector: sum_detected_intensity (func.cpp:2593-2641) called with
 * k1 = inner/lambda, k2 = outer/lambda (adfstem.cpp:2247-2253). */
#define MSF_ADF_REFERENCE 0 /* accept kx^2+ky^2 in (k1, k2], as the reference does */
#define MSF_ADF_PHYSICAL 1  /* accept kx^2+ky^2 in (k1^2, k2^2] */
typedef struct msf_detector {
    int32_t enabled;                   /* --images: compute the ADF image */
    double inner, outer;               /* detector angles [rad] */
    int32_t threshold_mode;            /* MSF_ADF_* */
} msf_detector;

/* FEM V_omega (d1): adfstem.cpp:947-1066, variance.cpp:64-300,
 * adfstem.cpp:2495-2601 and :2822-2866.
 * Detector MTF (--mtf_file, --mtf_sampling_rate; adfstem.cpp:238-453,
 * :2272-2385): when mtf and mtf_domain are set, mtf_n >= 1 and
 * mtf_sampling_rate != 0, every exit wave is replaced, after the ADF sum and
 * before the FEM statistics, by the MTF-filtered |psi| outside the aperture.
 * mtf_domain is in units of the Nyquist frequency, as read from the file
 * (msf_read_mtf); the arrays are read during the call only.  Zero-filled
 * trailing fields (an initialiser naming only enabled and dr) leave it off. */
typedef struct msf_fem {
    int32_t enabled;
    double dr;                         /* --dr azimuthal binning size factor (default 1/sqrt 2) */
    const double* mtf;                 /* mtf_n MTF values (second line of the MTF file) */
    const double* mtf_domain;          /* mtf_n spatial frequencies (first line) */
    uint32_t mtf_n;
    double mtf_sampling_rate;          /* --mtf_sampling_rate */
} msf_fem;

/* Transmission multiply of the slice loop (adfstem.cpp:1922-1952).  The
 * reference overwrites Re(psi) before computing Im(psi), so Im uses the
 * updated real part; MSF_TRANSMIT_REFERENCE reproduces that, _EXACT uses
 * the exact complex product. */
#define MSF_TRANSMIT_REFERENCE 0
#define MSF_TRANSMIT_EXACT 1

/* Execution settings (no reference counterpart: the reference runs one
 * position at a time on every MPI rank). */
typedef struct msf_config {
    uint32_t batch;                    /* probe positions per launch; 0 = automatic */
    int32_t transmit_mode;             /* MSF_TRANSMIT_* */
    int32_t profile_kernels;           /* 1: time every hot launch with HIP events */
    int32_t col_tile;                  /* columns per column-pass workgroup: 4, 8, 16; 0 = automatic */
    int32_t no_graph;                  /* 1: launch on the stream; 0: capture the job's launch
                                          sequence once per shape as a HIP graph and replay it */
    int32_t chains;                    /* batch chains in flight: 1, 2 (alternate batches on
                                          two streams), or 3 (batches in pairs on one stream,
                                          the row pass of one and the column pass of the other
                                          fused into one launch, on square 256/384/512 grids with
                                          their default radix plans and tiles; elsewhere the
                                          passes launch one after the other); FEM statistics
                                          still in raster order, results bit-identical;
                                          0 = automatic (two chains on the streams for
                                          launches of >= 4M pixels, else one chain as a graph) */
    int32_t autotune;                  /* with batch, col_tile and chains all automatic, the
                                          launch shape (positions per batch, column tile,
                                          chains 1, 2 or 3) comes from the process-wide tuning table
                                          (msf_tuning_*), measured on first use by timing whole
                                          batch steps of the job itself per device type and
                                          shape (nx, ny, kept columns, positions up to 256,
                                          outputs): 0 (default) for runs of at least 16384
                                          probe-slices, results bit-identical to the fixed
                                          rules; 1 for every run of two or more positions, and
                                          with fft_plan 0 the radix plans too (results then
                                          differ from the rules' by FP32 rounding); -1 the
                                          fixed rules */
    int32_t fft_plan;                  /* radix plans of the slice passes: 1 primary (every
                                          length), 2 the alternate plans with 16 values per lane
                                          and one fewer stage on lengths 256, 512 and 1024 (the
                                          primary elsewhere), 3 the alternate plan for the row
                                          passes (length ny) and the primary for the column
                                          passes, 4 the reverse; 0 = automatic (per length: the
                                          alternate at 256 and 1024, else the primary; or the
                                          autotuner's choice among these).  Plans differ by
                                          FP32 rounding only. */
} msf_config;

/* Timings of the last msf_run_stem() on the context. */
typedef struct msf_stats {
    double setup_ms;                   /* probe, shifts, detector tables */
    double slice_loop_ms;              /* all batches: slice passes + reductions */
    double total_ms;
    uint64_t probe_slices;             /* positions x slices propagated */
    uint32_t batch;                    /* batch depth used */
    uint32_t n_batches;
    /* per-kernel averages (profile_kernels = 1 only) */
    double row_pass_us, col_pass_us, first_row_pass_us, last_col_pass_us, reduce_us;
    uint64_t row_pass_count, col_pass_count;
    double row_pass_bytes, col_pass_bytes; /* compulsory HBM bytes of one launch */
    int32_t graph;                     /* 0: stream launches; 1: graph captured; 2: graph replayed */
    int32_t chains;                    /* batch chains in flight (msf_config.chains) */
    int32_t col_tile;                  /* columns per column-pass workgroup used */
    int32_t autotuned;                 /* 0: fixed rules; 1: tuned in this call; 2: tuned earlier */
    int32_t row_fft_plan;              /* radix plan of the row passes (length ny) and */
    int32_t col_fft_plan;              /* of the column passes (nx): 1 primary, 2 alternate */
    uint64_t fused_launches;           /* launches that ran a row pass and a column pass
                                          together (chains 3) */
    uint32_t group_runs;               /* msf_run_group: runs that took part */
    uint64_t group_launches;           /* msf_run_group: slice-pass launches, each over the
                                          current batches of up to 16 runs */
} msf_stats;

/* ------------------------------------------------------------ context */
int msf_create(msf_ctx** ctx, int device);
int msf_destroy(msf_ctx* ctx);
const char* msf_last_error(const msf_ctx* ctx);
int msf_set_config(msf_ctx* ctx, const msf_config* cfg);
int msf_get_stats(const msf_ctx* ctx, msf_stats* out);
/* 1 if transforms of length n are compiled in, else 0 */
int msf_supported_length(uint32_t n);

/* No reference counterpart (measurement aid): the rate of a streaming
 * device copy of `bytes` (rounded down to 16 KiB) with the library's own
 * copy kernel, read + write bytes per second in GB/s, the best of `reps`
 * event-timed launches after one warm-up.  The achievable HBM rate the
 * slice passes are compared with. */
int msf_copy_rate(msf_ctx* ctx, uint64_t bytes, uint32_t reps, double* gbs);

/* The tuning table (no direct reference counterpart; the reference keeps
 * FFTW wisdom in a file per grid, tem.cpp:588-616 and :1749-1765): the
 * autotuner's choices, process-wide, written to / read from a text file.
 * msf_tuning_load merges a file's entries (MSF_ERR_INVALID if it cannot be
 * read or a line is malformed; entries may be NULL); msf_tuning_clear forgets
 * every choice. */
int msf_tuning_save(const char* path);
int msf_tuning_load(const char* path, uint32_t* entries);
int msf_tuning_clear(void);

/* ------------------------------------------------------------ scalars */
/* tem.cpp:424-438 */
int msf_wavelength(double voltage, double* lambda, double* gamma);
/* scherzer.cpp:61-79 (uncorrected) and :141-157 (corrected to Cs5) */
double msf_scherzer_defocus(int32_t corrected, double lambda, double cs);
double msf_scherzer_alphamax(int32_t corrected, double lambda, double cs);
/* scherzer.cpp:131-139: Cs3 balancing a given Cs5 (adfstemcorr with --scherzer_cs3) */
double msf_scherzer_cs3(double lambda, double cs5);

/* ------------------------------------------------------------ specimen */
/* Replaces tem.cpp:944-1276: species templates (scatterer_pap_LUT,
 * scatterer_pap.hpp:57-231), assign_atoms_to_slices_z_auto
 * (slice.cpp:469-740), slice::update_propagator (slice.cpp:1293-1363) and
 * slice::update_transmission_function (slice.cpp:1451-1757) for every
 * slice.  q_xyz holds 3*n_atoms positions [A], Z the atomic numbers.  The
 * transmission functions stay resident on the device until msf_specimen_free. */
int msf_specimen_build(msf_ctx* ctx, const msf_grid* grid, const double* q_xyz,
                       const uint32_t* Z, uint32_t n_atoms, double min_slice,
                       msf_specimen** out);
int msf_specimen_free(msf_specimen* spec);
uint32_t msf_specimen_slices(const msf_specimen* spec);
/* lower_bound, thickness: n_slices doubles each; atoms: n_slices counts (any may be NULL) */
int msf_specimen_slice_info(const msf_specimen* spec, double* lower_bound, double* thickness,
                            uint32_t* atoms);
/* transmission function of one slice, FP64 as built (2*nx*ny doubles) */
int msf_specimen_transmission(const msf_specimen* spec, uint32_t slice, double* t);

/* propagator of one slice as built (slice.cpp:1293-1363): nx and ny complex
 * doubles (either may be NULL) */
int msf_specimen_propagator(const msf_specimen* spec, uint32_t slice, double* px, double* py);

/* ------------------------------------------------------------ probe */
/* Normalised real-space probe of adfstem.cpp:586-817 at (xx[0], yy[0]):
 * 2*nx*ny doubles. */
int msf_probe_realspace(msf_ctx* ctx, const msf_grid* grid, const msf_optics* optics,
                        double* probe);

/* Initial k-space wave of one position (adfstem.cpp:1590-1671,
 * :1831-1837): the normalised probe circularly shifted by (sx, sy) pixels,
 * FFT / sqrt(NxNy); 2*nx*ny doubles. */
int msf_probe_k(msf_ctx* ctx, const msf_grid* grid, const msf_optics* optics, int32_t sx,
                int32_t sy, double* psi_k);

/* One slice step in FP64 on `count` k-space waves in place
 * (adfstem.cpp:1866-2132 for slice `slice`: IFFT, transmission update in
 * the configured transmit_mode, FFT, band limit, propagator): a parity
 * entry point for single steps; the production loop is msf_run_stem. */
int msf_slice_step_batch(msf_ctx* ctx, const msf_specimen* spec, uint32_t slice, double* psi_k,
                         uint32_t count);

/* Parity entry point for the ADF ring sum alone (func.cpp:2593-2641, called at
 * adfstem.cpp:2245-2267): adf[p] = sum of |psi_p|^2 over the pixels with
 * k^2 in (k1, k2] (k1 = inner/lambda, k2 = outer/lambda, squared with
 * MSF_ADF_PHYSICAL), FP64 on the device, for count <= 65535 k-space waves
 * psi_k (count*nx*ny complex doubles, reference layout). */
int msf_adf_reduce(msf_ctx* ctx, const msf_grid* grid, const msf_detector* det,
                   const double* psi_k, uint32_t count, double* adf);

/* Parity entry point for the FEM ring sums and their accumulation
 * (variance.cpp:64-300 integrate_out_phi_fftw; adfstem.cpp:2518-2522
 * direct-beam zeroing, :2591-2599 accumulation, :2837-2856 V_omega), FP64 on
 * the device, on count <= 65535 k-space waves taken as the positions of one
 * raster in order.  radial: count*n_bins ring sums of |psi| (zero on
 * direct-beam bins); counts: n_bins ring pixel counts; v_omega: n_bins values
 * (msf_fem_bins).  Each output may be NULL. */
int msf_fem_accumulate(msf_ctx* ctx, const msf_grid* grid, const msf_fem* fem, double alpha_max,
                       const double* psi_k, uint32_t count, double* radial, int32_t* counts,
                       double* v_omega);

/* The 2-D detector MTF of fem's table on the grid (adfstem.cpp:300-439):
 * nx*ny complex doubles, reference layout.  Parity entry point. */
int msf_mtf_2d(msf_ctx* ctx, const msf_grid* grid, const msf_fem* fem, double* mtf2d);
/* The MTF application of adfstem.cpp:2272-2385 on count <= 65535 k-space
 * waves in place (FP64, reference layout): |psi| outside the aperture
 * (lambda^2 k^2 > alpha_max^2), BACKWARD FFT, times the 2-D MTF / sqrt(NxNy)
 * with the reference's in-place real/imaginary update, FORWARD FFT,
 * / sqrt(NxNy), bw_limit.  Parity entry point; the CLI uses it for the
 * final diffraction image of an MTF run. */
int msf_apply_mtf(msf_ctx* ctx, const msf_grid* grid, const msf_fem* fem, double alpha_max,
                  double* psi_k, uint32_t count);

/* ------------------------------------------------------------ run */
/* Number of V_omega bins (adfstem.cpp:947-972) and their lower edges
 * (n_bins doubles, may be NULL). */
int msf_fem_bins(const msf_grid* grid, double dr, uint32_t* n_bins, double* lower_edges);
/* Number of correlograph phi sectors (adfstem.cpp:969-985, :1013-1027): the
 * V_omega bin boundaries ii*delta_k (ii = 0..n_bins) lying in [0.2, 1.0] 1/A,
 * minus one; sector jj spans phi = PI + atan2(ky, kx) in
 * (jj, jj+1] * 2 PI / n_phi.  k_edges (n_bins + 1 doubles) and phi_edges
 * (n_phi + 1 doubles) receive the boundaries when not NULL. */
int msf_correlograph_bins(const msf_grid* grid, double dr, uint32_t* n_phi, double* k_edges,
                          double* phi_edges);

/* Replaces TEM_NS::adfstem() (adfstem.hpp:66-121, body adfstem.cpp:37-4096)
 * for the default modes: raster loop over x_p (outer) and y_p (inner),
 * probe shift, slice loop, ADF integration and FEM V_omega.
 * adf_image: n_xp*n_yp doubles, x-major (stem_image[py + px*n_yp],
 * adfstem.cpp:2264), written when det->enabled.
 * fem_v: n_bins doubles (msf_fem_bins), written when fem->enabled. */
int msf_run_stem(msf_ctx* ctx, const msf_specimen* spec, const msf_optics* optics,
                 const msf_detector* det, const msf_fem* fem,
                 const double* x_p, uint32_t n_xp, const double* y_p, uint32_t n_yp,
                 double* adf_image, double* fem_v);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, allocation-free serving loops):
 * msf_run_stem on a resident specimen with every device buffer of the run
 * (twiddle tables, pixel-group tables, vacuum waves, the batch wave buffer,
 * partial sums and accumulators) carved out of one static device array of
 * MSF_NOALLOC_SCRATCH_BYTES that each device the library runs on holds until
 * the process ends.  It calls no hipMalloc, hipFree or other allocating
 * routine, and takes no state from an earlier call: the array's contents
 * are rebuilt at every call (set-up is repeated: probe, vacuum waves, table
 * uploads), and the batch depth is the ordinary rule's or less, one chain,
 * stream launches.  Results equal msf_run_stem's bit for bit.  A run that
 * does not fit fails with MSF_ERR_INVALID before it launches its slice loop.
 * All calls on one device share its array: the library serialises them, so
 * calls that overlap on one device (from several contexts or threads) wait
 * for each other instead of racing. */
#define MSF_NOALLOC_SCRATCH_BYTES (256u << 20)
int msf_run_stem_noalloc(msf_ctx* ctx, const msf_specimen* spec, const msf_optics* optics,
                         const msf_detector* det, const msf_fem* fem,
                         const double* x_p, uint32_t n_xp, const double* y_p, uint32_t n_yp,
                         double* adf_image, double* fem_v);

/* Probe shift indices of every raster position, (sx, sy) pairs in raster
 * order (x outer, y inner): the nearest-pixel rule of adfstem.cpp:1590-1606
 * with indices persisting from the previous position when no grid point
 * qualifies (:1476-1477).  Host-only; MSF_ERR_DOMAIN below the domain. */
int msf_probe_shifts(const msf_grid* grid, const double* x_p, uint32_t n_xp, const double* y_p,
                     uint32_t n_yp, int32_t* shifts);

/* FEM outputs of one run, each n_bins doubles (msf_fem_bins) or NULL when
 * not wanted.  Modes of adfstem.cpp:2489-3605:
 *   v_omega       --V_omega (d1): N sum r^2 / (sum r)^2 - 1 of the ring means r of |psi|
 *   vbar_r        --Vbar_r (d2): mean over positions of <|psi|^2>_ring / <|psi|>_ring^2 - 1
 *   v_re          --V_re (d3): ring mean of <|psi|^2> over (ring mean of <|psi|>)^2, minus 1
 *   omega_vimage  --Omega_vimage (d4): ring mean of the variance image <|psi|^2>/<|psi|>^2 - 1
 *   v_gt17        --V_gt17: ring mean of <|psi|^2> over ring mean of <|psi|>^2
 *   v_rbar        --V_rbar: ring mean of <|psi|>^2 over (ring mean of <|psi|>)^2, minus 1
 * (<.> = mean over positions).  Empty and direct-beam bins follow the
 * reference (NaN or 0 as its arithmetic gives).
 *   correlograph_avg, correlograph_sqr_sum   --correlograph (adfstem.cpp:2603-2650,
 *       :3620-3632): n_bins * n_phi doubles (msf_correlograph_bins), index
 *       ii + jj * n_bins for ring ii and phi sector jj; per position
 *       c = count_ii * S_ii,jj / S_ii - 1 (S: sums of |psi| over the sector
 *       and the ring) where that ratio is >= 1, else 0; the mean of c and
 *       the sum of c^2 over positions.
 *   correlograph_each   (--correlograph_everyimage / _everytxt, adfstem.cpp:2655-2685):
 *       n_xp * n_yp * n_bins * n_phi doubles, c of every position in raster
 *       order, [p][ii + jj * n_bins]; computed with the correlograph.
 *   diffraction_mean_abs, diffraction_mean_sqr, variance_2d   (adfstem.cpp:2968-2990,
 *       :3060-3095, :3468-3520): nx * ny doubles each, reference layout: the
 *       mean over positions of |psi| and of |psi|^2 per pixel, and the 2-D
 *       variance image <|psi|^2> / <|psi|>^2 - 1 with the direct beam
 *       (lambda^2 k^2 <= alpha_max^2) set to 0 (0/0 elsewhere outside the band
 *       limit, as in the reference). */
typedef struct msf_fem_outputs {
    double* v_omega;
    double* vbar_r;
    double* v_re;
    double* omega_vimage;
    double* v_gt17;
    double* v_rbar;
    double* correlograph_avg;
    double* correlograph_sqr_sum;
    double* correlograph_each;
    double* diffraction_mean_abs;
    double* diffraction_mean_sqr;
    double* variance_2d;
} msf_fem_outputs;

/* msf_run_stem with every FEM mode above; fem->enabled must be set for any
 * FEM output to be computed. */
int msf_run_stem_fem(msf_ctx* ctx, const msf_specimen* spec, const msf_optics* optics,
                     const msf_detector* det, const msf_fem* fem,
                     const double* x_p, uint32_t n_xp, const double* y_p, uint32_t n_yp,
                     double* adf_image, const msf_fem_outputs* fem_out);

/* Exit waves psi(k) after the last slice for the same raster, FP32 complex,
 * n_xp*n_yp*nx*ny*2 floats in the reference layout (zeros outside the band
 * limit).  Parity entry point: the reference keeps these only in memory. */
int msf_exit_waves(msf_ctx* ctx, const msf_specimen* spec, const msf_optics* optics,
                   const double* x_p, uint32_t n_xp, const double* y_p, uint32_t n_yp,
                   float* psi);
/* The same for n explicit (sx, sy) shift pairs, e.g. single positions of a
 * raster whose shifts msf_probe_shifts gave: the reference's first- and
 * final-position diffraction images (adfstem.cpp:2140-2230, :3780-3797). */
int msf_exit_waves_shifted(msf_ctx* ctx, const msf_specimen* spec, const msf_optics* optics,
                           const int32_t* shifts, uint32_t n, float* psi);

/* ------------------------------------------------------------ serving */
/* One self-contained job: specimen atoms, optics, raster and detector/FEM
 * settings (inputs are copied by msf_submit), and the caller's output
 * arrays, which must stay valid until msf_wait returns for the job. */
typedef struct msf_job {
    msf_grid grid;
    const double* q_xyz;               /* 3 * n_atoms positions [A] */
    const uint32_t* Z;
    uint32_t n_atoms;
    double min_slice;
    msf_optics optics;
    const double* x_p;
    uint32_t n_xp;
    const double* y_p;
    uint32_t n_yp;
    msf_detector detector;
    msf_fem fem;
    double* adf_image;                 /* n_xp*n_yp doubles when detector.enabled */
    msf_fem_outputs fem_out;           /* optional FEM outputs */
} msf_job;

/* No reference counterpart: the reference runs one job per process.
 * Continuous batching: the first submission starts the context's serving
 * engine, `slots` workers each with its own internal context (stream, run
 * plan).  A free worker takes the job at the head of the queue once the
 * head job's estimated device footprint (msf_job_footprint) fits beside the
 * jobs already running within the memory limit, so jobs that arrive while
 * others run start as slots and memory free up and run concurrently on the
 * device.  A job larger than the limit runs when nothing else does.  Each
 * job runs as msf_specimen_build + msf_run_stem_fem + msf_specimen_free with
 * the submitting context's msf_config at submission; its results do not
 * depend on what runs beside it.  Direct calls on the context do not wait
 * for served jobs.  msf_destroy finishes the queue before it returns. */
int msf_submit(msf_ctx* ctx, const msf_job* job, uint64_t* id);
/* A run on a specimen that stays resident on the device: the transmission
 * functions of msf_specimen_build (replacing tem.cpp:1212-1276, done once per
 * specimen) serve every job that names it, as the reference runs its whole
 * raster on slices built once (adfstem.cpp:1556-2793).  Optics, raster,
 * detector, FEM settings and outputs as in msf_job. */
typedef struct msf_run_job {
    const msf_specimen* specimen;      /* built on the submitting context */
    msf_optics optics;
    const double* x_p;
    uint32_t n_xp;
    const double* y_p;
    uint32_t n_yp;
    msf_detector detector;
    msf_fem fem;
    double* adf_image;                 /* n_xp*n_yp doubles when detector.enabled */
    msf_fem_outputs fem_out;           /* optional FEM outputs */
} msf_run_job;

/* No reference counterpart.  Queues a run on a resident specimen on the
 * same serving engine as msf_submit (same admission, slots and results as
 * msf_run_stem_fem on that specimen).  The specimen's stacks are counted
 * against the memory limit once, while any queued or running job refers to
 * them, instead of per job; msf_specimen_free waits for those jobs. */
int msf_submit_run(msf_ctx* ctx, const msf_run_job* job, uint64_t* id);

/* No reference counterpart (the reference runs one job per process).  Runs
 * n_runs runs on specimens built on ctx whose grids share nx and ny
 * together: each slice step launches the row pass once over the current
 * batches of up to 16 runs, then the column pass ("grouped launches"), while
 * every run keeps its own specimen, vacuum waves, pixel groups, buffers and
 * raster-order statistics.  Each run's outputs equal those of
 * msf_run_stem_fem on it alone (autotune 0 or -1) bit for bit.  status
 * (n_runs entries, may be NULL) gets each run's code; returns MSF_OK when
 * every run succeeded, else the first failing run's code with its message in
 * msf_last_error.  msf_get_stats reports group_runs and group_launches. */
int msf_run_group(msf_ctx* ctx, const msf_run_job* runs, uint32_t n_runs, int32_t* status);

/* MSF_OK or the job's error code once finished, MSF_PENDING while queued or
 * running, MSF_ERR_INVALID for an unknown id */
int msf_poll(msf_ctx* ctx, uint64_t id);
/* blocks until the job has finished; returns its status (the error text is
 * then available from msf_last_error) and forgets the job */
int msf_wait(msf_ctx* ctx, uint64_t id);

/* Timings of a served job: queue_ms from submission to start, run_ms from
 * start to finish (specimen build, run, release), its admitted footprint and
 * the worker slot that ran it. */
typedef struct msf_job_info {
    double queue_ms;
    double run_ms;
    uint64_t footprint;
    uint32_t slot;
} msf_job_info;
/* msf_wait that also returns the job's timings (info may be NULL) */
int msf_wait_info(msf_ctx* ctx, uint64_t id, msf_job_info* info);

/* Serving engine settings, before the first msf_submit on the context:
 * slots = jobs run at once (0: 3, at most 16); mem_limit = device bytes the
 * running jobs may claim together by their estimated footprints (0: 80 % of
 * the device memory free when the engine starts). */
int msf_serve_configure(msf_ctx* ctx, uint32_t slots, uint64_t mem_limit);

typedef struct msf_serve_stats {
    uint32_t slots;
    uint32_t running;                  /* jobs running now */
    uint32_t max_running;              /* most jobs that ran at once */
    uint64_t queued;                   /* jobs waiting for admission */
    uint64_t jobs_done;
    uint64_t mem_limit;
    uint64_t admitted_bytes;           /* footprints of the running jobs */
    uint64_t peak_admitted_bytes;      /* most footprint + resident bytes claimed at once */
    uint64_t specimen_builds;          /* specimens built for msf_submit jobs so far */
    uint64_t resident_bytes;           /* stacks of the resident specimens queued or running
                                          msf_submit_run jobs refer to (counted once each) */
    uint64_t groups;                   /* job groups run with shared launches (msf_serve_group) */
    uint64_t grouped_jobs;             /* jobs that ran in those groups */
    uint64_t group_launches;           /* their grouped slice-pass launches */
} msf_serve_stats;
int msf_serve_get_stats(msf_ctx* ctx, msf_serve_stats* out);

/* No reference counterpart.  Grouping of served jobs, before the first
 * msf_submit on the context: a worker that takes the head job of the queue
 * also takes up to max_jobs - 1 further queued jobs of the same nx, ny and
 * settings that fit the memory limit beside it, and runs them as one group
 * (msf_run_group: shared slice-pass launches, each job's results as alone).
 * 0 or 1 (the default): every job runs alone; at most 16. */
int msf_serve_group(msf_ctx* ctx, uint32_t max_jobs);

/* The admission estimate of a job's device footprint in bytes (specimen
 * stacks, vacuum waves, batch buffers, set-up temporaries, +10 %); host only. */
uint64_t msf_job_footprint(const msf_job* job);

#ifdef __cplusplus
}
#endif
#endif
