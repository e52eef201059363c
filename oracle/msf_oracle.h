/*
 * msf_oracle.h -- FP64 CPU restatement of ms-stem-fem's multislice ADF-STEM /
 * FEM path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark call it as the checker.  The product
 * library never links, loads or calls anything in this directory.
 *
 * Every function names the reference lines it restates.  Complex arrays are
 * interleaved (re, im) doubles in the reference's row-major [i*Ny + j] layout,
 * i = x row, j = y column.  FFT sign convention is FFTW's: sign = -1 is
 * FFTW_FORWARD (e^{-2 pi i}), sign = +1 is FFTW_BACKWARD; both unnormalised.
 */
#ifndef MSF_ORACLE_H
#define MSF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One simulation job, the inputs tem.cpp hands to adfstem(). */
typedef struct orc_job {
    int32_t nx, ny;                  /* samples */
    double xperiod, yperiod, zperiod; /* cell periods [A] (no --dupe) */
    double xmin, ymin, zmin;          /* domain minima [A] */
    double voltage;                   /* [V] */
    double min_slice;                 /* --minslice [A] */
    int32_t corrected;                /* 0: adfstemuncorr, 1: adfstemcorr */
    double cs3, cs5, defocus, alpha_max;
    int32_t adf;                      /* --images: ADF image */
    double det_inner, det_outer;      /* [rad] */
    int32_t fem;                      /* V_omega (d1) */
    double dr;                        /* azimuthal binning size factor */
    int32_t transmit_mode;            /* 0: reference in-place update, 1: exact complex product */
    int32_t adf_mode;                 /* 0: reference k^2 vs angle/lambda, 1: |k| vs angle/lambda */
    /* detector MTF (--mtf_file, --mtf_sampling_rate): active when mtf_n >= 1
     * and mtf_rate != 0; mtf_domain in units of the Nyquist frequency as read */
    const double* mtf;
    const double* mtf_domain;
    int32_t mtf_n;
    double mtf_rate;
} orc_job;

/* ---- scalars and domains ---- */
void orc_wavelength(double voltage, double* lambda, double* gamma);
void orc_bwcutoffs(int nx, int ny, double xperiod, double yperiod,
                   double* bwcutoff_pap, double* bwcutoff_t);
void orc_domains(int nx, int ny, double xperiod, double yperiod,
                 double xmin, double ymin,
                 double* kx, double* ky, double* xx, double* yy);
double orc_scherzer_defocus_uncorrected(double lambda, double cs3);
double orc_scherzer_alphamax_uncorrected(double lambda, double cs3);
double orc_scherzer_defocus_correctedtoCs5(double lambda, double cs5);
double orc_scherzer_alphamax_correctedtoCs5(double lambda, double cs5);
double orc_scherzer_cs3_correctedtoCs5(double lambda, double cs5);
void orc_raster(int n, double period, double qmin, double* q);

/* ---- FFT (own mixed-radix, FP64) ---- */
int orc_fft1d(double* data, int n, int sign);
int orc_fft2d(double* data, int nx, int ny, int sign);

/* ---- potentials and slices ---- */
double orc_fe(int Z, double ksq);
int orc_template(int Z, double lambda, double gamma, double ab_inv, double cutoff,
                 const double* kx, int nx, const double* ky, int ny,
                 double xmin, double ymin, double* out);
int orc_autoslice(const double* q, int n, double zmin, double zperiod,
                  double min_thickness, int32_t* slice_of_atom,
                  double* lower_bound, double* thickness, int32_t* n_slices);
void orc_propagator(double lambda, double thickness, const double* k, int n,
                    double bwcutoff_t, double* p);
int orc_shift_index(const double* grid, int n, double q, int64_t previous);
int orc_transmission(const double* q, const int32_t* Z, const int32_t* atoms, int n_atoms,
                     const int32_t* zlist, int nz, const double* templates,
                     const double* kx, const double* ky, const double* xx, const double* yy,
                     int nx, int ny, double bwcutoff_t, double* t);

/* ---- probe ---- */
int orc_probe(const orc_job* job, const double* kx, const double* ky,
              const double* xx, const double* yy, double* probe);
void orc_probe_shifts(const double* xx, int nx, const double* yy, int ny,
                      const double* x_p, int nxp, const double* y_p, int nyp,
                      int64_t* sx, int64_t* sy);

/* ---- slice loop, detectors ---- */
int orc_initial_wave(const double* probe, int64_t sx, int64_t sy, int nx, int ny, double* psi);
int orc_slice_step(double* psi, const double* t, const double* px, const double* py,
                   const double* kx, const double* ky, int nx, int ny, double bwcutoff_t,
                   int transmit_mode);
int orc_propagate_position(const double* probe, int64_t sx, int64_t sy,
                           int n_slices, const double* t_stack,
                           const double* px, const double* py,
                           const double* kx, const double* ky, int nx, int ny,
                           double bwcutoff_t, int transmit_mode, double* psi);
double orc_adf(const double* psi, const double* kx, const double* ky,
               int nx, int ny, double k1, double k2);
int orc_fem_nbins(const double* kx, const double* ky, double dr, double bwcutoff_t,
                  double* delta_k);
int orc_fem_radial(const double* psi, const double* kx, const double* ky, int nx, int ny,
                   double delta_k, int nbins, double lambda, double alpha_max,
                   double* radial, int32_t* counts);
void orc_fem_finalize(const double* sum_r, const double* sum_r2, int nbins,
                      int64_t n_positions, double* v);
int orc_fem_modes(const double* mag2, long n_positions, const double* kx, const double* ky,
                  int nx, int ny, double delta_k, int nbins, double lambda, double alpha_max,
                  const double* radial1, const int32_t* counts1, double* vbar_r, double* v_re,
                  double* omega_vimage, double* v_gt17, double* v_rbar);

/* ---- whole job (tem.cpp setup + adfstem()) ---- */
typedef struct orc_specimen {
    int32_t n_slices;
    double* t_stack;     /* n_slices * nx*ny complex */
    double* px;          /* n_slices * nx complex */
    double* py;          /* n_slices * ny complex */
    double* lower_bound; /* n_slices */
    double* thickness;   /* n_slices */
} orc_specimen;

int orc_specimen_build(const orc_job* job, const double* q, const int32_t* Z, int n_atoms,
                       orc_specimen* out);
void orc_specimen_free(orc_specimen* s);
int orc_adfstem(const orc_job* job, const orc_specimen* spec,
                const double* x_p, int nxp, const double* y_p, int nyp,
                int n_threads, double* adf_image, double* fem_v, double* exit_psi);
/* the same with the other FEM modes (Vbar_r, V_re, Omega_vimage, V_gt17,
 * V_rbar); each output may be NULL */
int orc_adfstem_fem(const orc_job* job, const orc_specimen* spec,
                    const double* x_p, int nxp, const double* y_p, int nyp, int n_threads,
                    double* adf_image, double* fem_v, double* vbar_r, double* v_re,
                    double* omega_vimage, double* v_gt17, double* v_rbar, double* corr_avg,
                    double* corr_sqr, double* corr_each, double* exit_psi);
int orc_correlograph_nphi(double delta_k, int nbins);
/* the 2-D detector MTF (adfstem.cpp:300-439, tem.cpp:276-285), complex [nx][ny] */
int orc_mtf_2d(const double* mtf, const double* mtf_domain, int n, double rate, const double* kx,
               const double* ky, int nx, int ny, double xperiod, double yperiod,
               double bwcutoff_t, double* out);
/* its application to one k-space exit wave in place (adfstem.cpp:2272-2385) */
int orc_apply_mtf(double* psi, const double* mtf2d, const double* kx, const double* ky, int nx,
                  int ny, double lambda, double alpha_max, double bwcutoff_t);
int orc_correlograph(const double* mag2, long n_positions, const double* kx, const double* ky,
                     int nx, int ny, double delta_k, int nbins, double lambda, double alpha_max,
                     const double* radial1, const int32_t* counts1, double* avg, double* sqr,
                     double* each);

#ifdef __cplusplus
}
#endif
#endif
