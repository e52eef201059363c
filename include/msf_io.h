/*
 * msf_io.h -- host-side model readers and result writers of the
 * ms-stem-fem command line (no device code; usable without a GPU).
 *
 * Readers replace read_position_xyz_file (tem_io.cpp:4627-4956) and
 * read_position_lammps_data_file (tem_io.cpp:36-1028) with the same
 * acceptance rules and coordinate shifts.  Writers replace
 * writeImage::write_tif_grayscale<uint16_t> (write_tif.hpp:143-335, 16-bit
 * min-max grayscale, linear and log), output_variance_to_txt
 * (io_txt.cpp:30-92) and add a raw float64 sidecar.
 *
 * Every function returns MSF_OK (0) or a negative MSF_ERR_* code from
 * msf.h; msf_io_last_error() gives the message of the calling thread's
 * last failure.
 */
#ifndef MSF_IO_H
#define MSF_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Atoms of a model file after the reader's shifts: q holds 3*n_atoms
 * positions [A], Z the atomic numbers.  Owned by the library. */
typedef struct msf_model {
    uint32_t n_atoms;
    double* q;
    uint32_t* Z;
    double xperiod, yperiod, zperiod;
    double xmin, ymin, zmin;       /* always 0 after reading, as the reference sets them */
} msf_model;

const char* msf_io_last_error(void);

/* atom_element_abbrev_to_Z (tem_io.cpp:6171-6385): element symbol in any
 * case or decimal atomic number; 0 if not an element 1..103 */
uint32_t msf_element_z(const char* name);
/* guessZ (tem_io.cpp:1030-1257): atomic number from a LAMMPS mass */
int msf_guess_z(double mass, uint32_t* Z);

/* read_position_xyz_file (tem_io.cpp:4627-4956) */
int msf_read_xyz(const char* path, msf_model* out);
/* read_position_lammps_data_file (tem_io.cpp:36-1028) */
int msf_read_lammps_data(const char* path, msf_model* out);
void msf_model_free(msf_model* m);

/* A detector MTF table: n spatial frequencies (Nyquist units) and values.
 * Owned by the library. */
typedef struct msf_mtf_table {
    uint32_t n;
    double* domain;
    double* mtf;
} msf_mtf_table;

/* read_mtf_file (tem_io.cpp:5722-5800): the first line holds the domain, the
 * second the MTF values; commas count as spaces; each line is read with
 * repeated stream extraction until its end, so trailing blanks repeat the
 * line's last value.  As with every reader here a line counts only if a
 * newline follows it.  Unequal lengths are an error, and so are an empty
 * table (which the reference would index out of bounds) and a non-numeric
 * entry (on which the reference's read loop never ends). */
int msf_read_mtf(const char* path, msf_mtf_table* out);
void msf_mtf_table_free(msf_mtf_table* t);

/* write_tif_grayscale<uint16_t> (write_tif.hpp:143-335): data is a
 * row-major height x width array scaled from [min, max] to 0..65535
 * (logscale: log(x - min + 1) / log(max - min + 1)); rows are stored
 * bottom-up as the reference writes them.  resolution_unit 1/2/3 and
 * pixels per unit along x (width) and y (height). */
int msf_write_tiff16(const char* path, const double* data, uint32_t height, uint32_t width,
                     double min, double max, int logscale, uint32_t resolution_unit,
                     double x_resolution, double y_resolution);

/* output_variance_to_txt (io_txt.cpp:30-92): "%.10e %.10e\n" per bin:
 * lower edge, value */
int msf_write_variance_txt(const char* path, const double* values, const double* lower_edges,
                           uint32_t n_bins);

/* output_correlograph_to_txt (io_txt.cpp:96-165): first line "0 " and the
 * lower phi boundaries phi_edges[jj] (jj < n_phi), then one line per k bin ii
 * (ii < n_k): k_edges[ii] and data[ii + jj * n_k] for every jj; every value
 * "%.10e " and every line ending in a space before the newline. */
int msf_write_correlograph_txt(const char* path, const double* data, const double* k_edges,
                               uint32_t n_k, const double* phi_edges, uint32_t n_phi);

/* float64 sidecar: a one-line text header "msf-f64 <rows> <cols>\n"
 * followed by rows*cols little-endian doubles, row-major */
int msf_write_f64(const char* path, const double* data, uint32_t rows, uint32_t cols);

#ifdef __cplusplus
}
#endif
#endif
