// msf_host.hpp -- host-side FP64 rules of the multislice path.
//
// Everything here is O(N) or O(N^2) bookkeeping that decides *which* pixel
// or atom goes where: grids, band-limit membership, the nearest-pixel shift
// rule, the auto-slicer, propagator tables, detector and FEM bin pixel
// lists.  These decisions are taken on the host in IEEE FP64 with the same
// expressions and comparison operators as the reference, so the device
// kernels only ever consume integer index data and never decide membership
// in FP32.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace msf {

constexpr double kRefPi = 3.14159265369; // the reference's PI macro (slice.cpp:41 and others)

// libm pow(x, 2.0), as the reference's -O0 build calls it (build/Makefile:37-45);
// compilers fold a literal pow(x, 2) into x*x, which rounds differently.
double ref_pow2(double x);
// ref_pow2 of every element
std::vector<double> pow2_all(const std::vector<double>& k);

struct Domain {
    int nx = 0, ny = 0;
    double lambda = 0, gamma = 0;
    double bw_pap = 0, bw_t = 0;    // tem.cpp:683-695
    double sqrt_nxny = 0;           // adfstem.cpp:210-212
    double nxny = 0;                // tem.cpp:405
    std::vector<double> kx, ky, xx, yy;
};

// tem.cpp:424-438
void wavelength(double voltage, double& lambda, double& gamma);
// tem.cpp:683-695, :756-883, func.cpp:92-106, :147-167
Domain make_domain(int nx, int ny, double xperiod, double yperiod, double xmin, double ymin,
                   double voltage);

// Row-wise membership of a disk-like pixel set in FFT order: for row i the
// kept columns are [0, a_i) U [ny - b_i, ny).  Built by testing every pixel
// with the caller's predicate; ok == false if a row is not of that form.
struct RowRanges {
    std::vector<int32_t> ab; // 2*nx: a_i, b_i
    int amax = 0, bmax = 0;  // union over rows
    bool ok = true;
};
template <typename Pred> RowRanges row_ranges(int nx, int ny, Pred keep)
{
    RowRanges r;
    r.ab.assign(2 * (size_t)nx, 0);
    for (int i = 0; i < nx; ++i) {
        int a = 0;
        while (a < ny && keep(i, a)) ++a;
        int b = 0;
        if (a < ny)
            while (b < ny - a && keep(i, ny - 1 - b)) ++b;
        for (int j = a; j < ny - b; ++j)
            if (keep(i, j)) r.ok = false;
        r.ab[2 * i] = a;
        r.ab[2 * i + 1] = b;
        if (a > r.amax) r.amax = a;
        if (b > r.bmax) r.bmax = b;
    }
    return r;
}

// bw_limit.cpp:30-47 together with slice::propagate's test (slice.cpp:1388-1389):
// the pixels a slice step keeps.
RowRanges slice_keep_ranges(const Domain& d);
// The same kept set by column of the compact plane [nx][ncp] (compact
// column jc holds full column j = jc < ka ? jc : jc - (ka + kb) + ny): the
// kept rows of column jc are [0, c) U [nx - d, nx), returned as (c, d) pairs
// (2*ncp entries; padding columns keep nothing).  ok == false if a column is
// not of that form; the pairs then are not used.
RowRanges column_keep_ranges(const RowRanges& rows, int nx, int ny, int ncp);

// bw_limit.cpp:30-47 alone (transmission band limit, slice.cpp:1731-1737)
RowRanges bwlimit_keep_ranges(const Domain& d);
// scatterer_pap.hpp:108-109: template support |k|^2 < bwcutoff_pap^2
RowRanges template_keep_ranges(const Domain& d);
// probe_wavefunction.cpp:107-111: aperture lambda^2 k^2 < alpha_max^2
RowRanges aperture_keep_ranges(const Domain& d, double alpha_max);

// Nearest-pixel rule, slice.cpp:1554-1572 and adfstem.cpp:1590-1606: first i
// with grid[i] - q >= half spacing, minus one; `previous` if none.  Returns
// -1 where the reference's size_t index would underflow (q below the domain).
int64_t shift_index(const std::vector<double>& grid, double q, int64_t previous);

// slice.cpp:469-740 (assign_atoms_to_slices_z_auto)
struct Slices {
    std::vector<double> lower, thick;
    std::vector<std::vector<int32_t>> atoms; // per slice, in z-sorted order
};
Slices autoslice(const double* q, uint32_t n, double zmin, double zperiod, double min_thickness);

// slice.cpp:1293-1363 (update_propagator), one axis, FP64 (re, im) pairs
void propagator(double lambda, double thickness, const std::vector<double>& k, double bw_t,
                std::vector<double>& p);

// adfstem.cpp:947-972
int fem_bin_count(const Domain& d, double dr, double& delta_k);

// Pixel groups for the reductions: group 0 = ADF ring (func.cpp:2630-2637),
// groups 1..nbins = FEM rings (variance.cpp:160-228 membership,
// adfstem.cpp:1043-1062 sort key).  Offsets are into the compact exit-wave
// plane [nx][ncp], only pixels that can be non-zero after the slice step.
struct Groups {
    std::vector<int32_t> offsets; // G+1
    std::vector<int32_t> pixels;  // compact plane offsets
    std::vector<int32_t> kind;    // 0: sum |psi|^2, 1: sum |psi|
    std::vector<int32_t> counts;  // FEM: full-grid bin counts (k_bin_counts)
    std::vector<uint8_t> zeroed;  // FEM: direct-beam bins (adfstem.cpp:2518-2522)
    std::vector<double> edges;    // FEM: lower edges ii*delta_k
    int nbins = 0;
    int fem_abs = -1;             // first group of the FEM |psi| rings (-1: none)
    int fem_sqr = -1;             // first group of the FEM |psi|^2 rings (-1: none)
    int corr = -1;                // first correlograph sector group, [ii * nphi + jj] (-1: none)
    int nphi = 0;                 // correlograph phi sectors per ring
    // every pixel of every FEM bin in the bins' sorted order (bin ii owns
    // [bin_off[ii], bin_off[ii+1])): full-grid index i*ny + j and compact
    // offset, -1 where the slice step's band limit leaves the pixel zero
    std::vector<int32_t> bin_off, bin_full, bin_compact;
};
// group 0: the ADF ring; then, with fem, nbins |psi| rings and, with
// fem_sqr, nbins |psi|^2 rings over the same pixels
Groups build_groups(const Domain& d, const RowRanges& keep, int ncp, bool adf, double k1,
                    double k2, bool fem, double dr, double alpha_max, bool fem_sqr = false,
                    bool correlograph = false);
// adfstem.cpp:969-985: the correlograph's phi sector count for nbins rings of width delta_k
int correlograph_nphi(double delta_k, int nbins);

// The detector MTF sampled on the grid, real [nx][ny] (adfstem.cpp:300-376;
// the squared domain (0.5 rate d)^2 of tem.cpp:276-285 formed from the
// file's Nyquist units).  n >= 1.
std::vector<double> mtf_table_2d(const Domain& d, double xperiod, double yperiod,
                                 const double* mtf, const double* domain, int n, double rate);
// adfstem.cpp:2279-2290: 1 where lambda^2 k^2 > alpha_max^2 (kept as |psi|), else 0
std::vector<uint8_t> outside_aperture(const Domain& d, double alpha_max);

// Groups cut into segments of at most `seg` pixels, 64 segments per wavefront
// tile, pixel offsets interleaved [tile][k][lane] (-1 pads short segments) so
// that a wavefront's index loads are coalesced.  Segments of one group are
// consecutive; group g owns segments [seg_of_group[g], seg_of_group[g+1]).
struct SegmentTiles {
    std::vector<int32_t> pix;         // interleaved pixel offsets
    std::vector<int32_t> tile_off;    // n_tiles + 1, into pix (units of 64 entries)
    std::vector<int32_t> seg_kind;    // per segment: 0 |psi|^2, 1 |psi|
    std::vector<int32_t> seg_of_group;
    int n_segments = 0, n_tiles = 0;
};
SegmentTiles build_segments(const Groups& g, int seg);

} // namespace msf
