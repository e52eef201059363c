// kernels.hpp -- host-callable launchers of the device kernels.
//
// Wave layout of the hot path (FP32 complex), for a batch of B probe
// positions on an nx x ny grid whose slice step keeps the columns
// [0, ka) U [ny - kb, ny) (the 2/3 band limit; every other column of the
// (x, ky) representation is identically zero after the first slice):
//   psi[b][i][jc], jc < ncp, ncp >= ka + kb rounded up to the column tile.
// Between the two passes of a slice the wave is held with x along i and k
// along j.  A slice step is
//   row pass    : [IFFT_j] -> transmit -> FFT_j              (one pass over psi)
//   column pass : FFT_i -> mask * P_x * P_y / s -> [IFFT_i]  (one pass over psi)
// Only the scattered part phi = psi - psi0 of the wave is propagated in FP32;
// the vacuum wave psi0 (t = 1) of every slice is computed once per run in FP64.
// The reference's slice operator is linear over the reals (also with its
// in-place transmission update), so phi obeys
//   phi' = M P F[ T(F^-1 phi) + T(u0) - T_vac(u0) ],   u0 = s * shifted psi0(x),
// and FP32 round-off scales with |phi| instead of |psi|.  The first row pass
// has phi = 0; the last column pass adds the shifted vacuum exit wave back.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace msf {

// Pair layout of the rows the row pass reads beside the wave (ny even): a
// lane holding elements j and j + ny/2 of a row fetches both with one 16-byte
// read.  Transmission rows store element j at pair_pos(j, ny); u0 rows hold,
// for every shift r in [0, ny), the pair (r, (r + ny/2) mod ny).
__host__ __device__ inline int pair_pos(int j, int n)
{
    return j < n / 2 ? 2 * j : 2 * (j - n / 2) + 1;
}

struct RowPassArgs {
    float2* psi;            // [B][nx][ncp] scattered wave phi
    const float2* t;        // [nx][ny] (t_re/s - 1/s, t_im/s) of this slice, rows in the
                            // pair layout
    const float2* u0;       // [nx][ny][2] vacuum wave entering this slice, s * real space,
                            // rows as shift pairs (pair_pos)

    const int2* shift;      // [B] probe shifts of the batch
    const float2* tw;       // stage twiddle table of length ny (stage_twiddles)
    float inv_s;            // 1/sqrtNxNy
    int nx, ncp, ka, kb, batch;
    int prune;              // 1: columns [3ny/8, 5ny/8) are outside the kept band
    int plan;               // radix plan of length ny: 0 primary, 1 alternate (has_alt_plan)
};

struct ColPassArgs {
    float2* psi;            // [B][nx][ncp]
    const float2* px;       // [nx] propagator along x of this slice
    const float2* py;       // [ny] propagator along y of this slice, scaled by 1/s
    const int2* colkeep;    // [ncp] (c, d): kept rows [0, c) U [nx - d, nx) of compact column jc
    const float2* tw;       // stage twiddle table of length nx (stage_twiddles)
    const float2* psi0;     // [ny][nx] vacuum exit wave in k-space, column-major (last pass only)
    const int2* shift;      // [B] probe shifts (last pass only)
    const float2* twx_nat;  // [nx] roots W_nx^e (last pass only)
    const float2* twy;      // [ny] roots W_ny^e (last pass only)
    int ny, ncp, ka, kb, batch;
    int tile;               // columns per workgroup (4, 8 or 16; 0 = default)
    int prune;              // 1: rows [3nx/8, 5nx/8) are masked for every column
    int plan;               // radix plan of length nx: 0 primary, 1 alternate (has_alt_plan)
};

// One launch with the row pass of one batch (ra) and the column pass of
// another (ca), same transform length n = nx = ny: the rows' FIRST and the
// columns' LAST flags as for the separate passes.  hipErrorNotSupported when
// fused_supported() is false for these arguments.
hipError_t launch_fused_pass(int n, bool first, bool last, int tmode, const RowPassArgs& ra,
                             const ColPassArgs& ca, hipStream_t s);
bool fused_supported(int n, const RowPassArgs& ra, const ColPassArgs& ca);
// true when length n has fused pass pairs for row plan vr, column plan vc and
// column tile `tile` (0: the plan's default)
bool fused_shape_ok(int n, int vr, int vc, int tile);

// Grouped launches: one launch runs the same pass over the batches of up to
// kGroupMax runs of one transform length (each its own specimen, vacuum
// waves, kept columns and buffers).  Entry k owns launch blocks
// [off[k], off[k+1]), its first block a multiple of 8 (XCD banding within
// the entry); blocks past an entry's own count return at once.
constexpr int kGroupMax = 16;
struct RowGroupArgs {
    int n;
    int off[kGroupMax + 1];
    RowPassArgs a[kGroupMax];
};
struct ColGroupArgs {
    int n;
    int off[kGroupMax + 1];
    ColPassArgs a[kGroupMax];
};
// blocks of one entry (before the padding to a multiple of 8)
int row_group_blocks(int ny, const RowPassArgs& a);
int col_group_blocks(int nx, const ColPassArgs& a);
// Every entry: same length, plan, prune flag and (column pass) tile; the
// rows' FIRST and the columns' LAST flags shared by the launch.
hipError_t launch_row_group(int ny, bool first, int tmode, const RowGroupArgs& g, hipStream_t s);
hipError_t launch_col_group(int nx, bool last, const ColGroupArgs& g, hipStream_t s);
// the column tile a grouped column pass uses with a configured tile `tile`
// (0: the plan's default)
int col_group_tile(int nx, int v, int tile);

bool slice_lengths_ok(int n);
// concatenated per-stage twiddle tables of plan v (0 primary, 1 alternate) of
// length n, (re, im) FP64
bool stage_twiddles(int n, std::vector<double>& re_im, int v = 0);
// true when length n has an alternate radix plan for the slice passes
bool has_alt_plan(int n);
// the column count of the compact wave is padded to a multiple of this
int col_tile(int nx);
// the column-pass tile width used when none is configured, for plan v
int col_tile_default(int nx, int v = 0);
// the tile used when none is configured: the default, or one column per
// workgroup when the launch would not give every CU a workgroup
int col_tile_auto(int nx, int v, int batch, int ncp);
// true when the kept set leaves [3n/8, 5n/8) empty along the row (ka, kb) and
// every row i in [3nx/8, 5nx/8) fully masked (keep ranges)
bool row_prunable(int ny, int ka, int kb);
bool col_prunable(int nx, const int32_t* keep_ab);

// ev0/ev1 (optional): events the dispatch itself records at the kernel's
// start and end, for profiling without marker overhead
hipError_t launch_row_pass(int ny, bool first, int transmit_mode, const RowPassArgs& a,
                           hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_col_pass(int nx, bool last, const ColPassArgs& a, hipStream_t s,
                           hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// per (position, segment) FP64 sums over interleaved pixel tiles:
// kind 0 -> |psi|^2, 1 -> |psi|; then per (position, group) the fixed-order
// sum of the group's segments
hipError_t launch_reduce_segments(const float2* psi, int64_t plane, const int32_t* pix,
                                  const int32_t* tile_off, const int32_t* seg_kind, int n_tiles,
                                  int batch, double* seg_out, hipStream_t s);
// (small: the groups of at most 64 segments, one lane each; large: the rest,
// one wavefront each; combine_split forms both lists from seg_of_group)
hipError_t launch_combine_segments(const double* seg_in, int n_seg_stride,
                                   const int32_t* seg_of_group, const int32_t* small, int n_small,
                                   const int32_t* large, int n_large, int groups, int batch,
                                   double* out, hipStream_t s);
void combine_split(const std::vector<int32_t>& seg_of_group, std::vector<int32_t>& small,
                   std::vector<int32_t>& large);
// ADF column of the partial sums -> adf[pos0 + b]
hipError_t launch_collect_adf(const double* partial, int groups, int batch, double* adf,
                              hipStream_t s);
// FEM Welford update over the batch in raster order; state[3*nbins] = (n, mean, M2)
hipError_t launch_fem_welford(const double* partial, int groups, int batch, int nbins,
                              const int32_t* counts, const uint8_t* zeroed, double* state,
                              hipStream_t s);
hipError_t launch_fem_finalize(const double* state, int nbins, double* v, hipStream_t s);
// Vbar_r accumulation over a batch (per-position ring ratios, raster order)
hipError_t launch_fem_vbar(const double* partial, int groups, int batch, int nbins, int off_abs,
                           int off_sqr, const int32_t* counts, const uint8_t* zeroed,
                           double* vsum, hipStream_t s);
// FP64 sums over pixel groups of count full-grid FP64 waves: out[count][groups]
hipError_t launch_group_sums_f64(const double2* psi, int64_t plane, const int32_t* offsets,
                                 const int32_t* pixels, const int32_t* kind, int groups,
                                 int count, double* out, hipStream_t s);
// detector MTF steps: |psi| outside the aperture from the compact FP32 planes
// into full FP64 planes; the same in place on full FP64 planes; / s with the
// band limit back into compact FP32 planes; a /= d
hipError_t launch_mtf_magnitude(const float2* psi, int nx, int ny, int ncp, int ka, int kb,
                                const uint8_t* outside, int batch, double2* out, hipStream_t s);
hipError_t launch_mtf_magnitude_f64(double2* a, int64_t m, int count, const uint8_t* outside,
                                    hipStream_t s);
hipError_t launch_mtf_compact(const double2* full, int nx, int ny, int ncp, int ka, int kb,
                              const int2* keep, double s, int batch, float2* psi, hipStream_t st);
hipError_t launch_div_f64(double2* a, int64_t n, double d, hipStream_t s);
// correlograph sums of c and c^2 per (ring, phi sector) over a batch; with
// `each`, also c of every position of the batch ([b][ii + jj * nbins])
hipError_t launch_fem_correlograph(const double* partial, int groups, int batch, int nbins,
                                   int nphi, int off_abs, int off_corr, const int32_t* counts,
                                   const uint8_t* zeroed, double* csum, double* csqr,
                                   double* each, hipStream_t s);
// per-pixel FP64 sums of |psi| and |psi|^2 over the positions of a batch
hipError_t launch_pixel_moments(const float2* psi, int64_t plane, int batch, double* s1,
                                double* s2, hipStream_t s);
// compact exit wave -> full reference layout (FP32)
hipError_t launch_expand_exit(const float2* psi, int nx, int ny, int ncp, int ka, int kb,
                              int batch, float2* out, hipStream_t s);

// ---------------------------------------------------------------- FP64 set-up
bool setup_lengths_ok(int n);
// in-place 2-D FFT of `count` planes of nx x ny complex doubles (twx, twy:
// stage twiddle tables)
hipError_t fft2d_f64(double2* data, int nx, int ny, int count, bool inverse, const double2* twx,
                     const double2* twy, hipStream_t s);
// probe H(k) of probe_wavefunction.cpp:37-239 inside the aperture ranges
hipError_t launch_probe_k(double2* h, const double* kx, const double* ky, int nx, int ny,
                          const int2* aperture, int corrected, double cs3, double cs5,
                          double defocus, double lambda, double x0, double y0,
                          hipStream_t s);
// sum |a|^2 over n values (one workgroup, fixed order)
hipError_t launch_norm2(const double2* a, int64_t n, double* out, hipStream_t s);
// probe_s = (float) (a * scale)
hipError_t launch_scale_to_f32(const double2* a, int64_t n, const double* norm2,
                               double numerator, float2* out, hipStream_t s);
hipError_t launch_fill_f64(double2* a, int64_t n, hipStream_t s);
// out[0, n16) = in[0, n16), 16-byte vectors; n16 a multiple of 1024
hipError_t launch_stream_copy(const float4* in, float4* out, int64_t n16, hipStream_t s);
// atom count maps: maps[(slot*nsp + sp)][sx][sy] += 1
hipError_t launch_scatter_atoms(double2* maps, const int4* atoms, int n_atoms, int nx, int ny,
                                int nsp, hipStream_t s);
// sigma_k[slot] = sum_sp T_sp(k) C_sp(k)   (scatterer_pap.hpp:104-130 template)
hipError_t launch_combine_species(const double2* maps, double2* sigma, int slots, int nsp,
                                  const double* params /* nsp*12 */, const double* kx,
                                  const double* ky, int nx, int ny, const int2* support,
                                  double lgab, double xmin, double ymin, hipStream_t s);
// t = exp(-Im v) (cos Re v, sin Re v), v = sigma * scale
hipError_t launch_exp_transmit(double2* a, int64_t n, double scale, hipStream_t s);
// zero outside the kept ranges
hipError_t launch_apply_keep(double2* a, int nx, int ny, int count, const int2* keep,
                             hipStream_t s);
// t = a / nxny (FP64, kept for queries); t32 = (t_re/s - 1/s, t_im/s) in FP32,
// each row of ny in the pair layout (pair_pos)
hipError_t launch_finish_transmit(double2* a, int64_t n, int ny, double nxny, double s,
                                  float2* t32, hipStream_t st);
// one vacuum slice step in k-space: a *= keep ? P_x[i] P_y[j] * scale : 0
hipError_t launch_vacuum_step(double2* a, int nx, int ny, const double2* px, const double2* py,
                              const int2* keep, double scale, hipStream_t s);
// FP64 transmission update of `count` real-space waves by one slice's t
hipError_t launch_transmit_f64(double2* a, const double2* t, int64_t m, int count, int mode,
                               double s, hipStream_t st);
// FP64 [nx][ny] -> FP32 [ny][nx]
hipError_t launch_to_f32_transposed(const double2* a, int nx, int ny, float2* out, hipStream_t s);
// out[i][2r + h] = (float2) a[i][(r + h ny/2) mod ny], r < ny, h = 0, 1: the
// pair of elements r and r + ny/2 for every shift r of a row (u0 rows)
hipError_t launch_to_f32_shift_pairs(const double2* a, int nx, int ny, float2* out, hipStream_t s);
// a *= f / sqrt(*norm2)
hipError_t launch_scale_f64(double2* a, int64_t n, const double* norm2, double f, hipStream_t s);

} // namespace msf
