// kernels_setup.hip -- FP64 specimen and probe set-up on the device.
//
// Transmission functions (slice.cpp:1451-1757).  The reference sums, for every
// atom of a slice, its species template circularly shifted to the atom's
// nearest pixel (slice.cpp:1594-1648): sigma_v = sum_sp tmpl_sp (*) c_sp with
// c_sp the per-pixel atom count map of species sp.  With tmpl_sp =
// IFFT(T_sp)/sqrt(NxNy) (scatterer_pap.hpp:104-174) the same sum is
//     sigma_v = IFFT( sum_sp T_sp(k) * FFT(c_sp)(k) ) / sqrt(NxNy),
// which costs O(N^2 log N) per slice instead of O(atoms * N^2).  The rest
// follows the reference: t = exp(-Im v)(cos Re v, sin Re v), FORWARD FFT,
// bw_limit, BACKWARD FFT, / NxNy.  Everything here is FP64.
//
// Probe (adfstem.cpp:586-817): H(k) inside the aperture, BACKWARD FFT, norm.
#include "fft_engine.hpp"
#include "kernels.hpp"

namespace msf {

// ------------------------------------------------------------ 2-D FFT, FP64
template <int NY, bool INV>
__global__ void __launch_bounds__(256) fft64_rows_kernel(double2* data, int64_t rows,
                                                         const double2* tw)
{
    using P = Plan<NY>;
    constexpr int E = P::E, T = P::T;
    constexpr int RPW = 256 / T;
    constexpr int LEN = lds_len(NY);
    __shared__ double2 s_tw[stage_tw_len<NY>()];
    __shared__ double2 s_buf[RPW * LEN];
    for (int e = threadIdx.x; e < stage_tw_len<NY>(); e += blockDim.x) s_tw[e] = tw[e];
    __syncthreads();
    const int lr = threadIdx.x / T, t = threadIdx.x % T;
    const int64_t r = (int64_t)blockIdx.x * RPW + lr;
    const bool active = r < rows;
    double2* row = data + (active ? r : 0) * NY;
    double2 x[E];
#pragma unroll
    for (int p = 0; p < E; ++p) x[p] = active ? row[t + T * p] : double2{0.0, 0.0};
    P::template run<INV>(x, s_buf + lr * LEN, s_tw, t);
    if (active) {
#pragma unroll
        for (int p = 0; p < E; ++p) row[t + T * p] = x[p];
    }
}

template <int NX, bool INV>
__global__ void __launch_bounds__(256) fft64_cols_kernel(double2* data, int ny, int count,
                                                         const double2* tw)
{
    using P = Plan<NX>;
    constexpr int E = P::E, T = P::T;
    constexpr int W = (256 / T) < 4 ? (256 / T) : 4;
    constexpr int LEN = lds_len(NX);
    __shared__ double2 s_tw[stage_tw_len<NX>()];
    __shared__ double2 s_buf[W * LEN];
    for (int e = threadIdx.x; e < stage_tw_len<NX>(); e += blockDim.x) s_tw[e] = tw[e];
    __syncthreads();
    const int c = threadIdx.x % W, t = threadIdx.x / W;
    const int ntiles = ny / W;
    const int plane = blockIdx.x / ntiles;
    const int j = (blockIdx.x % ntiles) * W + c;
    const bool active = plane < count;
    double2* col = data + (int64_t)(active ? plane : 0) * NX * ny + j;
    double2 x[E];
#pragma unroll
    for (int p = 0; p < E; ++p) x[p] = active ? col[(int64_t)(t + T * p) * ny] : double2{0.0, 0.0};
    P::template run<INV>(x, s_buf + c * LEN, s_tw, t);
    if (active) {
#pragma unroll
        for (int p = 0; p < E; ++p) col[(int64_t)(t + T * p) * ny] = x[p];
    }
}

template <int N> static constexpr int col_w64() { return (256 / Plan<N>::T) < 4 ? (256 / Plan<N>::T) : 4; }

template <int NY>
static hipError_t rows64(double2* d, int64_t rows, bool inv, const double2* tw, hipStream_t s)
{
    constexpr int RPW = 256 / Plan<NY>::T;
    const dim3 grid((unsigned)((rows + RPW - 1) / RPW)), block(RPW * Plan<NY>::T);
    if (inv) hipLaunchKernelGGL((fft64_rows_kernel<NY, true>), grid, block, 0, s, d, rows, tw);
    else hipLaunchKernelGGL((fft64_rows_kernel<NY, false>), grid, block, 0, s, d, rows, tw);
    return hipGetLastError();
}

template <int NX>
static hipError_t cols64(double2* d, int ny, int count, bool inv, const double2* tw, hipStream_t s)
{
    constexpr int W = col_w64<NX>();
    if (ny % W) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((int64_t)count * (ny / W))), block(W * Plan<NX>::T);
    if (inv) hipLaunchKernelGGL((fft64_cols_kernel<NX, true>), grid, block, 0, s, d, ny, count, tw);
    else hipLaunchKernelGGL((fft64_cols_kernel<NX, false>), grid, block, 0, s, d, ny, count, tw);
    return hipGetLastError();
}

bool setup_lengths_ok(int n) { return slice_lengths_ok(n); }

hipError_t fft2d_f64(double2* data, int nx, int ny, int count, bool inverse, const double2* twx,
                     const double2* twy, hipStream_t s)
{
    hipError_t e = hipErrorInvalidValue;
    const int64_t rows = (int64_t)count * nx;
    switch (ny) {
#define MSF_CASE(N) case N: e = rows64<N>(data, rows, inverse, twy, s); break;
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    switch (nx) {
#define MSF_CASE(N) case N: return cols64<N>(data, ny, count, inverse, twx, s);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------ probe
constexpr double kPi = 3.14159265369; // the reference's PI

// probe_wavefunction.cpp:37-137 / :139-239 with lenses.cpp:249-290
__global__ void probe_k_kernel(double2* h, const double* kx, const double* ky, int nx, int ny,
                               const int2* ap, int corrected, double cs3, double cs5,
                               double defocus, double lambda, double x0, double y0)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nx * ny) return;
    const int i = (int)(idx / ny), j = (int)(idx % ny);
    const int2 r = ap[i];
    double2 v{0.0, 0.0};
    if (j < r.x || j >= ny - r.y) {
        const double lsq = lambda * lambda;
        const double ksqr = kx[i] * kx[i] + ky[j] * ky[j];
        double chi;
        if (corrected) {
            const double kl = ksqr * lsq;
            chi = kPi * lambda * ksqr * (-defocus + 0.5 * cs3 * kl + (1.0 / 3.0) * cs5 * (kl * kl));
        } else {
            chi = kPi * lambda * ksqr * (0.5 * cs3 * lsq * ksqr - defocus);
        }
        const double trig = chi + 2 * kPi * ((kx[i] * x0) + (ky[j] * y0));
        double sn, cs;
        sincos(trig, &sn, &cs);
        v = double2{cs, -1.0 * sn};
    }
    h[idx] = v;
}

hipError_t launch_probe_k(double2* h, const double* kx, const double* ky, int nx, int ny,
                          const int2* aperture, int corrected, double cs3, double cs5,
                          double defocus, double lambda, double x0, double y0, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny;
    hipLaunchKernelGGL(probe_k_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, h, kx,
                       ky, nx, ny, aperture, corrected, cs3, cs5, defocus, lambda, x0, y0);
    return hipGetLastError();
}

// single workgroup, fixed thread count and tree: deterministic
__global__ void __launch_bounds__(1024) norm2_kernel(const double2* a, int64_t n, double* out)
{
    __shared__ double s[1024];
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += 1024) acc += a[k].x * a[k].x + a[k].y * a[k].y;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

hipError_t launch_norm2(const double2* a, int64_t n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, s, a, n, out);
    return hipGetLastError();
}

__global__ void scale_to_f32_kernel(const double2* a, int64_t n, const double* norm2,
                                    double numerator, float2* out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double f = numerator / sqrt(*norm2);
    out[k] = float2{(float)(a[k].x * f), (float)(a[k].y * f)};
}

hipError_t launch_scale_to_f32(const double2* a, int64_t n, const double* norm2,
                               double numerator, float2* out, hipStream_t s)
{
    hipLaunchKernelGGL(scale_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a,
                       n, norm2, numerator, out);
    return hipGetLastError();
}

__global__ void fill_f64_kernel(double2* a, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = double2{0.0, 0.0};
}

hipError_t launch_fill_f64(double2* a, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

// ------------------------------------------------------------ measurement
// A streaming copy at the rate HBM sustains for a plain read + write: each
// lane moves four 16-byte vectors, the four loads issued before the stores.
// n16 is a multiple of 4 * 256 (the caller sizes the buffers so).
__global__ void __launch_bounds__(256) stream_copy_kernel(const float4* __restrict__ in,
                                                           float4* __restrict__ out)
{
    const int64_t base = (int64_t)blockIdx.x * (4 * 256) + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = in[base + k * 256];
#pragma unroll
    for (int k = 0; k < 4; ++k) out[base + k * 256] = v[k];
}

hipError_t launch_stream_copy(const float4* in, float4* out, int64_t n16, hipStream_t s)
{
    if (n16 <= 0 || n16 % (4 * 256)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)(n16 / (4 * 256))), dim3(256), 0, s, in,
                       out);
    return hipGetLastError();
}

// ------------------------------------------------------------ transmission
// atoms[k] = (slot, species, sx, sy); counts are small integers, so the FP64
// atomic sums are exact and order-independent
__global__ void scatter_atoms_kernel(double2* maps, const int4* atoms, int n_atoms, int nx,
                                     int ny, int nsp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_atoms) return;
    const int4 a = atoms[k];
    double* dst = &maps[(((int64_t)a.x * nsp + a.y) * nx + a.z) * ny + a.w].x;
    atomicAdd(dst, 1.0);
}

hipError_t launch_scatter_atoms(double2* maps, const int4* atoms, int n_atoms, int nx, int ny,
                                int nsp, hipStream_t s)
{
    if (n_atoms <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_atoms_kernel, dim3((n_atoms + 255) / 256), dim3(256), 0, s, maps,
                       atoms, n_atoms, nx, ny, nsp);
    return hipGetLastError();
}

// projected_atomic_potential.cpp:169-225: lambda*gamma/(ab) * f_e(|k|^2) *
// e^{2 pi i k.(xmin, ymin)} inside the template support
__global__ void combine_species_kernel(const double2* maps, double2* sigma, int slots, int nsp,
                                       const double* params, const double* kx, const double* ky,
                                       int nx, int ny, const int2* support, double lgab,
                                       double xmin, double ymin)
{
    const int64_t plane = (int64_t)nx * ny;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane * slots) return;
    const int64_t pix = idx % plane;
    const int64_t slot = idx / plane;
    const int i = (int)(pix / ny), j = (int)(pix % ny);
    const int2 r = support[i];
    double2 acc{0.0, 0.0};
    if (j < r.x || j >= ny - r.y) {
        const double ksq = (kx[i] * kx[i]) + (ky[j] * ky[j]);
        double sn, cs;
        sincos(2 * kPi * (kx[i] * xmin + ky[j] * ymin), &sn, &cs);
        for (int sp = 0; sp < nsp; ++sp) {
            const double* p = params + 12 * sp;
            const double fe = p[0] / (ksq + p[3]) + p[1] / (ksq + p[4]) + p[2] / (ksq + p[5])
                              + p[6] * exp(-p[9] * ksq) + p[7] * exp(-p[10] * ksq)
                              + p[8] * exp(-p[11] * ksq);
            const double2 tk{lgab * fe * cs, lgab * fe * sn};
            const double2 c = maps[(slot * nsp + sp) * plane + pix];
            acc.x += tk.x * c.x - tk.y * c.y;
            acc.y += tk.x * c.y + tk.y * c.x;
        }
    }
    sigma[idx] = acc;
}

hipError_t launch_combine_species(const double2* maps, double2* sigma, int slots, int nsp,
                                  const double* params, const double* kx, const double* ky,
                                  int nx, int ny, const int2* support, double lgab, double xmin,
                                  double ymin, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny * slots;
    hipLaunchKernelGGL(combine_species_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       maps, sigma, slots, nsp, params, kx, ky, nx, ny, support, lgab, xmin, ymin);
    return hipGetLastError();
}

// slice.cpp:1700-1725
__global__ void exp_transmit_kernel(double2* a, int64_t n, double scale)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double re = a[k].x * scale, im = a[k].y * scale;
    const double e = exp(-im);
    double sn, cs;
    sincos(re, &sn, &cs);
    a[k] = double2{e * cs, e * sn};
}

hipError_t launch_exp_transmit(double2* a, int64_t n, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(exp_transmit_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a,
                       n, scale);
    return hipGetLastError();
}

__global__ void apply_keep_kernel(double2* a, int nx, int ny, int64_t total, const int2* keep)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t pix = idx % ((int64_t)nx * ny);
    const int i = (int)(pix / ny), j = (int)(pix % ny);
    const int2 r = keep[i];
    if (!(j < r.x || j >= ny - r.y)) a[idx] = double2{0.0, 0.0};
}

hipError_t launch_apply_keep(double2* a, int nx, int ny, int count, const int2* keep,
                             hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny * count;
    hipLaunchKernelGGL(apply_keep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, nx,
                       ny, n, keep);
    return hipGetLastError();
}

// slice.cpp:1741-1748 (/ NxNy).  The FP32 copy used by the slice loop holds
// t' = t/s (adfstem.cpp:1936,1949 divide by sqrtNxNy) as (t're - 1/s, t'im):
// the deviation from vacuum keeps full FP32 relative precision.
__global__ void finish_transmit_kernel(double2* a, int64_t n, int ny, double nxny, double s,
                                       float2* t32)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double2 t{a[k].x / nxny, a[k].y / nxny};
    a[k] = t;
    const int j = (int)(k % ny);
    t32[k - j + pair_pos(j, ny)] = float2{(float)(t.x / s - 1.0 / s), (float)(t.y / s)};
}

hipError_t launch_finish_transmit(double2* a, int64_t n, int ny, double nxny, double s,
                                  float2* t32, hipStream_t st)
{
    hipLaunchKernelGGL(finish_transmit_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       a, n, ny, nxny, s, t32);
    return hipGetLastError();
}

// vacuum slice step (t = 1) in k-space, slice.cpp:1388-1417 with the FFT
// scale folded into `scale`
__global__ void vacuum_step_kernel(double2* a, int nx, int ny, const double2* px,
                                   const double2* py, const int2* keep, double scale)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nx * ny) return;
    const int i = (int)(idx / ny), j = (int)(idx % ny);
    const int2 r = keep[i];
    double2 v{0.0, 0.0};
    if (j < r.x || j >= ny - r.y) {
        const double pr = px[i].x * py[j].x - px[i].y * py[j].y;
        const double pi = px[i].x * py[j].y + px[i].y * py[j].x;
        const double fr = a[idx].x * scale, fi = a[idx].y * scale;
        v = double2{pr * fr - pi * fi, pr * fi + pi * fr};
    }
    a[idx] = v;
}

hipError_t launch_vacuum_step(double2* a, int nx, int ny, const double2* px, const double2* py,
                              const int2* keep, double scale, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny;
    hipLaunchKernelGGL(vacuum_step_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, nx,
                       ny, px, py, keep, scale);
    return hipGetLastError();
}



// out[j * nx + i] = a[i * ny + j]: the column-major copy a column pass reads
// along its column
__global__ void to_f32_transposed_kernel(const double2* a, int nx, int ny, float2* out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const int64_t j = k / nx;
    const int i = (int)(k % nx);
    const double2 v = a[(int64_t)i * ny + j];
    out[k] = float2{(float)v.x, (float)v.y};
}

hipError_t launch_to_f32_transposed(const double2* a, int nx, int ny, float2* out, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny;
    hipLaunchKernelGGL(to_f32_transposed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       s, a, nx, ny, out);
    return hipGetLastError();
}

__global__ void to_f32_shift_pairs_kernel(const double2* a, int nx, int ny, float2* out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * 2 * ny) return;
    const int64_t i = k / (2 * ny);
    const int m = (int)(k % (2 * ny));
    // pair r = m / 2 holds a[i][r] and a[i][(r + ny/2) mod ny]
    int j = m / 2 + (m % 2) * (ny / 2);
    if (j >= ny) j -= ny;
    const double2 v = a[i * ny + j];
    out[k] = float2{(float)v.x, (float)v.y};
}

hipError_t launch_to_f32_shift_pairs(const double2* a, int nx, int ny, float2* out, hipStream_t s)
{
    const int64_t n = (int64_t)nx * 2 * ny;
    hipLaunchKernelGGL(to_f32_shift_pairs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       s, a, nx, ny, out);
    return hipGetLastError();
}

__global__ void scale_f64_kernel(double2* a, int64_t n, const double* norm2, double f)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double g = f / sqrt(*norm2);
    a[k] = double2{a[k].x * g, a[k].y * g};
}

hipError_t launch_scale_f64(double2* a, int64_t n, const double* norm2, double f, hipStream_t s)
{
    hipLaunchKernelGGL(scale_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n,
                       norm2, f);
    return hipGetLastError();
}

// FP64 transmission update of one slice on `count` real-space waves
// (adfstem.cpp:1922-1952): mode 0 forms the imaginary part from the already
// updated real part, as the reference's two statements do; mode 1 is the
// complex product.  Both divide by s = sqrt(NxNy).
__global__ void transmit_f64_kernel(double2* a, const double2* t, int64_t m, int count, int mode,
                                    double s)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * count) return;
    const double2 tt = t[idx % m];
    const double2 u = a[idx];
    double2 v;
    if (mode == 0) {
        v.x = (tt.x * u.x - tt.y * u.y) / s;
        v.y = (tt.x * u.y + tt.y * v.x) / s;
    } else {
        v.x = (tt.x * u.x - tt.y * u.y) / s;
        v.y = (tt.x * u.y + tt.y * u.x) / s;
    }
    a[idx] = v;
}

hipError_t launch_transmit_f64(double2* a, const double2* t, int64_t m, int count, int mode,
                               double s, hipStream_t st)
{
    const int64_t n = m * count;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(transmit_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                       t, m, count, mode, s);
    return hipGetLastError();
}

} // namespace msf
