// kernels_slice.hip -- launches of the slice-step passes (slice_passes.hpp)
// one batch at a time, fused pass pairs, and the FP64 detector reductions.
//
// Reference behaviour restated (paths relative to the reference root):
//   reductions  = sum_detected_intensity (func.cpp:2593-2641) and
//                 integrate_out_phi_fftw (variance.cpp:64-300), FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include <hip/hip_ext.h>

#include "slice_passes.hpp"

namespace msf {

// ------------------------------------------------------- fused pass pair
// One launch runs a row pass of one batch and a column pass of another
// (independent buffers), so that the VALU-bound row workgroups and the
// memory-bound column workgroups share the CUs instead of following each
// other.  Both passes need workgroups of the same size (RT = W * T).  Launch
// blocks alternate in groups of eight, eight of each pass per sixteen, so
// that a pass's logical block k runs on XCD k % 8 as its XCD banding
// assumes (the row pass's block count, groups * nx, is a multiple of 8).
__device__ __forceinline__ bool fused_decode(int b, int nr, int nc, int& id)
{
    const int m8 = (nr < nc ? nr : nc) & ~7;
    if (b < 2 * m8) {
        const int g = b >> 4, k = b & 15;
        id = g * 8 + (k & 7);
        return k < 8;
    }
    const int r = b - 2 * m8;
    if (r < nr - m8) {
        id = m8 + r;
        return true;
    }
    id = m8 + (r - (nr - m8));
    return false;
}

template <int N, int VR, int VC, bool FIRST, bool LAST, int TMODE, bool PR, bool PC, int W, int RT>
__global__ void __launch_bounds__(RT) fused_pass_kernel(RowPassArgs ra, ColPassArgs ca, int nr, int nc)
{
    constexpr int LR = row_smem_len<N, VR, RT>(), LC = col_smem_len<N, VC, LAST, W>();
    __shared__ __align__(16) float2 smem[LR > LC ? LR : LC];
    int id = 0;
    if (fused_decode((int)blockIdx.x, nr, nc, id))
        row_body<N, VR, FIRST, TMODE, PR, RT>(ra, id, smem);
    else
        col_body<N, VC, LAST, W, PC>(ca, id, nc, smem);
}

// ---------------------------------------------------------------- dispatch
bool slice_lengths_ok(int n)
{
    switch (n) {
    case 16: case 32: case 64: case 96: case 128: case 192: case 256: case 384: case 512:
    case 768: case 1024:
        return true;
    default:
        return false;
    }
}

template <typename P> static void stage_table(std::vector<double>& out)
{
    out.clear();
    int ns = 1;
    for (int s = 0; s < P::kStages; ++s) {
        const int r = P::kRadix[s];
        if (ns > 1)
            for (int q = 1; q < r; ++q)
                for (int k = 0; k < ns; ++k) {
                    const long m = (long)ns * r;
                    const double a = 2.0 * M_PI * (double)(((long)q * k) % m) / (double)m;
                    out.push_back(std::cos(a));
                    out.push_back(-std::sin(a));
                }
        ns *= r;
    }
}

template <int N> static bool stage_table_v(int v, std::vector<double>& out)
{
    if (v == 0) {
        stage_table<Plan<N>>(out);
        return (int)out.size() == 2 * plan_tw_len<Plan<N>>();
    }
    if constexpr (AltPlan<N>::exists) {
        stage_table<AltPlan<N>>(out);
        return (int)out.size() == 2 * plan_tw_len<AltPlan<N>>();
    }
    return false;
}

bool stage_twiddles(int n, std::vector<double>& re_im, int v)
{
    switch (n) {
#define MSF_CASE(N) case N: return stage_table_v<N>(v, re_im);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return false;
    }
}

bool has_alt_plan(int n)
{
    switch (n) {
#define MSF_CASE(N) case N: return AltPlan<N>::exists;
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return false;
    }
}

bool row_prunable(int ny, int ka, int kb)
{
    const int lo = (3 * ny + 7) / 8, hi = (5 * ny) / 8;
    return ka <= lo && ny - kb >= hi && ka >= ny / 4 && kb >= ny / 4;
}

bool col_prunable(int nx, const int32_t* ab)
{
    const int lo = (3 * nx + 7) / 8, hi = (5 * nx) / 8;
    for (int i = lo; i < hi; ++i)
        if (ab[2 * i] != 0 || ab[2 * i + 1] != 0) return false;
    return true;
}

int col_tile(int nx)
{
    switch (nx) {
#define MSF_CASE(N) case N: return ColTile<N>::WMAX;
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return 0;
    }
}

// Launches of so few positions that the default tiles would leave CUs idle
// (one position at 512^2: 44 workgroups of 8 columns for 256 CUs) take one
// column per workgroup instead.
int col_tile_auto(int nx, int v, int batch, int ncp)
{
    const int w = col_tile_default(nx, v);
    constexpr int64_t kCUs = 256;
    return (int64_t)batch * (ncp / std::max(w, 1)) < kCUs ? 1 : w;
}

int col_tile_default(int nx, int v)
{
    switch (nx) {
#define MSF_CASE(N) case N: return v ? ColTile<N, 1>::W : ColTile<N>::W;
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return 0;
    }
}

// With ev0/ev1 the dispatch packet itself stamps the kernel's start and end
// (hipExtLaunchKernelGGL), so profiled durations carry no marker overhead.
template <int NY, int V, int RT>
static hipError_t row_pass_rt(bool first, int tmode, const RowPassArgs& a, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    constexpr int RPW = RT / PlanV<NY, V>::T;
    const int64_t groups = (a.batch + RPW - 1) / RPW;
    const dim3 grid((unsigned)(groups * a.nx));
    const dim3 block(RT);
#define MSF_ROW(F, M, PR)                                                                          \
    do {                                                                                           \
        if (ev0) hipExtLaunchKernelGGL((row_pass_kernel<NY, V, F, M, PR, RT>), grid, block, 0, s,  \
                                       ev0, ev1, 0, a);                                            \
        else hipLaunchKernelGGL((row_pass_kernel<NY, V, F, M, PR, RT>), grid, block, 0, s, a);    \
    } while (0)
    if (a.prune) {
        if (first) { if (tmode == 0) MSF_ROW(true, 0, true); else MSF_ROW(true, 1, true); }
        else { if (tmode == 0) MSF_ROW(false, 0, true); else MSF_ROW(false, 1, true); }
    } else {
        if (first) { if (tmode == 0) MSF_ROW(true, 0, false); else MSF_ROW(true, 1, false); }
        else { if (tmode == 0) MSF_ROW(false, 0, false); else MSF_ROW(false, 1, false); }
    }
#undef MSF_ROW
    return hipGetLastError();
}

// rows per workgroup follow the batch: a batch smaller than a full
// workgroup's rows launches one row per workgroup
template <int NY, int V>
static hipError_t row_pass_nv(bool first, int tmode, const RowPassArgs& a, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    constexpr int RT = row_threads<NY, V>(), T = PlanV<NY, V>::T;
    if constexpr (RT > T) {
        if (a.batch < RT / T) return row_pass_rt<NY, V, T>(first, tmode, a, s, ev0, ev1);
    }
    return row_pass_rt<NY, V, RT>(first, tmode, a, s, ev0, ev1);
}

template <int NY>
static hipError_t row_pass_n(bool first, int tmode, const RowPassArgs& a, hipStream_t s,
                             hipEvent_t ev0, hipEvent_t ev1)
{
    if (a.plan == 0) return row_pass_nv<NY, 0>(first, tmode, a, s, ev0, ev1);
    if constexpr (AltPlan<NY>::exists) {
        if (a.plan == 1) return row_pass_nv<NY, 1>(first, tmode, a, s, ev0, ev1);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_row_pass(int ny, bool first, int tmode, const RowPassArgs& a, hipStream_t s,
                           hipEvent_t ev0, hipEvent_t ev1)
{
    switch (ny) {
#define MSF_CASE(N) case N: return row_pass_n<N>(first, tmode, a, s, ev0, ev1);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <int NX, int V, int W>
static hipError_t col_pass_w(bool last, const ColPassArgs& a, hipStream_t s, hipEvent_t ev0,
                             hipEvent_t ev1)
{
    if (a.ncp % W) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((int64_t)a.batch * ((a.ka + a.kb + W - 1) / W)));
    const dim3 block(W * PlanV<NX, V>::T);
#define MSF_COL(L, PR)                                                                             \
    do {                                                                                           \
        if (ev0) hipExtLaunchKernelGGL((col_pass_kernel<NX, V, L, W, PR>), grid, block, 0, s,      \
                                       ev0, ev1, 0, a);                                            \
        else hipLaunchKernelGGL((col_pass_kernel<NX, V, L, W, PR>), grid, block, 0, s, a);        \
    } while (0)
    if (a.prune) {
        if (last) MSF_COL(true, true);
        else MSF_COL(false, true);
    } else {
        if (last) MSF_COL(true, false);
        else MSF_COL(false, false);
    }
#undef MSF_COL
    return hipGetLastError();
}

template <int NX, int V>
static hipError_t col_pass_nv(bool last, const ColPassArgs& a, hipStream_t s, hipEvent_t ev0,
                              hipEvent_t ev1)
{
    int w = col_tile_auto(NX, V, a.batch, a.ncp);
    if (a.tile == 4 || a.tile == 8 || a.tile == 16) w = a.tile;
    if (w == 1) return col_pass_w<NX, V, 1>(last, a, s, ev0, ev1);
    if constexpr (col_fits<NX, V, 16>()) {
        if (w == 16) return col_pass_w<NX, V, 16>(last, a, s, ev0, ev1);
    }
    if (w == 4) return col_pass_w<NX, V, 4>(last, a, s, ev0, ev1);
    return col_pass_w<NX, V, 8>(last, a, s, ev0, ev1);
}

template <int NX>
static hipError_t col_pass_n(bool last, const ColPassArgs& a, hipStream_t s, hipEvent_t ev0,
                             hipEvent_t ev1)
{
    if (a.plan == 0) return col_pass_nv<NX, 0>(last, a, s, ev0, ev1);
    if constexpr (AltPlan<NX>::exists) {
        if (a.plan == 1) return col_pass_nv<NX, 1>(last, a, s, ev0, ev1);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_col_pass(int nx, bool last, const ColPassArgs& a, hipStream_t s,
                           hipEvent_t ev0, hipEvent_t ev1)
{
    switch (nx) {
#define MSF_CASE(N) case N: return col_pass_n<N>(last, a, s, ev0, ev1);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return hipErrorInvalidValue;
    }
}

// Fused pass pairs exist for the lengths whose default row and column
// workgroups have the same size: 256 (alternate plans: 16 rows x 16 lanes,
// 16 columns x 16 lanes), 384 (8 x 32, 8 x 32) and 512 (8 x 64, 8 x 64),
// with both passes pruned.
template <int N> struct FusedShape { static constexpr bool exists = false; };
template <> struct FusedShape<256> { static constexpr bool exists = true; static constexpr int VR = 1, VC = 1; };
template <> struct FusedShape<384> { static constexpr bool exists = true; static constexpr int VR = 0, VC = 0; };
template <> struct FusedShape<512> { static constexpr bool exists = true; static constexpr int VR = 0, VC = 0; };

template <int N> static bool fused_ok(const RowPassArgs& ra, const ColPassArgs& ca)
{
    if constexpr (!FusedShape<N>::exists) {
        return false;
    } else {
        constexpr int VR = FusedShape<N>::VR, VC = FusedShape<N>::VC;
        constexpr int RT = row_threads<N, VR>(), W = ColTile<N, VC>::W;
        static_assert(RT == W * PlanV<N, VC>::T, "fused workgroups must match");
        const int w = (ca.tile == 4 || ca.tile == 8 || ca.tile == 16) ? ca.tile : W;
        return ra.plan == VR && ca.plan == VC && w == W && ra.prune && ca.prune &&
               ra.batch >= RT / PlanV<N, VR>::T && ca.ncp % W == 0 && ra.nx == N && ca.ny == N;
    }
}

template <int N> static bool fused_shape_n(int vr, int vc, int tile)
{
    if constexpr (!FusedShape<N>::exists) {
        return false;
    } else {
        constexpr int W = ColTile<N, FusedShape<N>::VC>::W;
        return vr == FusedShape<N>::VR && vc == FusedShape<N>::VC && (tile == 0 || tile == W);
    }
}

bool fused_shape_ok(int n, int vr, int vc, int tile)
{
    switch (n) {
    case 256: return fused_shape_n<256>(vr, vc, tile);
    case 384: return fused_shape_n<384>(vr, vc, tile);
    case 512: return fused_shape_n<512>(vr, vc, tile);
    default: return false;
    }
}

bool fused_supported(int n, const RowPassArgs& ra, const ColPassArgs& ca)
{
    switch (n) {
    case 256: return fused_ok<256>(ra, ca);
    case 384: return fused_ok<384>(ra, ca);
    case 512: return fused_ok<512>(ra, ca);
    default: return false;
    }
}

template <int N>
static hipError_t fused_n(bool first, bool last, int tmode, const RowPassArgs& ra,
                          const ColPassArgs& ca, hipStream_t s)
{
    if constexpr (!FusedShape<N>::exists) {
        return hipErrorNotSupported;
    } else {
        if (!fused_ok<N>(ra, ca)) return hipErrorNotSupported;
        constexpr int VR = FusedShape<N>::VR, VC = FusedShape<N>::VC;
        constexpr int RT = row_threads<N, VR>(), W = ColTile<N, VC>::W;
        constexpr int RPW = RT / PlanV<N, VR>::T;
        const int nr = ((ra.batch + RPW - 1) / RPW) * ra.nx;
        const int nc = ca.batch * ((ca.ka + ca.kb + W - 1) / W);
        const dim3 grid((unsigned)(nr + nc)), block(RT);
#define MSF_FUSED(F, L, M)                                                                         \
    hipLaunchKernelGGL((fused_pass_kernel<N, VR, VC, F, L, M, true, true, W, RT>), grid, block, 0,  \
                       s, ra, ca, nr, nc)
        if (first) {
            if (last) { if (tmode == 0) MSF_FUSED(true, true, 0); else MSF_FUSED(true, true, 1); }
            else { if (tmode == 0) MSF_FUSED(true, false, 0); else MSF_FUSED(true, false, 1); }
        } else {
            if (last) { if (tmode == 0) MSF_FUSED(false, true, 0); else MSF_FUSED(false, true, 1); }
            else { if (tmode == 0) MSF_FUSED(false, false, 0); else MSF_FUSED(false, false, 1); }
        }
#undef MSF_FUSED
        return hipGetLastError();
    }
}

hipError_t launch_fused_pass(int n, bool first, bool last, int tmode, const RowPassArgs& ra,
                             const ColPassArgs& ca, hipStream_t s)
{
    switch (n) {
    case 256: return fused_n<256>(first, last, tmode, ra, ca, s);
    case 384: return fused_n<384>(first, last, tmode, ra, ca, s);
    case 512: return fused_n<512>(first, last, tmode, ra, ca, s);
    default: return hipErrorNotSupported;
    }
}

// ------------------------------------------------------------- reductions
// Each lane sums one segment of a ring's pixel list in FP64, in a fixed order;
// the segments of a ring are then added in segment order.  Neither depends on
// the batch composition.
__global__ void __launch_bounds__(256) reduce_segments_kernel(const float2* psi, int64_t plane,
                                                               const int32_t* pix,
                                                               const int32_t* tile_off,
                                                               const int32_t* seg_kind,
                                                               int n_tiles, double* seg_out)
{
    const int lane = threadIdx.x & 63;
    // Tiles follow the groups in ring order, and neighbouring rings read the
    // same cache lines (a FEM ring is narrower than a pixel): give each XCD a
    // contiguous range of (position, tile group) so those reads meet in one L2
    const int gx = (int)gridDim.x;
    const int lin = xcd_band((int)(blockIdx.y * gridDim.x + blockIdx.x), gx * (int)gridDim.y);
    const int b = lin / gx;
    const int w = (lin - b * gx) * 4 + (threadIdx.x >> 6);
    if (w >= n_tiles) return;
    const float2* wave = psi + (int64_t)b * plane;
    const int k0 = tile_off[w], k1 = tile_off[w + 1];
    const bool amp = seg_kind[w * 64 + lane] != 0;
    double acc = 0.0;
    // chunks of 8 pixels: the index loads, then the wave loads (empty slots
    // read element 0), are issued back to back so their latencies overlap;
    // the sum runs over the pixels in the same order as one at a time
    constexpr int C = 8;
    for (int k = k0; k < k1; k += C) {
        int o[C];
#pragma unroll
        for (int j = 0; j < C; ++j) o[j] = (k + j < k1) ? pix[(int64_t)(k + j) * 64 + lane] : -1;
        float2 v[C];
#pragma unroll
        for (int j = 0; j < C; ++j) v[j] = wave[o[j] >= 0 ? o[j] : 0];
#pragma unroll
        for (int j = 0; j < C; ++j) {
            if (o[j] >= 0) {
                const double re = v[j].x, im = v[j].y;
                const double i2 = re * re + im * im;
                acc += amp ? sqrt(i2) : i2;
            }
        }
    }
    seg_out[(int64_t)b * n_tiles * 64 + w * 64 + lane] = acc;
}

hipError_t launch_reduce_segments(const float2* psi, int64_t plane, const int32_t* pix,
                                  const int32_t* tile_off, const int32_t* seg_kind, int n_tiles,
                                  int batch, double* seg_out, hipStream_t s)
{
    if (n_tiles <= 0 || batch <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n_tiles + 3) / 4), (unsigned)batch);
    hipLaunchKernelGGL(reduce_segments_kernel, grid, dim3(256), 0, s, psi, plane, pix, tile_off,
                       seg_kind, n_tiles, seg_out);
    return hipGetLastError();
}

// Group sums from the segment sums, per (group, position), in an order that
// depends on the group only.  Groups of at most kWaveGroup segments: one lane
// adds them in segment order.  Larger groups (the ADF annulus at 1024^2:
// ~5,500 segments): one wavefront, lane l adding segments l, l + 64, ... in
// order, the lanes' sums meeting in a fixed butterfly.
constexpr int kWaveGroup = 64;

__global__ void combine_small_kernel(const double* seg_in, int n_seg_stride,
                                     const int32_t* seg_of_group, const int32_t* ids, int n_ids,
                                     int groups, int batch, double* out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n_ids * batch) return;
    const int g = ids[(int)(idx % n_ids)];
    const int b = (int)(idx / n_ids);
    const double* in = seg_in + (int64_t)b * n_seg_stride;
    double acc = 0.0;
    for (int k = seg_of_group[g]; k < seg_of_group[g + 1]; ++k) acc += in[k];
    out[(int64_t)b * groups + g] = acc;
}

__global__ void __launch_bounds__(256) combine_large_kernel(const double* seg_in, int n_seg_stride,
                                                            const int32_t* seg_of_group,
                                                            const int32_t* ids, int n_ids,
                                                            int groups, int batch, double* out)
{
    const int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= (int64_t)n_ids * batch) return; // whole wavefronts
    const int lane = threadIdx.x & 63;
    const int g = ids[(int)(pair % n_ids)];
    const int b = (int)(pair / n_ids);
    const double* in = seg_in + (int64_t)b * n_seg_stride;
    const int s1 = seg_of_group[g + 1];
    double acc = 0.0;
    for (int k = seg_of_group[g] + lane; k < s1; k += 64) acc += in[k];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) out[(int64_t)b * groups + g] = acc;
}

void combine_split(const std::vector<int32_t>& seg_of_group, std::vector<int32_t>& small,
                   std::vector<int32_t>& large)
{
    small.clear();
    large.clear();
    for (size_t g = 0; g + 1 < seg_of_group.size(); ++g)
        (seg_of_group[g + 1] - seg_of_group[g] > kWaveGroup ? large : small).push_back((int32_t)g);
}

hipError_t launch_combine_segments(const double* seg_in, int n_seg_stride,
                                   const int32_t* seg_of_group, const int32_t* small, int n_small,
                                   const int32_t* large, int n_large, int groups, int batch,
                                   double* out, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    const int64_t ns = (int64_t)n_small * batch, nl = (int64_t)n_large * batch;
    if (ns > 0)
        hipLaunchKernelGGL(combine_small_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                           s, seg_in, n_seg_stride, seg_of_group, small, n_small, groups, batch,
                           out);
    if (nl > 0)
        hipLaunchKernelGGL(combine_large_kernel, dim3((unsigned)((nl + 3) / 4)), dim3(256), 0, s,
                           seg_in, n_seg_stride, seg_of_group, large, n_large, groups, batch, out);
    return hipGetLastError();
}

__global__ void collect_adf_kernel(const double* partial, int groups, int batch, double* adf)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch) adf[b] = partial[(int64_t)b * groups];
}

hipError_t launch_collect_adf(const double* partial, int groups, int batch, double* adf,
                              hipStream_t s)
{
    hipLaunchKernelGGL(collect_adf_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, partial,
                       groups, batch, adf);
    return hipGetLastError();
}

// r = radial sum / count (adfstem.cpp:2591-2599), direct-beam bins zeroed
// first (:2518-2522); Welford over positions in raster order.
__global__ void fem_welford_kernel(const double* partial, int groups, int batch, int nbins,
                                   const int32_t* counts, const uint8_t* zeroed, double* state)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbins) return;
    double n = state[3 * k], mean = state[3 * k + 1], m2 = state[3 * k + 2];
    const double c = (double)counts[k];
    const bool z = zeroed[k] != 0;
    for (int b = 0; b < batch; ++b) {
        const double r = z ? 0.0 : partial[(int64_t)b * groups + 1 + k];
        const double x = r / c;
        n += 1.0;
        const double d = x - mean;
        mean += d / n;
        m2 += d * (x - mean);
    }
    state[3 * k] = n;
    state[3 * k + 1] = mean;
    state[3 * k + 2] = m2;
}

hipError_t launch_fem_welford(const double* partial, int groups, int batch, int nbins,
                              const int32_t* counts, const uint8_t* zeroed, double* state,
                              hipStream_t s)
{
    if (nbins <= 0) return hipSuccess;
    hipLaunchKernelGGL(fem_welford_kernel, dim3((nbins + 255) / 256), dim3(256), 0, s, partial,
                       groups, batch, nbins, counts, zeroed, state);
    return hipGetLastError();
}

// V = Npos sum r^2 / (sum r)^2 - 1 = M2 / (n mean^2); 0 when sum r == 0
// (adfstem.cpp:2837-2856)
__global__ void fem_finalize_kernel(const double* state, int nbins, double* v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbins) return;
    const double n = state[3 * k], mean = state[3 * k + 1], m2 = state[3 * k + 2];
    v[k] = (mean == 0.0) ? 0.0 : m2 / (n * mean * mean);
}

hipError_t launch_fem_finalize(const double* state, int nbins, double* v, hipStream_t s)
{
    if (nbins <= 0) return hipSuccess;
    hipLaunchKernelGGL(fem_finalize_kernel, dim3((nbins + 255) / 256), dim3(256), 0, s, state,
                       nbins, v);
    return hipGetLastError();
}

// Vbar_r (adfstem.cpp:2688-2750): per position and bin
//   ((r2 / c) / ((r1 r1) / (c c))) - 1
// with r1, r2 the ring sums of |psi| and |psi|^2 (zero on the direct beam)
// and c*c an integer product, summed over positions in raster order.
__global__ void fem_vbar_kernel(const double* partial, int groups, int batch, int nbins,
                                int off_abs, int off_sqr, const int32_t* counts,
                                const uint8_t* zeroed, double* vsum)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbins) return;
    const int c = counts[k];
    const double cc = (double)(c * c);
    const bool z = zeroed[k] != 0;
    double acc = vsum[k];
    for (int b = 0; b < batch; ++b) {
        const double r1 = z ? 0.0 : partial[(int64_t)b * groups + off_abs + k];
        const double r2 = z ? 0.0 : partial[(int64_t)b * groups + off_sqr + k];
        acc += ((r2 / (double)c) / ((r1 * r1) / cc)) - 1.0;
    }
    vsum[k] = acc;
}

hipError_t launch_fem_vbar(const double* partial, int groups, int batch, int nbins, int off_abs,
                           int off_sqr, const int32_t* counts, const uint8_t* zeroed,
                           double* vsum, hipStream_t s)
{
    if (nbins <= 0 || batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(fem_vbar_kernel, dim3((nbins + 255) / 256), dim3(256), 0, s, partial,
                       groups, batch, nbins, off_abs, off_sqr, counts, zeroed, vsum);
    return hipGetLastError();
}

// Correlograph (adfstem.cpp:2603-2650): per position, ring ii and phi sector
// jj, c = count_ii * I_ii,jj / r_ii - 1 when that ratio is >= 1 and count,
// I and r are all non-zero, else 0; I and r are the sector and ring sums of
// |psi| (zero on direct-beam rings).  Sums of c and c^2 over positions in
// raster order, stored [ii + jj * nbins].
__global__ void fem_correlograph_kernel(const double* partial, int groups, int batch, int nbins,
                                        int nphi, int off_abs, int off_corr,
                                        const int32_t* counts, const uint8_t* zeroed,
                                        double* csum, double* csqr, double* each)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x; // ii * nphi + jj
    if (k >= nbins * nphi) return;
    const int ii = k / nphi, jj = k - ii * nphi;
    const int64_t o = (int64_t)ii + (int64_t)jj * nbins;
    const int c = counts[ii];
    const bool z = zeroed[ii] != 0;
    double a = csum[o], q = csqr[o];
    for (int b = 0; b < batch; ++b) {
        const double r = z ? 0.0 : partial[(int64_t)b * groups + off_abs + ii];
        const double I = z ? 0.0 : partial[(int64_t)b * groups + off_corr + k];
        const double t = (c * I / r);
        double v = 0.0;
        if ((t >= 1) && (c != 0) && (I != 0) && (r != 0)) v = t - 1;
        a += v;
        q += v * v;
        if (each) each[(int64_t)b * nbins * nphi + o] = v; // per position (optional)
    }
    csum[o] = a;
    csqr[o] = q;
}

hipError_t launch_fem_correlograph(const double* partial, int groups, int batch, int nbins,
                                   int nphi, int off_abs, int off_corr, const int32_t* counts,
                                   const uint8_t* zeroed, double* csum, double* csqr,
                                   double* each, hipStream_t s)
{
    if (nbins <= 0 || nphi <= 0 || batch <= 0) return hipSuccess;
    const int n = nbins * nphi;
    hipLaunchKernelGGL(fem_correlograph_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partial,
                       groups, batch, nbins, nphi, off_abs, off_corr, counts, zeroed, csum, csqr,
                       each);
    return hipGetLastError();
}

// FP64 group sums on full-grid FP64 waves (parity entry points): one
// workgroup per (group, wave), strided pixel loop and a fixed-order tree, so
// results do not depend on scheduling.  kind 0 sums |psi|^2, kind 1 |psi|.
__global__ void __launch_bounds__(256) group_sums_f64_kernel(const double2* psi, int64_t plane,
                                                             const int32_t* offsets,
                                                             const int32_t* pixels,
                                                             const int32_t* kind, int groups,
                                                             double* out)
{
    __shared__ double red[256];
    const int g = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
    const double2* w = psi + (int64_t)b * plane;
    const bool amp = kind[g] == 1;
    double acc = 0.0;
    for (int k = offsets[g] + t; k < offsets[g + 1]; k += 256) {
        const double2 v = w[pixels[k]];
        const double i2 = v.x * v.x + v.y * v.y;
        acc += amp ? sqrt(i2) : i2;
    }
    red[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) out[(int64_t)b * groups + g] = red[0];
}

hipError_t launch_group_sums_f64(const double2* psi, int64_t plane, const int32_t* offsets,
                                 const int32_t* pixels, const int32_t* kind, int groups,
                                 int count, double* out, hipStream_t s)
{
    if (groups <= 0 || count <= 0) return hipSuccess;
    if (groups > 65535 || count > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_sums_f64_kernel, dim3((unsigned)groups, (unsigned)count), dim3(256),
                       0, s, psi, plane, offsets, pixels, kind, groups, out);
    return hipGetLastError();
}

// Detector MTF, first step (adfstem.cpp:2279-2292): |psi| of the compact
// FP32 exit waves outside the aperture (outside[i*ny+j] != 0), else 0, into
// full-grid FP64 planes with zero imaginary part.
__global__ void mtf_magnitude_kernel(const float2* psi, int nx, int ny, int ncp, int ka, int kb,
                                     const uint8_t* outside, int64_t total, double2* out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int j = (int)(idx % ny);
    const int64_t r = idx / ny; // b * nx + i
    const int i = (int)(r % nx);
    double v = 0.0;
    if (outside[(int64_t)i * ny + j] && (j < ka || j >= ny - kb)) {
        const int jc = (j < ka) ? j : j - ny + ka + kb;
        const float2 w = psi[r * ncp + jc];
        const double re = w.x, im = w.y;
        v = sqrt((re * re) + (im * im));
    }
    out[idx] = double2{v, 0.0};
}

hipError_t launch_mtf_magnitude(const float2* psi, int nx, int ny, int ncp, int ka, int kb,
                                const uint8_t* outside, int batch, double2* out, hipStream_t s)
{
    const int64_t n = (int64_t)batch * nx * ny;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mtf_magnitude_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       psi, nx, ny, ncp, ka, kb, outside, n, out);
    return hipGetLastError();
}

// the same on full-grid FP64 waves in place
__global__ void mtf_magnitude_f64_kernel(double2* a, int64_t m, int64_t total,
                                         const uint8_t* outside)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const double2 w = a[idx];
    const double v = outside[idx % m] ? sqrt((w.x * w.x) + (w.y * w.y)) : 0.0;
    a[idx] = double2{v, 0.0};
}

hipError_t launch_mtf_magnitude_f64(double2* a, int64_t m, int count, const uint8_t* outside,
                                    hipStream_t s)
{
    const int64_t n = m * count;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mtf_magnitude_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       s, a, m, n, outside);
    return hipGetLastError();
}

// Detector MTF, last step (adfstem.cpp:2369-2383): / sqrt(NxNy), bw_limit
// (keep ranges per row), back into the compact FP32 planes; the padding
// columns are written as zero.
__global__ void mtf_compact_kernel(const double2* full, int nx, int ny, int ncp, int ka, int kb,
                                   const int2* keep, double s, int64_t total, float2* psi)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int c = (int)(idx % ncp);
    const int64_t r = idx / ncp; // b * nx + i
    const int i = (int)(r % nx);
    float2 v{0.f, 0.f};
    if (c < ka + kb) {
        const int j = (c < ka) ? c : c + ny - ka - kb;
        const int2 k = keep[i];
        if (j < k.x || j >= ny - k.y) {
            const double2 w = full[r * ny + j];
            v = float2{(float)(w.x / s), (float)(w.y / s)};
        }
    }
    psi[idx] = v;
}

hipError_t launch_mtf_compact(const double2* full, int nx, int ny, int ncp, int ka, int kb,
                              const int2* keep, double s, int batch, float2* psi, hipStream_t st)
{
    const int64_t n = (int64_t)batch * nx * ncp;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mtf_compact_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       full, nx, ny, ncp, ka, kb, keep, s, n, psi);
    return hipGetLastError();
}

// a /= d on both parts (the reference divides, it does not scale by 1/d)
__global__ void div_f64_kernel(double2* a, int64_t n, double d)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const double2 w = a[idx];
    a[idx] = double2{w.x / d, w.y / d};
}

hipError_t launch_div_f64(double2* a, int64_t n, double d, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(div_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, d);
    return hipGetLastError();
}

// Per-pixel sums over positions of |psi| and |psi|^2 in raster order, FP64
// (adfstem.cpp:2753-2781), on the compact plane of the exit waves.
__global__ void pixel_moments_kernel(const float2* psi, int64_t plane, int batch, double* s1,
                                     double* s2)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= plane) return;
    double a = s1[p], q = s2[p];
    for (int b = 0; b < batch; ++b) {
        const float2 v = psi[(int64_t)b * plane + p];
        const double re = v.x, im = v.y;
        const double m2 = re * re + im * im;
        a += sqrt(m2);
        q += m2;
    }
    s1[p] = a;
    s2[p] = q;
}

hipError_t launch_pixel_moments(const float2* psi, int64_t plane, int batch, double* s1,
                                double* s2, hipStream_t s)
{
    if (plane <= 0 || batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(pixel_moments_kernel, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0,
                       s, psi, plane, batch, s1, s2);
    return hipGetLastError();
}

__global__ void expand_exit_kernel(const float2* psi, int nx, int ny, int ncp, int ka, int kb,
                                   int batch, float2* out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)batch * nx * ny;
    if (idx >= total) return;
    const int j = (int)(idx % ny);
    const int64_t r = idx / ny; // b * nx + i
    float2 v{0.f, 0.f};
    if (j < ka) v = psi[r * ncp + j];
    else if (j >= ny - kb) v = psi[r * ncp + j - ny + ka + kb];
    out[idx] = v;
}

hipError_t launch_expand_exit(const float2* psi, int nx, int ny, int ncp, int ka, int kb,
                              int batch, float2* out, hipStream_t s)
{
    const int64_t total = (int64_t)batch * nx * ny;
    hipLaunchKernelGGL(expand_exit_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       psi, nx, ny, ncp, ka, kb, batch, out);
    return hipGetLastError();
}

} // namespace msf
