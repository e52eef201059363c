// kernels_group.hip -- grouped launches of the slice passes: one launch runs
// the row (or column) pass over the current batches of up to kGroupMax runs
// of one transform length, each entry with its own arguments (specimen
// slice, vacuum waves, kept columns, shifts, wave buffer).  A workgroup finds
// its entry from the block offsets and runs the same per-workgroup body as
// the single-batch kernels (slice_passes.hpp), so every position's
// arithmetic is that of a run on its own.  No reference counterpart: the
// reference runs one job per process (adfstem.cpp:1556-2793).
#include "slice_passes.hpp"

namespace msf {

// entry of launch block b: the last k with off[k] <= b (workgroup-uniform,
// read from the kernel arguments)
__device__ __forceinline__ int group_entry(const int* off, int n, int b)
{
    int k = 0;
    for (int e = 1; e < n; ++e) k = (b >= off[e]) ? e : k;
    return k;
}

template <int NY, int V, bool FIRST, int TMODE, bool PRUNE, int RT>
__global__ void __launch_bounds__(RT)
__attribute__((amdgpu_waves_per_eu((row_waves<NY, V>())))) row_group_kernel(RowGroupArgs g)
{
    __shared__ __align__(16) float2 smem[row_smem_len<NY, V, RT>()];
    constexpr int RPW = RT / PlanV<NY, V>::T;
    const int k = group_entry(g.off, g.n, (int)blockIdx.x);
    const RowPassArgs& a = g.a[k];
    const int blk = (int)blockIdx.x - g.off[k];
    if (blk >= ((a.batch + RPW - 1) / RPW) * a.nx) return; // padding block
    row_body<NY, V, FIRST, TMODE, PRUNE, RT>(a, blk, smem);
}

template <int NX, int V, bool LAST, int W, bool PRUNE>
__global__ void __launch_bounds__((col_bound<NX, V, W>())) col_group_kernel(ColGroupArgs g)
{
    __shared__ float2 smem[col_smem_len<NX, V, LAST, W>()];
    const int k = group_entry(g.off, g.n, (int)blockIdx.x);
    const ColPassArgs& a = g.a[k];
    const int blk = (int)blockIdx.x - g.off[k];
    const int nblk = a.batch * ((a.ka + a.kb + W - 1) / W);
    if (blk >= nblk) return; // padding block
    col_body<NX, V, LAST, W, PRUNE>(a, blk, nblk, smem);
}

// ------------------------------------------------------------------ launches
template <int NY, int V> static constexpr int group_rpw()
{
    return row_threads<NY, V>() / PlanV<NY, V>::T;
}

template <int NY> static int row_blocks_n(const RowPassArgs& a)
{
    const int rpw = a.plan ? group_rpw<NY, AltPlan<NY>::exists ? 1 : 0>() : group_rpw<NY, 0>();
    return ((a.batch + rpw - 1) / rpw) * a.nx;
}

int row_group_blocks(int ny, const RowPassArgs& a)
{
    switch (ny) {
#define MSF_CASE(N) case N: return row_blocks_n<N>(a);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return 0;
    }
}

int col_group_tile(int nx, int v, int tile)
{
    return (tile == 4 || tile == 8 || tile == 16) ? tile : col_tile_default(nx, v);
}

int col_group_blocks(int nx, const ColPassArgs& a)
{
    const int w = col_group_tile(nx, a.plan, a.tile);
    return w > 0 ? a.batch * ((a.ka + a.kb + w - 1) / w) : 0;
}

template <int NY, int V>
static hipError_t row_group_nv(bool first, int tmode, const RowGroupArgs& g, hipStream_t s)
{
    constexpr int RT = row_threads<NY, V>();
    const dim3 grid((unsigned)g.off[g.n]), block(RT);
    const bool pr = g.a[0].prune != 0;
#define MSF_ROWG(F, M, PR)                                                                         \
    hipLaunchKernelGGL((row_group_kernel<NY, V, F, M, PR, RT>), grid, block, 0, s, g)
    if (pr) {
        if (first) { if (tmode == 0) MSF_ROWG(true, 0, true); else MSF_ROWG(true, 1, true); }
        else { if (tmode == 0) MSF_ROWG(false, 0, true); else MSF_ROWG(false, 1, true); }
    } else {
        if (first) { if (tmode == 0) MSF_ROWG(true, 0, false); else MSF_ROWG(true, 1, false); }
        else { if (tmode == 0) MSF_ROWG(false, 0, false); else MSF_ROWG(false, 1, false); }
    }
#undef MSF_ROWG
    return hipGetLastError();
}

template <int NY>
static hipError_t row_group_n(bool first, int tmode, const RowGroupArgs& g, hipStream_t s)
{
    if (g.a[0].plan == 0) return row_group_nv<NY, 0>(first, tmode, g, s);
    if constexpr (AltPlan<NY>::exists) {
        if (g.a[0].plan == 1) return row_group_nv<NY, 1>(first, tmode, g, s);
    }
    return hipErrorInvalidValue;
}

// the arguments every entry must share, and the block offsets
template <class Args, class Blocks>
static bool group_ok(int n_len, int n, const int* off, const Args* a, Blocks blocks)
{
    if (n < 1 || n > kGroupMax || off[0] != 0) return false;
    for (int k = 0; k < n; ++k) {
        if (a[k].plan != a[0].plan || a[k].prune != a[0].prune || a[k].batch < 1) return false;
        if (off[k] % 8 || off[k + 1] - off[k] < blocks(n_len, a[k])) return false;
    }
    return true;
}

hipError_t launch_row_group(int ny, bool first, int tmode, const RowGroupArgs& g, hipStream_t s)
{
    if (!group_ok(ny, g.n, g.off, g.a, row_group_blocks)) return hipErrorInvalidValue;
    for (int k = 0; k < g.n; ++k)
        if (g.a[k].nx != g.a[0].nx) return hipErrorInvalidValue;
    switch (ny) {
#define MSF_CASE(N) case N: return row_group_n<N>(first, tmode, g, s);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <int NX, int V, int W>
static hipError_t col_group_w(bool last, const ColGroupArgs& g, hipStream_t s)
{
    for (int k = 0; k < g.n; ++k)
        if (g.a[k].ncp % W) return hipErrorInvalidValue;
    const dim3 grid((unsigned)g.off[g.n]), block(W * PlanV<NX, V>::T);
    const bool pr = g.a[0].prune != 0;
#define MSF_COLG(L, PR)                                                                            \
    hipLaunchKernelGGL((col_group_kernel<NX, V, L, W, PR>), grid, block, 0, s, g)
    if (pr) {
        if (last) MSF_COLG(true, true);
        else MSF_COLG(false, true);
    } else {
        if (last) MSF_COLG(true, false);
        else MSF_COLG(false, false);
    }
#undef MSF_COLG
    return hipGetLastError();
}

template <int NX, int V>
static hipError_t col_group_nv(bool last, const ColGroupArgs& g, hipStream_t s)
{
    const int w = col_group_tile(NX, V, g.a[0].tile);
    for (int k = 0; k < g.n; ++k)
        if (col_group_tile(NX, V, g.a[k].tile) != w) return hipErrorInvalidValue;
    if constexpr (col_fits<NX, V, 16>()) {
        if (w == 16) return col_group_w<NX, V, 16>(last, g, s);
    }
    if (w == 4) return col_group_w<NX, V, 4>(last, g, s);
    if (w == 8) return col_group_w<NX, V, 8>(last, g, s);
    return hipErrorInvalidValue;
}

template <int NX>
static hipError_t col_group_n(bool last, const ColGroupArgs& g, hipStream_t s)
{
    if (g.a[0].plan == 0) return col_group_nv<NX, 0>(last, g, s);
    if constexpr (AltPlan<NX>::exists) {
        if (g.a[0].plan == 1) return col_group_nv<NX, 1>(last, g, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_col_group(int nx, bool last, const ColGroupArgs& g, hipStream_t s)
{
    if (!group_ok(nx, g.n, g.off, g.a, col_group_blocks)) return hipErrorInvalidValue;
    for (int k = 0; k < g.n; ++k)
        if (g.a[k].ny != g.a[0].ny) return hipErrorInvalidValue;
    switch (nx) {
#define MSF_CASE(N) case N: return col_group_n<N>(last, g, s);
        MSF_CASE(16) MSF_CASE(32) MSF_CASE(64) MSF_CASE(96) MSF_CASE(128) MSF_CASE(192)
        MSF_CASE(256) MSF_CASE(384) MSF_CASE(512) MSF_CASE(768) MSF_CASE(1024)
#undef MSF_CASE
    default:
        return hipErrorInvalidValue;
    }
}

} // namespace msf
