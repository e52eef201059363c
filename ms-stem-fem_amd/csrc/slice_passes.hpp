// slice_passes.hpp -- device code of the FP32 slice-step passes: the row
// pass (inverse transform along y, transmission, forward transform along y)
// and the column pass (forward transform along x, band limit and propagator,
// inverse transform along x, or the exit wave on the last slice), as device
// functions over one workgroup's share and as kernels.  Included by the units
// that launch them one batch at a time (kernels_slice.hip) and grouped over
// several runs (kernels_group.hip).
//
// Reference behaviour restated (paths relative to the reference root):
//   row pass    = adfstem.cpp:1891 (BACKWARD FFT, the y half) + :1922-1952
//                 (transmission multiply, /sqrtNxNy) + :1986 (FORWARD FFT,
//                 y half); the first pass starts from the circularly shifted
//                 real-space probe of adfstem.cpp:1632-1671.
//   column pass = adfstem.cpp:1986-1992 (FORWARD FFT, x half, /sqrtNxNy) +
//                 bw_limit (bw_limit.cpp:30-47) + slice::propagate
//                 (slice.cpp:1366-1421) + the next slice's BACKWARD FFT (x half).
#pragma once

#include <hip/hip_runtime.h>

#include "fft_engine.hpp"
#include "kernels.hpp"

namespace msf {

// Element k of an array whose byte offsets fit 32 bits (a row or one
// position's plane of the wave): the offset is formed in 32 bits, so the
// load or store takes a uniform base and a 32-bit per-lane offset rather
// than a 64-bit address per element.
template <typename T> __device__ __forceinline__ T& elem32(T* base, uint32_t k)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + k * (uint32_t)sizeof(T));
}

// ------------------------------------------------------------------ row pass
// Transmission of one row.  With t' = t/s = (tau + 1/s, ti) and the
// reference's in-place update (TMODE 0: Im uses the updated Re), for the
// scattered wave u and the vacuum wave u0:
//   a  = t're u_re - ti u_im,            b  = t're u_im + ti a
//   Da = tau u0_re - ti u0_im,           Db = tau u0_im + ti (u0_re/s + Da)
// and TMODE 1 (exact product) uses b = t're u_im + ti u_re, Db = tau u0_im + ti u0_re.
// Lane element p (indices t + T*p) lies entirely inside the middle band
// [3N/8, 5N/8), which the 2/3 band limit leaves empty in k-space (checked on
// the host before a pruned variant is launched).
template <int N, int E> __host__ __device__ constexpr bool in_zero_band(int p)
{
    constexpr int T = N / E;
    return T * p >= (3 * N + 7) / 8 && T * (p + 1) <= (5 * N) / 8;
}

// With PRUNE the host also guarantees ka >= N/4 and kb >= N/4: lane elements
// entirely below N/4 are always in the low kept band, those entirely at or
// above 3N/4 always in the high band; only the elements straddling a band edge
// need a run-time test.
template <int N, int E> __host__ __device__ constexpr bool always_low(int p)
{
    return (N / E) * (p + 1) <= N / 4;
}
template <int N, int E> __host__ __device__ constexpr bool always_high(int p)
{
    return (N / E) * p >= N - N / 4;
}

template <int N> __device__ __forceinline__ int wrap_sub(int j, int s)
{
    if constexpr ((N & (N - 1)) == 0) {
        return (j - s) & (N - 1);
    } else {
        const int d = j - s;
        return d < 0 ? d + N : d;
    }
}

// Workgroup -> (row i, group of RPW positions).  Consecutive launch indices go
// to different XCDs round-robin; the permutation gives each XCD a contiguous
// band of rows so that the t and u0 rows a band needs are fetched into that
// XCD's L2 once.
__device__ __forceinline__ int xcd_band(int L, int G)
{
    // Workgroups are dispatched round-robin over the 8 XCDs (L % 8).  Give
    // XCD x the contiguous logical range [x*q + min(x, r), ... + q + (x < r))
    // with G = 8q + r, so neighbouring tiles (which share cache lines) meet
    // in one L2.  A bijection on [0, G) for every G.
    constexpr int kXcd = 8;
    const int q = G / kXcd, r = G % kXcd;
    const int x = L % kXcd, k = L / kXcd;
    return x * q + min(x, r) + k;
}

// lanes of a row-pass workgroup: RPW = row_threads / T rows share the staged
// twiddle table and transmission row.  At 512 (primary plan) eight rows per
// workgroup halve that staging per row (row pass 41.1 -> 40.4 us per 64
// positions); at 1024 the larger LDS footprint costs more than it saves.
template <int NY, int V> __host__ __device__ constexpr int row_threads()
{
    // the alternate 1024 plan (16 values on 64 lanes, ~120 registers): two
    // rows per 128-lane workgroup, 71 us against 76 us on 256 lanes per 17
    // positions
    return (NY == 512 && V == 0) ? 512 : (NY == 1024 && V == 1) ? 128 : 256;
}

// Lean row passes read the twiddle table and the transmission row from
// global memory (L2-resident, shared by the workgroups of a row band) instead
// of staging them in LDS, so that more workgroups fit on a CU.
template <int NY, int V> __host__ __device__ constexpr bool row_lean() { return NY == 1024; }

// Occupancy the row pass asks of the register allocator: 6 waves per SIMD for
// the 8-value plans; the plans with 12 to 24 values per lane need more
// registers, and ask for what they reach (their allocation is unchanged).
template <int NY, int V> __host__ __device__ constexpr int row_waves()
{
    constexpr int E = PlanV<NY, V>::E;
    return E <= 8 ? 6 : (E == 16 ? (NY == 512 ? 3 : 4) : (E == 12 ? 3 : 2));
}

// RT lanes per workgroup: row_threads() for batches that fill it, or one
// row's T lanes when the batch has fewer positions than a workgroup has rows
// (no lane then computes a row it does not store).
// LDS of a row-pass workgroup in float2 units: the staged twiddle table and
// transmission row (unless lean), and one exchange buffer per row.
template <int NY, int V, int RT> __host__ __device__ constexpr bool row_is_lean()
{
    // one row per workgroup: nothing to share, tables are read from L2
    return row_lean<NY, V>() || RT / PlanV<NY, V>::T == 1;
}
// padding of a row's exchange buffer: one slot every 2^row_pad_shift values.
// The alternate 1024 row pass pads every 32 (8.3 KB per row instead of
// 9.2 KB): its pass takes as long, but the smaller workgroups leave room for
// the other batch chain's column pass on the same CUs (1024^2 job +3 %).
template <int NY, int V> __host__ __device__ constexpr int row_pad_shift()
{
    return (NY == 1024 && V == 1) ? 5 : 3;
}
// Row passes that read every transmission and u0 pair of a lane before the
// first product (the alternate 1024 plan: row pass 57.5 -> 53.4 us; the
// primary 512 plan: the 512^2 job +0.7 %); the others interleave reads and
// products, which keeps their registers lower (384: 128 against 149, the
// alternate 256 plan: 101 against 128 with spills; the uneven 256-512 job
// set +1 %).
template <int NY, int V> __host__ __device__ constexpr bool row_reads_ahead()
{
    return (NY == 1024 && V == 1) || (NY == 512 && V == 0);
}
template <int NY, int V, int RT> __host__ __device__ constexpr int row_smem_len()
{
    return (row_is_lean<NY, V, RT>() ? 0 : plan_tw_len<PlanV<NY, V>>() + NY) +
           (RT / PlanV<NY, V>::T) * lds_len<row_pad_shift<NY, V>()>(NY);
}

// One row-pass workgroup: logical block `blk` of the pass (XCD-banded over
// the pass's own blocks), LDS at smem.
template <int NY, int V, bool FIRST, int TMODE, bool PRUNE, int RT>
__device__ __forceinline__ void row_body(const RowPassArgs& a, int blk, float2* smem)
{
    using P = PlanV<NY, V>;
    constexpr int E = P::E, T = P::T;
    constexpr int RPW = RT / T;
    constexpr int PS = row_pad_shift<NY, V>();
    constexpr int LEN = lds_len<PS>(NY);
    constexpr bool LEAN = row_is_lean<NY, V, RT>();
    float2* s_tw = smem;
    float2* s_t = smem + (LEAN ? 0 : plan_tw_len<P>());
    float2* s_buf = smem + (LEAN ? 0 : plan_tw_len<P>() + NY);

    // workgroup-uniform decode (scalar unit)
    const int groups = (a.batch + RPW - 1) / RPW;
    const int g = xcd_band(blk, groups * a.nx);
    const int i = g / groups;
    const int b0 = (g - i * groups) * RPW;

    // twiddles and the transmission row i, shared by the RPW rows of the group
    const float2* trow = a.t + (int64_t)i * NY;
    if constexpr (!LEAN) {
        for (int e = threadIdx.x; e < NY; e += blockDim.x) {
            if (e < plan_tw_len<P>()) s_tw[e] = a.tw[e];
            s_t[e] = trow[e];
        }
        __syncthreads();
    }
    const float2* tw = LEAN ? a.tw : s_tw;
    const float2* tt = LEAN ? trow : s_t;

    // with T == 64 each row is one wavefront: keep its index in a scalar register
    const int lr = (T == 64) ? __builtin_amdgcn_readfirstlane(threadIdx.x / T) : threadIdx.x / T;
    const int t = threadIdx.x % T;
    const bool active = b0 + lr < a.batch;
    const int b = active ? b0 + lr : a.batch - 1;
    float2* row = a.psi + ((int64_t)b * a.nx + i) * a.ncp;
    float2* lds = s_buf + lr * LEN;
    const int hi = NY - a.kb;           // first kept column of the negative-ky band
    const int off = a.ka + a.kb - NY;   // compact index = j + off for j >= hi

    // Transmission and u0 rows are stored as pairs (pair_pos): a lane's
    // elements p and p + E/2 are j and j + NY/2, and one 16-byte read fetches
    // both, of t at j and of u0 at the shifted r = (j - sy) mod NY.
    const int2 sh = a.shift[b];
    int ii = i - sh.x;
    if (ii < 0) ii += a.nx;
    const float2* urow = a.u0 + (int64_t)ii * (2 * NY);
    auto pair_at = [&](int p, float4& tv, float4& uv) {
        const int j = t + T * p;
        int r = j - sh.y;
        if (r < 0) r += NY;
        tv = *reinterpret_cast<const float4*>(tt + 2 * j);
        uv = *reinterpret_cast<const float4*>(urow + 2 * r);
    };

    float2 x[E];
    if constexpr (!FIRST) {
#pragma unroll
        for (int p = 0; p < E; ++p) {
            const uint32_t j = (uint32_t)(t + T * p);
            if (PRUNE && in_zero_band<NY, E>(p)) {
                x[p] = float2{0.f, 0.f};
            } else if (PRUNE && always_low<NY, E>(p)) {
                x[p] = elem32(row, j);
            } else if (PRUNE && always_high<NY, E>(p)) {
                x[p] = elem32(row, j + (uint32_t)off);
            } else {
                const bool lo = (int)j < a.ka, up = (int)j >= hi;
                // lanes outside the kept band all read element 0 (one shared line)
                const float2 v = elem32(row, lo ? j : (up ? j + (uint32_t)off : 0u));
                x[p] = (lo || up) ? v : float2{0.f, 0.f};
            }
        }
        P::template run<true, float2, (T <= 64), PS>(x, lds, tw, t);
    }

    const float is = a.inv_s;
    auto transmit = [&](int p, float2 tv, float2 u0) {
        const float da = tv.x * u0.x - tv.y * u0.y;
        const float db = (TMODE == 0) ? tv.x * u0.y + tv.y * (u0.x * is + da) : tv.x * u0.y + tv.y * u0.x;
        if constexpr (FIRST) {
            x[p] = float2{da, db};
        } else {
            const float tr = tv.x + is;
            const float2 u = x[p];
            const float re = tr * u.x - tv.y * u.y;
            const float im = (TMODE == 0) ? tr * u.y + tv.y * re : tr * u.y + tv.y * u.x;
            x[p] = float2{re + da, im + db};
        }
    };
    if constexpr (row_reads_ahead<NY, V>()) {
        // all reads issued before the first product: more of them in flight,
        // for more registers
        float4 tq[E / 2], uq[E / 2];
#pragma unroll
        for (int p = 0; p < E / 2; ++p) pair_at(p, tq[p], uq[p]);
#pragma unroll
        for (int p = 0; p < E / 2; ++p) {
            transmit(p, float2{tq[p].x, tq[p].y}, float2{uq[p].x, uq[p].y});
            transmit(p + E / 2, float2{tq[p].z, tq[p].w}, float2{uq[p].z, uq[p].w});
        }
    } else {
#pragma unroll
        for (int p = 0; p < E / 2; ++p) {
            float4 tv, uv;
            pair_at(p, tv, uv);
            transmit(p, float2{tv.x, tv.y}, float2{uv.x, uv.y});
            transmit(p + E / 2, float2{tv.z, tv.w}, float2{uv.z, uv.w});
        }
    }

    P::template run<false, float2, (T <= 64), PS>(x, lds, tw, t);

    if (active) {
#pragma unroll
        for (int p = 0; p < E; ++p) {
            const uint32_t j = (uint32_t)(t + T * p);
            if (PRUNE && in_zero_band<NY, E>(p)) continue;
            if (PRUNE && always_low<NY, E>(p)) {
                elem32(row, j) = x[p];
            } else if (PRUNE && always_high<NY, E>(p)) {
                elem32(row, j + (uint32_t)off) = x[p];
            } else {
                if ((int)j < a.ka) elem32(row, j) = x[p];
                else if ((int)j >= hi) elem32(row, j + (uint32_t)off) = x[p];
            }
        }
    }
}

template <int NY, int V, bool FIRST, int TMODE, bool PRUNE, int RT>
__global__ void __launch_bounds__(RT)
__attribute__((amdgpu_waves_per_eu((row_waves<NY, V>())))) row_pass_kernel(RowPassArgs a)
{
    __shared__ __align__(16) float2 smem[row_smem_len<NY, V, RT>()];
    row_body<NY, V, FIRST, TMODE, PRUNE, RT>(a, (int)blockIdx.x, smem);
}

// --------------------------------------------------------------- column pass
// A workgroup owns W adjacent kept columns of one position.  The tile is
// loaded with lanes running along the columns (coalesced W*8-byte segments
// per row) into a column-major LDS tile, each column is then transformed by
// T lanes of one wavefront (wave-local exchanges, no workgroup barrier
// between radix stages), and the tile is written back the same coalesced way.
template <int NX, int V = 0> struct ColTile {
    static constexpr int T = PlanV<NX, V>::T;
    static constexpr int WMAX = 16;  // ncp is a multiple of this
    // default tile: keep LDS (W * lds_len * 8 B) near 40 KiB so that several
    // workgroups share a CU
    // The alternate 1024 plan (16 values on 64 lanes, ~190 registers) fits
    // one 8-column tile per CU; 4-column tiles with lean tables fit two
    // (column pass 83 -> 74.5 us per 17 positions).
    static constexpr int W = (NX == 1024 && V == 1) ? 4 : (16 * lds_len(NX) * 8 > 48 * 1024) ? 8 : 16;
};

// padding of a column in the tile: one slot every 2^col_pad_shift values.
// The alternate 1024 column pass pads every 16 (a 4-column tile of 34.9 KB
// instead of 37 KB): the pass 45.5 -> 44.0 us and the 1024^2 job +3-4 %
// (every 32: +2 %), the smaller tiles sharing CUs with the other batch
// chain's row pass; the primary 512 column pass likewise (the 512^2 job
// +1.3 %; every 32: no better), and the alternate 256 one (its pass
// 40.3-42.4 -> 39.0-39.5 us per 256 positions).
template <int NX, int V> __host__ __device__ constexpr int col_pad_shift()
{
    return ((NX == 1024 && V == 1) || (NX == 512 && V == 0) || (NX == 256 && V == 1)) ? 4 : 3;
}
// column stride of the LDS tile: a 16-lane group of the coalesced load/store
// phases covers W' = min(W, 16) columns x 16/W' rows; a stride congruent to
// 16/W' modulo 16 (8-byte words) puts those 16 lanes on distinct bank pairs
template <int NX, int V, int W> __host__ __device__ constexpr int col_len()
{
    constexpr int base = lds_len<col_pad_shift<NX, V>()>(NX);
    constexpr int want = (16 / (W < 16 ? W : 16)) % 16;
    return base + ((want - base % 16) + 16) % 16;
}

// Lean tiles: when the twiddle and propagator tables would keep two
// workgroups from sharing a CU's 160 KB of LDS (1024 rows: a 74 KB tile),
// the tables are read from global memory (L2-resident, shared by every
// workgroup) and only the tile stays in LDS.
template <int NX, int V, int W> __host__ __device__ constexpr bool col_lean()
{
    if (W == 1) return true;      // one column: nothing to share, tables are read from L2
    if (W == 4) return NX >= 512; // small tiles: tables would outweigh the tile
    return 2 * ((size_t)W * col_len<NX, V, W>() * 8 + (size_t)plan_tw_len<PlanV<NX, V>>() * 8 +
                (size_t)NX * 8) > 160 * 1024;
}

// Plans with 12 or more values per lane (the alternate plans, and the
// primary plans of 96, 192, 384 and 768) take the tile's real size as their
// launch bound, so that a 128- to 512-lane tile may use more than the 128
// registers a 1024-lane bound leaves per lane (384: 116 B of spills per lane
// under that bound).
template <int NX, int V, int W> __host__ __device__ constexpr int col_bound()
{
    return (V || PlanV<NX, V>::E >= 12) ? W * PlanV<NX, V>::T : 1024;
}

// LDS of a column-pass workgroup in float2 units: the staged twiddle and
// propagator tables (unless lean), the exit roots (last slice) and the tile.
template <int NX, int V, bool LAST, int W> __host__ __device__ constexpr int col_smem_len()
{
    return (col_lean<NX, V, W>() ? 0 : plan_tw_len<PlanV<NX, V>>() + NX) + (LAST ? NX : 0) +
           W * col_len<NX, V, W>();
}

// tile sizes whose LDS and workgroup fit a CU
template <int NX, int V, int W> __host__ __device__ constexpr bool col_fits()
{
    using P = PlanV<NX, V>;
    const size_t tables = col_lean<NX, V, W>() ? (size_t)NX * 8  // s_nat of the last slice
                                                : (size_t)NX * 16 + (size_t)plan_tw_len<P>() * 8;
    return (size_t)W * col_len<NX, V, W>() * 8 + tables <= 160 * 1024 && W * P::T <= 1024;
}

// One column-pass workgroup: logical block `blk` of `nblk` (XCD-banded over
// the pass's own blocks), LDS at smem.
template <int NX, int V, bool LAST, int W, bool PRUNE>
__device__ __forceinline__ void col_body(const ColPassArgs& a, int blk, int nblk, float2* smem)
{
    using P = PlanV<NX, V>;
    constexpr int E = P::E, T = P::T;
    constexpr int LEN = col_len<NX, V, W>();
    constexpr int PS = col_pad_shift<NX, V>();
    constexpr bool LEAN = col_lean<NX, V, W>();
    // a column transform lies in one wavefront (wave-local exchanges) or
    // spans whole wavefronts (workgroup barriers)
    static_assert(64 % T == 0 || T % 64 == 0, "a column transform must align with wavefronts");
    constexpr bool WAVE = T <= 64;
    float2* s_tw = smem;
    float2* s_px = smem + (LEAN ? 0 : plan_tw_len<P>());
    float2* s_nat = smem + (LEAN ? 0 : plan_tw_len<P>() + NX); // roots W_nx^e (last slice)
    float2* s_buf = s_nat + (LAST ? NX : 0);

    for (int e = threadIdx.x; e < NX; e += blockDim.x) {
        if constexpr (!LEAN) {
            if (e < plan_tw_len<P>()) s_tw[e] = a.tw[e];
            s_px[e] = a.px[e];
        }
        if constexpr (LAST) s_nat[e] = a.twx_nat[e];
    }
    const float2* tw = LEAN ? a.tw : s_tw;
    const float2* tpx = LEAN ? a.px : s_px;

    // neighbouring tiles share 128-byte lines (W = 8 reads 64 B per row): keep
    // them on one XCD, next to each other in dispatch order
    // tiles cover the stored columns [0, nc); tiles of padding columns only
    // (ncp is rounded up to 16) are not launched
    const int nc = a.ka + a.kb;
    const int ntiles = (nc + W - 1) / W;
    const int g = xcd_band(blk, nblk);
    const int b = g / ntiles;
    const int j0 = (g - b * ntiles) * W;
    float2* tile = a.psi + (int64_t)b * NX * a.ncp + j0;

    {   // coalesced load of the tile (32-bit offsets from the uniform tile base).
        // Every load is in bounds (b < batch by the grid, padding columns
        // jc >= nc exist in the buffer), so all E loads are issued before the
        // first LDS write and overlap; padding columns are zeroed after.
        const int c = threadIdx.x % W, r0 = threadIdx.x / W;
        const bool ok = j0 + c < nc;
        const uint32_t o0 = (uint32_t)(r0 * a.ncp + c);
        const uint32_t step = (uint32_t)(T * a.ncp);
        float2 v[E];
#pragma unroll
        for (int p = 0; p < E; ++p) v[p] = elem32(tile, o0 + (uint32_t)p * step);
#pragma unroll
        for (int p = 0; p < E; ++p)
            s_buf[c * LEN + lds_pad_at<T, PS>(r0, p)] = ok ? v[p] : float2{0.f, 0.f};
    }
    __syncthreads();

    const int c = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int jc = j0 + c;
    const bool active = (b < a.batch) && (jc < nc);
    const int j = (jc < a.ka) ? jc : jc - nc + a.ny;
    float2* lds = s_buf + c * LEN;
    float2 x[E];
#pragma unroll
    for (int p = 0; p < E; ++p) x[p] = lds[lds_pad_at<T, PS>(t, p)];

    P::template run<false, float2, WAVE, PS>(x, lds, tw, t);

    // band limit (bw_limit + propagate's k^2 test, decided on the host) and
    // propagator P_x[i] * P_y[j]; P_y carries the 1/sqrtNxNy of the FFT
    const float2 pyj = active ? a.py[j] : float2{0.f, 0.f};
    // kept rows of this column, decided on the host (one pair per column
    // instead of a range table per row)
    const int2 ck = active ? a.colkeep[jc] : int2{0, 0};
#pragma unroll
    for (int p = 0; p < E; ++p) {
        if (PRUNE && in_zero_band<NX, E>(p)) {
            x[p] = float2{0.f, 0.f};
            continue;
        }
        const int i = t + T * p;
        const bool keep = (i < ck.x) || (i >= NX - ck.y);
        const float2 pxi = tpx[i];
        const float2 pr = float2{pxi.x * pyj.x - pxi.y * pyj.y, pxi.x * pyj.y + pxi.y * pyj.x};
        const float2 f = x[p];
        x[p] = keep ? float2{pr.x * f.x - pr.y * f.y, pr.x * f.y + pr.y * f.x} : float2{0.f, 0.f};
    }

    if constexpr (!LAST) {
        P::template run<true, float2, WAVE, PS>(x, lds, tw, t);
    } else if (active) {
        // add the vacuum exit wave, shifted: psi0(k) e^{-2 pi i (i sx / nx + j sy / ny)}
        const int2 sh = a.shift[b];
        const float2 wy = a.twy[((uint32_t)j * (uint32_t)sh.y) % (uint32_t)a.ny];
#pragma unroll
        for (int p = 0; p < E; ++p) {
            if (PRUNE && in_zero_band<NX, E>(p)) continue;
            const uint32_t i = (uint32_t)(t + T * p);
            const float2 wx = s_nat[(i * (uint32_t)sh.x) % (uint32_t)NX];
            const float2 w = float2{wx.x * wy.x - wx.y * wy.y, wx.x * wy.y + wx.y * wy.x};
            const float2 v0 = a.psi0[(int64_t)j * NX + i];
            x[p].x += v0.x * w.x - v0.y * w.y;
            x[p].y += v0.x * w.y + v0.y * w.x;
        }
    }

    lds_sync<WAVE>();
#pragma unroll
    for (int p = 0; p < E; ++p) lds[lds_pad_at<T, PS>(t, p)] = x[p];
    __syncthreads();

    {   // coalesced store of the tile
        const int c2 = threadIdx.x % W, r0 = threadIdx.x / W;
        const bool ok = (b < a.batch) && (j0 + c2 < nc);
        const uint32_t o0 = (uint32_t)(r0 * a.ncp + c2);
        const uint32_t step = (uint32_t)(T * a.ncp);
        if (ok) {
#pragma unroll
            for (int p = 0; p < E; ++p)
                elem32(tile, o0 + (uint32_t)p * step) = s_buf[c2 * LEN + lds_pad_at<T, PS>(r0, p)];
        }
    }
}

template <int NX, int V, bool LAST, int W, bool PRUNE>
__global__ void __launch_bounds__((col_bound<NX, V, W>())) col_pass_kernel(ColPassArgs a)
{
    __shared__ float2 smem[col_smem_len<NX, V, LAST, W>()];
    col_body<NX, V, LAST, W, PRUNE>(a, (int)blockIdx.x, (int)gridDim.x, smem);
}

} // namespace msf
