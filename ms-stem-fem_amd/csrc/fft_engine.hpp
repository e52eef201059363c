// fft_engine.hpp -- register/LDS Stockham FFT for CDNA4 (gfx950).
//
// One N-point transform is computed by T = N/E cooperating lanes, each holding
// E complex values.  On entry and exit lane t holds elements t + T*p,
// p = 0..E-1 (the "distributed natural order"), so global loads and stores of
// a row are coalesced and a transform can be followed directly by an
// element-wise operation and by another transform without any LDS traffic.
// Between radix stages the values go through an LDS buffer of padded length
// lds_len(N).  Stage s with radix R and span NS (product of the previous
// radices) is the autosort Stockham step
//     y[(j/NS)*NS*R + j%NS + q*NS] = sum_r x[j + r*N/R] W_{NS R}^{r (j%NS)} W_R^{rq}
// for butterfly j = t + T*m, m = 0..E/R-1.
//
// Twiddles come from per-stage tables (see fft_stage) built on the host in
// FP64 and rounded once to the working precision, held in LDS or read from
// L2; the plans with 16 values per lane load a few entries per stage and
// derive the rest by one complex product each.
// FFT sign conventions match FFTW: forward = exp(-2 pi i), inverse =
// exp(+2 pi i), both unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace msf {

template <typename R> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b)
{
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(a.x) s) { return C{a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C mul_mi(C a)
{
    return INV ? C{-a.y, a.x} : C{a.y, -a.x};
}

// ------------------------------------------------------------ butterflies
template <int R, bool INV, typename C> struct Dft;

template <bool INV, typename C> struct Dft<2, INV, C> {
    static __device__ __forceinline__ void run(C* v)
    {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV, typename C> struct Dft<4, INV, C> {
    static __device__ __forceinline__ void run(C* v)
    {
        const C s0 = cadd(v[0], v[2]), s1 = csub(v[0], v[2]);
        const C s2 = cadd(v[1], v[3]), s3 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(s0, s2);
        v[1] = cadd(s1, s3);
        v[2] = csub(s0, s2);
        v[3] = csub(s1, s3);
    }
};

template <bool INV, typename C> struct Dft<8, INV, C> {
    static __device__ __forceinline__ void run(C* v)
    {
        using R = decltype(v[0].x);
        const R h = R(0.70710678118654752440);
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV, C>::run(e);
        Dft<4, INV, C>::run(o);
        // o[k] *= W8^k
        C o1, o3;
        if (!INV) {
            o1 = C{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};
            o3 = C{(o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h};
        } else {
            o1 = C{(o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h};
            o3 = C{-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h};
        }
        const C o2 = mul_mi<INV>(o[2]);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <bool INV, typename C> struct Dft<3, INV, C> {
    static __device__ __forceinline__ void run(C* v)
    {
        using R = decltype(v[0].x);
        const R s3 = R(0.86602540378443864676);
        const C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        const C m = C{v[0].x - R(0.5) * t1.x, v[0].y - R(0.5) * t1.y};
        const C n = cscale(mul_mi<INV>(t2), s3);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, n);
        v[2] = csub(m, n);
    }
};

template <bool INV, typename C> struct Dft<6, INV, C> {
    static __device__ __forceinline__ void run(C* v)
    {
        using R = decltype(v[0].x);
        const R s3 = R(0.86602540378443864676);
        C e[3] = {v[0], v[2], v[4]};
        C o[3] = {v[1], v[3], v[5]};
        Dft<3, INV, C>::run(e);
        Dft<3, INV, C>::run(o);
        // W6 = 1/2 -/+ i sqrt(3)/2, W6^2 = -1/2 -/+ i sqrt(3)/2 (forward/inverse)
        C o1, o2;
        if (!INV) {
            o1 = C{R(0.5) * o[1].x + s3 * o[1].y, R(0.5) * o[1].y - s3 * o[1].x};
            o2 = C{-R(0.5) * o[2].x + s3 * o[2].y, -R(0.5) * o[2].y - s3 * o[2].x};
        } else {
            o1 = C{R(0.5) * o[1].x - s3 * o[1].y, R(0.5) * o[1].y + s3 * o[1].x};
            o2 = C{-R(0.5) * o[2].x - s3 * o[2].y, -R(0.5) * o[2].y + s3 * o[2].x};
        }
        v[0] = cadd(e[0], o[0]);
        v[3] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[4] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[5] = csub(e[2], o2);
    }
};

// multiplication by W16^e = exp(-/+ 2 pi i e / 16) (forward / inverse)
__host__ __device__ constexpr double cos16(int e)
{
    switch (((e % 16) + 16) % 16) {
    case 0: return 1.0;
    case 1: case 15: return 0.92387953251128675613;
    case 2: case 14: return 0.70710678118654752440;
    case 3: case 13: return 0.38268343236508977173;
    case 4: case 12: return 0.0;
    case 5: case 11: return -0.38268343236508977173;
    case 6: case 10: return -0.70710678118654752440;
    case 7: case 9: return -0.92387953251128675613;
    default: return -1.0;
    }
}
template <int EXP, bool INV, typename C> __device__ __forceinline__ C mul_w16(C a)
{
    using R = decltype(a.x);
    constexpr int e = EXP % 16;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (e == 4) {
        return mul_mi<INV>(a);
    } else if constexpr (e == 8) {
        return C{-a.x, -a.y};
    } else {
        constexpr R c = R(cos16(e)), s = R(cos16(e - 4)); // sin(2 pi e / 16)
        return INV ? C{a.x * c - a.y * s, a.y * c + a.x * s} : C{a.x * c + a.y * s, a.y * c - a.x * s};
    }
}

// multiplication by W16^e for an exponent known after unrolling
template <bool INV, typename C> __device__ __forceinline__ C mul_w16_at(C a, int e)
{
    using R = decltype(a.x);
    e &= 15;
    if (e == 0) return a;
    if (e == 4) return mul_mi<INV>(a);
    if (e == 8) return C{-a.x, -a.y};
    if (e == 12) return mul_mi<!INV>(a);
    const R c = R(cos16(e)), s = R(cos16(e - 4));
    return INV ? C{a.x * c - a.y * s, a.y * c + a.x * s} : C{a.x * c + a.y * s, a.y * c - a.x * s};
}

// radix 16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2,
// X[k] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 W4^(n1 k1) x[4 n1 + n2]
template <bool INV, typename C> struct Dft<16, INV, C> {
    static __device__ __forceinline__ void run(C* v)
    {
        C a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[4 * n1 + n2];
            Dft<4, INV, C>::run(a[n2]);
        }
        a[1][1] = mul_w16<1, INV>(a[1][1]);
        a[1][2] = mul_w16<2, INV>(a[1][2]);
        a[1][3] = mul_w16<3, INV>(a[1][3]);
        a[2][1] = mul_w16<2, INV>(a[2][1]);
        a[2][2] = mul_w16<4, INV>(a[2][2]);
        a[2][3] = mul_w16<6, INV>(a[2][3]);
        a[3][1] = mul_w16<3, INV>(a[3][1]);
        a[3][2] = mul_w16<6, INV>(a[3][2]);
        a[3][3] = mul_w16<9, INV>(a[3][3]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            C b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            Dft<4, INV, C>::run(b);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
        }
    }
};

// ------------------------------------------------------------ LDS layout
// one padding slot every 8 elements breaks the power-of-two strides of the
// Stockham write pattern across LDS banks
// (PS: one padding slot every 2^PS elements; 3 unless a kernel asks for a
// sparser padding to fit more rows in LDS)
template <int PS = 3> __host__ __device__ constexpr int lds_pad(int e) { return e + (e >> PS); }
template <int PS = 3> __host__ __device__ constexpr int lds_len(int n) { return n + (n >> PS) + 1; }
// lds_pad(r + S * p) for r >= 0: with S a multiple of 2^PS the padding of
// the p-th element is (r >> PS) + p S / 2^PS, so the address is one per-lane
// base plus a constant per p (an immediate offset, no address register per p)
template <int S, int PS = 3> __device__ __forceinline__ int lds_pad_at(int r, int p)
{
    if constexpr (S % (1 << PS) == 0) {
        return (r + (r >> PS)) + p * (S + (S >> PS));
    } else {
        return lds_pad<PS>(r + S * p);
    }
}
// twiddle tables in LDS: stage reads W^(q k N/(NS R)) step through the table
// with power-of-two strides; one slot every 16 entries spreads them over banks
__host__ __device__ constexpr int tw_pad(int e) { return e + (e >> 4); }
__host__ __device__ constexpr int tw_len(int n) { return n + (n >> 4) + 1; }

// ------------------------------------------------------------ sync
// WAVE = true when all lanes of a transform (and its LDS region) belong to one
// wavefront: LDS instructions of a wavefront execute in order, so only the
// compiler must be kept from reordering the exchange.
template <bool WAVE> __device__ __forceinline__ void lds_sync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// ------------------------------------------------------------ one stage
// Twiddles of stage s (radix R, span NS) come from a per-stage table laid out
// [q-1][k], entry W_{NS R}^{q k}, q = 1..R-1, k = 0..NS-1, at offset TWO in
// the concatenated table of all stages (stage_twiddles): lanes with
// consecutive k read consecutive words, and the q offsets are immediates.
template <typename C, int N, int E, int R, int NS, int TWO, bool INV, bool LAST, bool WAVE,
          int PS = 3>
__device__ __forceinline__ void fft_stage(C (&x)[E], C* lds, const C* tw, int t)
{
    constexpr int T = N / E;
    constexpr int BPT = E / R; // butterflies per lane
    static_assert(E % R == 0, "radix must divide the per-lane element count");
    // Derived twiddles (16-value plans): a lane's butterflies m >= 1 of a
    // stage whose spans hold them whole take W^{q(t + T m)} = W^{q t} times
    // the constant W16^{16 q T m / (NS R)} (MDER), or the twiddles of
    // butterfly 0 when T is a multiple of NS (SAME: (t + T m) % NS = t % NS);
    // a radix-16 stage with one butterfly per lane takes W^{q k}, q = 4a + b,
    // as W^{4a k} W^{b k} (QDER).  Each derived twiddle is one product of
    // table entries or of a table entry and a W16 constant (about one ulp
    // more than the table's); the 1024 alternate plan loads 9 instead of 27
    // per transform, and its row pass, bound by load latency at 4 waves per
    // SIMD, goes from 71.7 to 59.9 us per 17 positions.
    constexpr bool DERIVE = E >= 16 && NS > 1;
    constexpr bool SAME = DERIVE && BPT > 1 && T % NS == 0;
    constexpr bool MDER = DERIVE && !SAME && BPT > 1 && T * BPT <= NS && (16 * T) % (NS * R) == 0;
    constexpr bool QDER = DERIVE && R == 16 && BPT == 1;
    C w0[R];
    C y[E];
#pragma unroll
    for (int m = 0; m < BPT; ++m) {
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = x[m + q * BPT];
        if constexpr (MDER || SAME) {
            if (m == 0) {
                const C* twk = tw + TWO + t % NS;
#pragma unroll
                for (int q = 1; q < R; ++q) w0[q] = twk[(q - 1) * NS];
            }
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const C w = SAME ? w0[q] : mul_w16_at<false>(w0[q], (16 * q * T * m) / (NS * R));
                v[q] = INV ? cmulc(v[q], w) : cmul(v[q], w);
            }
        } else if constexpr (QDER) {
            const C* twk = tw + TWO + t % NS;
            C wb[4], wa[4];
#pragma unroll
            for (int b = 1; b < 4; ++b) wb[b] = twk[(b - 1) * NS];
#pragma unroll
            for (int a4 = 1; a4 < 4; ++a4) wa[a4] = twk[(4 * a4 - 1) * NS];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const int a4 = q / 4, b = q % 4;
                const C w = a4 == 0 ? wb[b] : (b == 0 ? wa[a4] : cmul(wa[a4], wb[b]));
                v[q] = INV ? cmulc(v[q], w) : cmul(v[q], w);
            }
        } else if constexpr (NS > 1) {
            const C* twk = tw + TWO + (t + T * m) % NS;
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const C w = twk[(q - 1) * NS];
                v[q] = INV ? cmulc(v[q], w) : cmul(v[q], w);
            }
        }
        Dft<R, INV, C>::run(v);
#pragma unroll
        for (int q = 0; q < R; ++q) y[m * R + q] = v[q];
    }
    if constexpr (LAST) {
#pragma unroll
        for (int m = 0; m < BPT; ++m)
#pragma unroll
            for (int q = 0; q < R; ++q) x[m + q * BPT] = y[m * R + q];
    } else {
        lds_sync<WAVE>(); // previous readers of this buffer are done
        // Plans with 16 values per lane address the exchange as one base
        // plus constant offsets (fewer address registers: the 1024 column
        // pass 177 -> 120 registers); the 8-value plans measure faster with
        // an address per element (512 column pass 39.6 -> 37.1 us).
        constexpr bool FOLD = E >= 16;
#pragma unroll
        for (int m = 0; m < BPT; ++m) {
            const int j = t + T * m;
            const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
            for (int q = 0; q < R; ++q)
                lds[FOLD ? lds_pad_at<NS, PS>(base, q) : lds_pad<PS>(base + q * NS)] = y[m * R + q];
        }
        lds_sync<WAVE>();
#pragma unroll
        for (int p = 0; p < E; ++p)
            x[p] = lds[FOLD ? lds_pad_at<T, PS>(t, p) : lds_pad<PS>(t + T * p)];
    }
}

// stage offsets in the concatenated twiddle table: stage s with span NS > 1
// adds (R-1)*NS entries; the first stage (NS = 1) needs none
template <typename C, int N, int E, bool INV, bool WAVE, int PS, int NS, int TWO, int R,
          int... Rest>
__device__ __forceinline__ void run_stages(C (&x)[E], C* lds, const C* tw, int t)
{
    fft_stage<C, N, E, R, NS, TWO, INV, sizeof...(Rest) == 0, WAVE, PS>(x, lds, tw, t);
    if constexpr (sizeof...(Rest) > 0)
        run_stages<C, N, E, INV, WAVE, PS, NS * R, (NS > 1 ? TWO + (R - 1) * NS : 0), Rest...>(
            x, lds, tw, t);
}

// ------------------------------------------------------------ plans
// E = values per lane; radices applied first to last.  kRadix lists them for
// the host code that builds the twiddle tables.  Plan<N> is the primary
// schedule of each length (also used by the FP64 transforms); AltPlan<N>, where
// it exists, a second schedule of the FP32 slice passes with 16 values per lane
// and one fewer radix stage, chosen per run (configuration or autotuner).
template <int N> struct Plan;
template <int N> struct AltPlan { static constexpr bool exists = false; };
#define MSF_PLAN_AS(NAME, N_, E_, ...)                                                         \
    template <> struct NAME<N_> {                                                              \
        static constexpr bool exists = true;                                                   \
        static constexpr int N = N_, E = E_, T = N_ / E_;                                      \
        static constexpr int kRadix[] = {__VA_ARGS__};                                         \
        static constexpr int kStages = sizeof(kRadix) / sizeof(int);                           \
        template <bool INV, typename C, bool WAVE = false, int PS = 3>                         \
        static __device__ __forceinline__ void run(C (&x)[E_], C* lds, const C* tw, int t)     \
        {                                                                                      \
            run_stages<C, N_, E_, INV, WAVE, PS, 1, 0, __VA_ARGS__>(x, lds, tw, t);            \
        }                                                                                      \
    };
#define MSF_PLAN(N_, E_, ...) MSF_PLAN_AS(Plan, N_, E_, __VA_ARGS__)
#define MSF_ALT_PLAN(N_, E_, ...) MSF_PLAN_AS(AltPlan, N_, E_, __VA_ARGS__)
MSF_PLAN(16, 8, 8, 2)
MSF_PLAN(32, 8, 8, 4)
MSF_PLAN(64, 8, 8, 8)
MSF_PLAN(96, 12, 6, 4, 4)
MSF_PLAN(128, 8, 8, 4, 4)
MSF_PLAN(192, 24, 6, 8, 4)
MSF_PLAN(256, 8, 8, 8, 4)
MSF_PLAN(384, 12, 6, 4, 4, 4)
MSF_PLAN(512, 8, 8, 8, 8)
MSF_PLAN(768, 24, 6, 8, 4, 4)
MSF_PLAN(1024, 8, 8, 8, 4, 4)
MSF_ALT_PLAN(256, 16, 16, 16)
MSF_ALT_PLAN(512, 16, 16, 8, 4)
MSF_ALT_PLAN(1024, 16, 16, 16, 4)
#undef MSF_ALT_PLAN
#undef MSF_PLAN
#undef MSF_PLAN_AS

// plan V of length N: 0 = primary, 1 = alternate
template <int N, int V> struct PlanSel { using type = Plan<N>; };
template <int N> struct PlanSel<N, 1> { using type = AltPlan<N>; };
template <int N, int V> using PlanV = typename PlanSel<N, V>::type;

// entries of the concatenated stage tables: sum (R-1) NS over stages 1.. = N - R_0
template <typename P> constexpr int plan_tw_len() { return P::N - P::kRadix[0]; }
template <int N> constexpr int stage_tw_len() { return plan_tw_len<Plan<N>>(); }

} // namespace msf
