// msf_api.cpp -- C ABI (include/msf.h): specimen set-up, probe, raster loop.
//
// Host orchestration of the device kernels.  Reference seams replaced:
//   tem.cpp:944-1276 (species templates, auto-slicing, propagators,
//   transmission functions)  -> msf_specimen_build
//   TEM_NS::adfstem (adfstem.hpp:66-121)  -> msf_run_stem
#include "../../include/msf.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "msf_host.hpp"

namespace {

const double kKirkland[103][12] = {
#include "kirkland_params.inc"
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        release();
        if (n == 0) n = 16;
        bytes = n;
        return hipMalloc(&p, n);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

std::atomic<uint64_t> g_specimen_ids{1};

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

} // namespace

// Everything a raster run needs besides the positions, cached on the context
// and rebuilt only when the specimen, optics, detector or FEM settings change.
struct RunPlan {
    uint64_t spec_id = 0;
    msf_optics optics{};
    msf_detector det{};
    msf_fem fem{};
    bool valid = false;
    int ncp = 0, G = 0, nbins = 0;
    DevBuf U0, Psi0;                          // vacuum waves
    DevBuf seg_pix, tile_off, seg_kind, seg_of_group, cnt, zero;  // reduction groups
    int n_tiles = 0;
    // the pixel groups depend on the grid, the aperture and the detector/FEM
    // settings only: they are kept across specimens on the same grid
    bool groups_ok = false;
    msf_grid ggrid{};
    double galpha = 0.0;
    bool fem_sqr = false;                     // groups carry the |psi|^2 FEM rings (Vbar_r)
    bool corr = false;                        // groups carry the correlograph sectors
    msf::Groups hg;                           // host copy for the per-pixel FEM modes
    int64_t cap_b = 0, cap_p = 0;
    DevBuf psi, part, segp, state, femv, adf, shift, exit;
    DevBuf psi2, part2, segp2, mtf_full2;     // the second batch chain's work buffers
    DevBuf vsum, pm1, pm2;                    // Vbar_r sums; per-pixel |psi|, |psi|^2 sums
    DevBuf csum, csqr;                        // correlograph sums of c and c^2
    DevBuf ceach;                             // correlograph of every position
    // detector MTF: the table it was built from, the 2-D MTF [nx][ny], the
    // outside-aperture map and the full-grid FP64 work planes [B][nx][ny]
    bool mtf_on = false;
    std::vector<double> mtf_v, mtf_d;
    double mtf_rate = 0.0;
    DevBuf mtf2d, mtf_out, mtf_full;
    // the job's launch sequence as an instantiated HIP graph, replayed while
    // every pointer and size it was captured with is unchanged
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_sig;
    RunPlan() = default;
    RunPlan(const RunPlan&) = delete;
    RunPlan& operator=(const RunPlan&) = delete;
    ~RunPlan() { drop_graph(); }
    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_sig.clear();
    }
    // Frees the device memory of the last run except the pixel-group tables
    // (a few MB, keyed by grid and settings): the next run rebuilds the rest,
    // and reuses the groups when its grid and settings match.
    void release_work()
    {
        drop_graph();
        valid = false;
        spec_id = 0;
        for (DevBuf* b : {&U0, &Psi0, &psi, &part, &segp, &state, &femv, &adf, &shift, &exit, &psi2,
                          &part2, &segp2, &mtf_full2, &vsum, &pm1, &pm2, &csum, &csqr, &ceach,
                          &mtf_full})
            b->release();
        cap_b = cap_p = 0;
    }
};

// A job of the serving queue: inputs copied at submission, outputs written
// to the caller's arrays.
struct ServeJob {
    uint64_t id = 0;
    msf_grid grid{};
    std::vector<double> q;
    std::vector<uint32_t> Z;
    double min_slice = 0;
    msf_optics optics{};
    std::vector<double> xp, yp;
    msf_detector det{};
    msf_fem fem{};
    double* adf = nullptr;
    msf_fem_outputs fo{};
    std::vector<double> mtf, mtf_dom; // copies of the MTF table (fem.mtf* point here)
    msf_config cfg{};                 // the submitting context's settings at submission
    uint64_t footprint = 0;           // estimated device bytes (admission)
    double t_submit = 0;
};

struct ServeDone {
    int rc = MSF_OK;
    std::string msg;
    msf_job_info info{};
};

// Continuous batching: `slots` workers, each with its own internal context
// (stream, run plan), take jobs from the head of the queue as soon as a slot
// is free and the head job's estimated footprint fits beside the jobs
// already running (a job runs alone if it exceeds the limit by itself).
struct ServeState {
    std::vector<std::thread> workers;
    std::vector<msf_ctx*> slot_ctx;
    std::mutex m;
    std::condition_variable work_cv, done_cv;
    std::deque<ServeJob> queue;
    std::set<uint64_t> pending;                              // queued or running
    std::map<uint64_t, ServeDone> done;
    uint64_t next_id = 1;
    bool stop = false;
    uint32_t slots = 0;
    uint64_t limit = 0, admitted = 0, peak_admitted = 0, jobs_done = 0;
    uint32_t running = 0, max_running = 0;
};

struct msf_ctx {
    std::recursive_mutex mu;    // serialises device work of the context
    std::mutex serve_mu;        // guards the creation of the serving engine
    std::unique_ptr<ServeState> serve;
    uint32_t serve_slots = 0;   // msf_serve_configure (0: automatic)
    uint64_t serve_limit = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream and the events that order two batch chains (created on first use)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fem[2] = {nullptr, nullptr};
    std::string err;
    msf_config cfg{0, MSF_TRANSMIT_REFERENCE, 0, 0, 0, 0, 0, 0};
    msf_stats stats{};
    std::map<int, DevBuf> tw32, tw64, twnat, tw32alt;
    // autotuned launch shape per (nx, ny, ncp, min(P, 256)): batch, column tile,
    // chains, radix plan
    std::map<std::array<int, 4>, std::array<int, 4>> tuned;
    RunPlan plan;
    std::set<msf_specimen*> specimens; // live specimens built on this context
};

struct msf_specimen {
    msf_ctx* ctx = nullptr;
    msf_grid grid{};
    msf::Domain dom;
    int n_slices = 0;
    std::vector<double> lower, thick;
    std::vector<uint32_t> counts;
    DevBuf t32;  // [S][nx][ny] float2, t / sqrtNxNy
    DevBuf t64;  // [S][nx][ny] double2, t as built
    DevBuf px;   // [S][nx] float2
    DevBuf py;   // [S][ny] float2, P_y / sqrtNxNy
    DevBuf px64; // [S][nx] double2
    DevBuf py64; // [S][ny] double2
    uint64_t id = 0;           // unique per build, keys the run-plan cache
    msf::RowRanges keep;       // columns the slice step keeps (bw_limit + propagate)
    DevBuf dkeep;              // keep as int2 [nx]
};

namespace {

int fail(msf_ctx* c, int code, const std::string& m)
{
    if (c) c->err = m;
    return code;
}

// Every entry point turns a C++ exception (a failed host allocation, a length
// error from a malformed size) into an error code: none may cross the C ABI.
#define MSF_CATCH_ALL(c)                                                                      \
    catch (const std::bad_alloc&)                                                             \
    {                                                                                         \
        return fail((c), MSF_ERR_MEMORY, "host allocation failed");                          \
    }                                                                                         \
    catch (const std::exception& e_)                                                          \
    {                                                                                         \
        return fail((c), MSF_ERR_INVALID, std::string("unexpected error: ") + e_.what());    \
    }                                                                                         \
    catch (...)                                                                               \
    {                                                                                         \
        return fail((c), MSF_ERR_INVALID, "unexpected error");                               \
    }

// device work of a context is serialised (direct calls and the serving worker)
#define MSF_LOCK(c) std::lock_guard<std::recursive_mutex> msf_lock_((c)->mu)

// A failed call is reported once: the runtime's last-error state is cleared,
// so the next launch's hipGetLastError() check reports only its own status.
#define HIPCHK(ctx, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (void)hipGetLastError();                                                          \
            return fail((ctx), e_ == hipErrorOutOfMemory ? MSF_ERR_MEMORY : MSF_ERR_DEVICE,    \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
        }                                                                                     \
    } while (0)

template <typename T> hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t s)
{
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

// Per-stage twiddle tables of the length-n transform plan (FP64 and FP32)
// and the natural roots W_n^e = exp(-2 pi i e / n) in FP32, built in FP64 on
// the host and rounded once.
int twiddles(msf_ctx* c, int n, const float2** f32, const double2** f64,
             const float2** nat32 = nullptr)
{
    if (!c->tw32.count(n)) {
        std::vector<double> st;
        if (!msf::stage_twiddles(n, st)) return fail(c, MSF_ERR_UNSUPPORTED, "no plan for this length");
        std::vector<double2> d(st.size() / 2);
        std::vector<float2> f(st.size() / 2), nat(n);
        for (size_t e = 0; e < d.size(); ++e) {
            d[e] = double2{st[2 * e], st[2 * e + 1]};
            f[e] = float2{(float)st[2 * e], (float)st[2 * e + 1]};
        }
        for (int e = 0; e < n; ++e) {
            const double a = 2.0 * M_PI * (double)e / (double)n;
            nat[e] = float2{(float)std::cos(a), (float)-std::sin(a)};
        }
        HIPCHK(c, upload(c->tw64[n], d, c->stream));
        HIPCHK(c, upload(c->tw32[n], f, c->stream));
        HIPCHK(c, upload(c->twnat[n], nat, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (f32) *f32 = c->tw32[n].as<float2>();
    if (f64) *f64 = c->tw64[n].as<double2>();
    if (nat32) *nat32 = c->twnat[n].as<float2>();
    return MSF_OK;
}

// Stage twiddles of the alternate radix plan of length n (FP32), for the
// slice passes only; null when the length has none.
int alt_twiddles(msf_ctx* c, int n, const float2** f32)
{
    *f32 = nullptr;
    if (!msf::has_alt_plan(n)) return MSF_OK;
    if (!c->tw32alt.count(n)) {
        std::vector<double> st;
        if (!msf::stage_twiddles(n, st, 1)) return fail(c, MSF_ERR_UNSUPPORTED, "alternate plan table");
        std::vector<float2> f(st.size() / 2);
        for (size_t e = 0; e < f.size(); ++e) f[e] = float2{(float)st[2 * e], (float)st[2 * e + 1]};
        HIPCHK(c, upload(c->tw32alt[n], f, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    *f32 = c->tw32alt[n].as<float2>();
    return MSF_OK;
}

// Radix plan of one axis of the slice passes when none is configured or
// tuned (0 primary, 1 alternate), by measurement on the 256^2 Si, 512^2 Cu/Zr
// and 1024^2 Si workloads: the alternate plans (16 values per lane, one
// fewer stage) are 6-7 % faster over a job at 256 and 1024 and 10 % slower
// at 512.
int fft_plan_rule(int n)
{
    return (msf::has_alt_plan(n) && n != 512) ? 1 : 0;
}

int check_grid(msf_ctx* c, const msf_grid* g)
{
    if (!g) return fail(c, MSF_ERR_INVALID, "grid is NULL");
    if (!msf::slice_lengths_ok((int)g->nx) || !msf::slice_lengths_ok((int)g->ny))
        return fail(c, MSF_ERR_UNSUPPORTED,
                    "no compiled transform for samples " + std::to_string(g->nx) + " x " +
                        std::to_string(g->ny));
    if (!(g->xperiod > 0.0) || !(g->yperiod > 0.0) || !(g->zperiod > 0.0))
        return fail(c, MSF_ERR_INVALID, "periods must be positive (tem_io.cpp:4681-4719)");
    if (!(g->voltage > 0.0)) return fail(c, MSF_ERR_INVALID, "microscope_voltage must be positive");
    return MSF_OK;
}

std::vector<int2> to_int2(const msf::RowRanges& r)
{
    std::vector<int2> v(r.ab.size() / 2);
    for (size_t i = 0; i < v.size(); ++i) v[i] = int2{r.ab[2 * i], r.ab[2 * i + 1]};
    return v;
}

// The 2-D detector MTF (adfstem.cpp:300-439): the table sampled on the grid
// on the host in FP64, then on the device BACKWARD FFT, bw_limit applied to
// that real-space array by k index (as the reference does), FORWARD FFT,
// / NxNy.  out: [nx][ny] double2.
int build_mtf2d(msf_ctx* c, const msf::Domain& d, double xperiod, double yperiod,
                const msf_fem* fem, DevBuf& out)
{
    const std::vector<double> m = msf::mtf_table_2d(d, xperiod, yperiod, fem->mtf,
                                                    fem->mtf_domain, (int)fem->mtf_n,
                                                    fem->mtf_sampling_rate);
    std::vector<double2> h(m.size());
    for (size_t k = 0; k < m.size(); ++k) h[k] = double2{m[k], 0.0};
    const double2 *twx, *twy;
    if (int r = twiddles(c, d.nx, nullptr, &twx)) return r;
    if (int r = twiddles(c, d.ny, nullptr, &twy)) return r;
    DevBuf keep;
    HIPCHK(c, upload(out, h, c->stream));
    HIPCHK(c, upload(keep, to_int2(msf::bwlimit_keep_ranges(d)), c->stream));
    double2* v = out.as<double2>();
    HIPCHK(c, msf::fft2d_f64(v, d.nx, d.ny, 1, true, twx, twy, c->stream));
    HIPCHK(c, msf::launch_apply_keep(v, d.nx, d.ny, 1, keep.as<int2>(), c->stream));
    HIPCHK(c, msf::fft2d_f64(v, d.nx, d.ny, 1, false, twx, twy, c->stream));
    HIPCHK(c, msf::launch_div_f64(v, (int64_t)m.size(), d.nxny, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream)); // keep and h are released on return
    return MSF_OK;
}

int check_mtf(msf_ctx* c, const msf_fem* fem)
{
    if (!fem || !fem->mtf || !fem->mtf_domain || fem->mtf_n < 1)
        return fail(c, MSF_ERR_INVALID, "no MTF table");
    if (fem->mtf_sampling_rate == 0.0)
        return fail(c, MSF_ERR_INVALID, "mtf_sampling_rate must be non-zero");
    return MSF_OK;
}

// H(k) -> BACKWARD FFT; leaves h = IFFT(H) (unnormalised) and its squared norm
int probe_device(msf_ctx* c, const msf::Domain& d, const msf_optics* o, DevBuf& h, DevBuf& norm2)
{
    const int nx = d.nx, ny = d.ny;
    const msf::RowRanges ap = msf::aperture_keep_ranges(d, o->alpha_max);
    if (!ap.ok) return fail(c, MSF_ERR_INVALID, "aperture is not a disk in FFT order");
    const double2 *twx, *twy;
    if (int r = twiddles(c, nx, nullptr, &twx)) return r;
    if (int r = twiddles(c, ny, nullptr, &twy)) return r;
    DevBuf dkx, dky, dap;
    HIPCHK(c, upload(dkx, d.kx, c->stream));
    HIPCHK(c, upload(dky, d.ky, c->stream));
    HIPCHK(c, upload(dap, to_int2(ap), c->stream));
    HIPCHK(c, h.alloc((size_t)nx * ny * sizeof(double2)));
    HIPCHK(c, norm2.alloc(sizeof(double)));
    HIPCHK(c, msf::launch_probe_k(h.as<double2>(), dkx.as<double>(), dky.as<double>(), nx, ny,
                                  dap.as<int2>(), o->corrected, o->cs3, o->cs5, o->defocus,
                                  d.lambda, d.xx[0], d.yy[0], c->stream));
    HIPCHK(c, msf::fft2d_f64(h.as<double2>(), nx, ny, 1, true, twx, twy, c->stream));
    HIPCHK(c, msf::launch_norm2(h.as<double2>(), (int64_t)nx * ny, norm2.as<double>(), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MSF_OK;
}

int check_optics(msf_ctx* c, const msf_optics* o)
{
    if (!o) return fail(c, MSF_ERR_INVALID, "optics is NULL");
    if (!(o->alpha_max > 0.0)) return fail(c, MSF_ERR_INVALID, "alphamax must be positive");
    if (o->corrected != 0 && o->corrected != 1)
        return fail(c, MSF_ERR_INVALID, "corrected must be 0 or 1");
    return MSF_OK;
}

struct RunOut {
    double* adf = nullptr;
    double* fem = nullptr;      // V_omega (d1)
    double* vbar_r = nullptr;   // d2
    double* v_re = nullptr;     // d3
    double* omega_vimage = nullptr; // d4
    double* v_gt17 = nullptr;
    double* v_rbar = nullptr;   // rva
    double* corr_avg = nullptr; // correlograph, [ii + jj * nbins]
    double* corr_sqr = nullptr;
    double* corr_each = nullptr; // [p][ii + jj * nbins]
    double* mean_abs = nullptr;  // per-pixel means over positions, [i * ny + j]
    double* mean_sqr = nullptr;
    double* var2d = nullptr;
    float* exit = nullptr;
};

// V_re, Omega_vimage, V_gt17 and V_rbar from the per-pixel sums of |psi|
// (s1) and |psi|^2 (s2) over P positions, on the compact plane
// (adfstem.cpp:2942-2995 averages; :3169-3466 ring ratios; :3468-3605 the
// variance image and its annular mean).  Ring sums run over every pixel of
// a bin in the bins' sorted order; pixels the band limit leaves zero
// contribute zero sums (and 0/0 to the variance image), as in the reference.
void fem_pixel_modes(const msf::Domain& d, const msf::Groups& g, const std::vector<double>& s1,
                     const std::vector<double>& s2, int64_t P, double alpha_max, const RunOut& out)
{
    const int nb = g.nbins;
    const double np = (double)P;
    const double l2 = d.lambda * d.lambda, a2 = alpha_max * alpha_max; // adfstem.cpp:95-96
    const std::vector<double> kx2 = msf::pow2_all(d.kx), ky2 = msf::pow2_all(d.ky);
    for (int ii = 0; ii < nb; ++ii) {
        double S2 = 0.0, Q2 = 0.0, S1 = 0.0, D = 0.0;
        for (int32_t k = g.bin_off[ii]; k < g.bin_off[ii + 1]; ++k) {
            const int32_t cp = g.bin_compact[k];
            const double mag = (cp >= 0 ? s1[cp] : 0.0) / np;
            const double sqr = (cp >= 0 ? s2[cp] : 0.0) / np;
            S2 += sqr;
            Q2 += mag * mag;
            S1 += mag;
            if (out.omega_vimage) {
                const int i = g.bin_full[k] / d.ny, j = g.bin_full[k] % d.ny;
                double v = (sqr / (mag * mag)) - 1.0;
                const double ksqr = kx2[i] + ky2[j];
                if (l2 * ksqr <= a2) v = 0.0;
                D += v;
            }
        }
        if (g.zeroed[ii]) S2 = Q2 = S1 = D = 0.0;
        const int c = g.counts[ii];
        const double cc = (double)(c * c);
        if (out.v_gt17) out.v_gt17[ii] = (S2 / c) / (Q2 / c);
        if (out.v_re) out.v_re[ii] = ((S2 / c) / (S1 * S1 / cc)) - 1.0;
        if (out.v_rbar) out.v_rbar[ii] = ((Q2 / c) / (S1 * S1 / cc)) - 1.0;
        if (out.omega_vimage) out.omega_vimage[ii] = D / c;
    }
}

// The per-pixel means of |psi| and |psi|^2 over positions on the full grid
// (adfstem.cpp:2968-2990) and the 2-D variance image (adfstem.cpp:3468-3492):
// <|psi|^2> / <|psi|>^2 - 1, 0 where lambda^2 k^2 <= alpha_max^2.  Pixels
// outside the stored columns have zero sums.
void fem_pixel_images(const msf::Domain& d, const msf::RowRanges& keep, int ncp,
                      const std::vector<double>& s1, const std::vector<double>& s2, int64_t P,
                      double alpha_max, const RunOut& out)
{
    const int A = keep.amax, B = keep.bmax;
    const double np = (double)P;
    const double l2 = d.lambda * d.lambda, a2 = alpha_max * alpha_max; // adfstem.cpp:95-96
    const std::vector<double> kx2 = msf::pow2_all(d.kx), ky2 = msf::pow2_all(d.ky);
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j) {
            const int jc = (j < A) ? j : (j >= d.ny - B ? j - d.ny + A + B : -1);
            const int64_t cp = jc >= 0 ? (int64_t)i * ncp + jc : -1;
            const double mag = (cp >= 0 ? s1[cp] : 0.0) / np;
            const double sqr = (cp >= 0 ? s2[cp] : 0.0) / np;
            const size_t k = (size_t)i * d.ny + j;
            if (out.mean_abs) out.mean_abs[k] = mag;
            if (out.mean_sqr) out.mean_sqr[k] = sqr;
            if (out.var2d) {
                double v = (sqr / (mag * mag)) - 1.0;
                const double ksqr = kx2[i] + ky2[j];
                if (l2 * ksqr <= a2) v = 0.0;
                out.var2d[k] = v;
            }
        }
}

// Probe shifts in raster order (x outer, y inner), each index persisting
// from the previous position when no grid point qualifies
// (adfstem.cpp:1476-1477, :1590-1606).  False if a position lies below the
// domain, where the reference's size_t index would wrap.
bool raster_shifts(const msf::Domain& d, const double* x_p, uint32_t n_xp, const double* y_p,
                   uint32_t n_yp, std::vector<int2>& shifts)
{
    shifts.resize((size_t)n_xp * n_yp);
    int64_t sx = 0, sy = 0;
    for (uint32_t a = 0; a < n_xp; ++a)
        for (uint32_t b = 0; b < n_yp; ++b) {
            sx = msf::shift_index(d.xx, x_p[a], sx);
            sy = msf::shift_index(d.yy, y_p[b], sy);
            if (sx < 0 || sy < 0) return false;
            shifts[(size_t)a * n_yp + b] = int2{(int)sx, (int)sy};
        }
    return true;
}

int run_internal(msf_ctx* c, const msf_specimen* sp, const msf_optics* o,
                 const msf_detector* det, const msf_fem* fem, const std::vector<int2>& shifts,
                 RunOut out)
{
    const double t_start = now_ms();
    if (!sp) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
    if (sp->ctx != c) return fail(c, MSF_ERR_INVALID, "specimen was built on another context");
    if (int r = check_optics(c, o)) return r;
    if (shifts.empty()) return fail(c, MSF_ERR_INVALID, "empty raster");
    const msf::Domain& d = sp->dom;
    const int nx = d.nx, ny = d.ny, S = sp->n_slices;
    const bool want_adf = det && det->enabled;
    const bool fem_on = fem && fem->enabled;
    const bool want_d1 = fem_on && out.fem;
    const bool want_d2 = fem_on && out.vbar_r;
    const bool want_pix = fem_on && (out.v_re || out.omega_vimage || out.v_gt17 || out.v_rbar ||
                                     out.mean_abs || out.mean_sqr || out.var2d);
    const bool want_corr = fem_on && (out.corr_avg || out.corr_sqr || out.corr_each);
    const bool want_fem = want_d1 || want_d2 || want_pix || want_corr;
    // the detector MTF changes only what the FEM statistics see
    const bool want_mtf = want_fem && fem->mtf && fem->mtf_domain && fem->mtf_n >= 1 &&
                          fem->mtf_sampling_rate != 0.0;
    if (want_adf) {
        if (det->inner < 0.0 || det->outer < 0.0)
            return fail(c, MSF_ERR_INVALID, "detector angles must be >= 0 (tem.cpp:710-720)");
        if (det->inner > d.bw_t * d.lambda)
            return fail(c, MSF_ERR_INVALID,
                        "detector inner angle falls beyond bandwidth limits (tem.cpp:721-732)");
    }
    if (want_fem && !(fem->dr > 0.0)) return fail(c, MSF_ERR_INVALID, "dr must be positive");
    const int64_t P = (int64_t)shifts.size();
    const int tmode = c->cfg.transmit_mode;
    for (const int2& sh : shifts)
        if (sh.x < 0 || sh.x >= nx || sh.y < 0 || sh.y >= ny)
            return fail(c, MSF_ERR_DOMAIN, "probe shift outside the grid");

    const msf::RowRanges& keep = sp->keep;
    const int W = msf::col_tile(nx);
    const int ka = keep.amax, kb = keep.bmax;
    const int ncp = std::max(W, ((ka + kb + W - 1) / W) * W);
    const int64_t M = (int64_t)nx * ny;
    const int64_t plane = (int64_t)nx * ncp;

    RunPlan& pl = c->plan;
    // cache keys built member by member (no padding bytes from the caller);
    // the MTF table is compared by content below
    msf_detector dkey;
    msf_fem fkey;
    std::memset(&dkey, 0, sizeof dkey);
    std::memset(&fkey, 0, sizeof fkey);
    if (want_adf) {
        dkey.enabled = det->enabled;
        dkey.inner = det->inner;
        dkey.outer = det->outer;
        dkey.threshold_mode = det->threshold_mode;
    }
    if (want_fem) {
        fkey.enabled = fem->enabled;
        fkey.dr = fem->dr;
    }
    if (want_mtf) {
        fkey.mtf_n = fem->mtf_n;
        fkey.mtf_sampling_rate = fem->mtf_sampling_rate;
    }
    bool same_mtf = pl.mtf_on == want_mtf;
    if (same_mtf && want_mtf)
        same_mtf = std::equal(pl.mtf_v.begin(), pl.mtf_v.end(), fem->mtf) &&
                   std::equal(pl.mtf_d.begin(), pl.mtf_d.end(), fem->mtf_domain) &&
                   pl.mtf_v.size() == fem->mtf_n;
    const bool same_optics = pl.valid && pl.spec_id == sp->id &&
                             std::memcmp(&pl.optics, o, sizeof(msf_optics)) == 0;
    msf_grid gkey;
    std::memset(&gkey, 0, sizeof gkey);
    gkey.nx = sp->grid.nx;
    gkey.ny = sp->grid.ny;
    gkey.xperiod = sp->grid.xperiod;
    gkey.yperiod = sp->grid.yperiod;
    gkey.zperiod = sp->grid.zperiod;
    gkey.voltage = sp->grid.voltage;
    gkey.xmin = sp->grid.xmin;
    gkey.ymin = sp->grid.ymin;
    gkey.zmin = sp->grid.zmin;
    const bool same_groups = pl.groups_ok && std::memcmp(&pl.ggrid, &gkey, sizeof gkey) == 0 &&
                             pl.galpha == o->alpha_max &&
                             std::memcmp(&pl.det, &dkey, sizeof dkey) == 0 &&
                             std::memcmp(&pl.fem, &fkey, sizeof fkey) == 0 &&
                             pl.fem_sqr == want_d2 && pl.corr == want_corr && same_mtf;
    if (!same_optics || !same_groups) pl.valid = false;

    if (!same_groups) {
        // the tables below are replaced one by one: until all of them and
        // their keys are written, no later call may take them as valid
        pl.groups_ok = false;
        double k1 = 0.0, k2 = 0.0;
        if (want_adf) {
            k1 = det->inner / d.lambda;
            k2 = det->outer / d.lambda;
            if (det->threshold_mode == MSF_ADF_PHYSICAL) {
                k1 = k1 * k1;
                k2 = k2 * k2;
            }
        }
        const msf::Groups g = msf::build_groups(d, keep, ncp, want_adf, k1, k2, want_fem,
                                                want_fem ? fem->dr : 1.0, o->alpha_max, want_d2,
                                                want_corr);
        pl.G = (int)g.kind.size();
        pl.nbins = g.nbins;
        pl.fem_sqr = want_d2;
        pl.corr = want_corr;
        const msf::SegmentTiles st = msf::build_segments(g, 32);
        pl.n_tiles = st.n_tiles;
        HIPCHK(c, upload(pl.seg_pix, st.pix, c->stream));
        HIPCHK(c, upload(pl.tile_off, st.tile_off, c->stream));
        HIPCHK(c, upload(pl.seg_kind, st.seg_kind, c->stream));
        HIPCHK(c, upload(pl.seg_of_group, st.seg_of_group, c->stream));
        HIPCHK(c, upload(pl.cnt, g.counts, c->stream));
        HIPCHK(c, upload(pl.zero, g.zeroed, c->stream));
        pl.cap_b = 0; // partial-sum buffers depend on G
        pl.hg = g;
        pl.ggrid = gkey;
        pl.galpha = o->alpha_max;
        pl.det = dkey;
        pl.fem = fkey;
        pl.mtf_on = false;
        pl.mtf_v.clear();
        pl.mtf_d.clear();
        if (want_mtf) {
            // the MTF's bw_limit must keep exactly the compact layout's pixels
            if (msf::bwlimit_keep_ranges(d).ab != keep.ab)
                return fail(c, MSF_ERR_UNSUPPORTED,
                            "MTF band limit differs from the slice band limit on this grid");
            if (int r = build_mtf2d(c, d, sp->grid.xperiod, sp->grid.yperiod, fem, pl.mtf2d))
                return r;
            HIPCHK(c, upload(pl.mtf_out, msf::outside_aperture(d, o->alpha_max), c->stream));
            pl.mtf_v.assign(fem->mtf, fem->mtf + fem->mtf_n);
            pl.mtf_d.assign(fem->mtf_domain, fem->mtf_domain + fem->mtf_n);
            pl.mtf_rate = fem->mtf_sampling_rate;
            pl.mtf_on = true;
        }
        pl.groups_ok = true;
    }
    const int G = pl.G;
    const int nbins = pl.nbins;
    const int nphi = want_corr ? pl.hg.nphi : 0;
    const size_t ncorr = (size_t)std::max(nbins * nphi, 1);

    if (!same_optics) {
        // vacuum wave (t = 1) of every slice, FP64, from the probe of
        // adfstem.cpp:586-817: u0_0 = s * A_p * IFFT(H)/s; then per slice
        // w = FFT(u0_n) / s^2 * M P_n (adfstem.cpp:1986-2043), u0_{n+1} = IFFT(w)
        DevBuf h, norm2;
        if (int r = probe_device(c, d, o, h, norm2)) return r;
        const double2 *twx64, *twy64;
        if (int r = twiddles(c, nx, nullptr, &twx64)) return r;
        if (int r = twiddles(c, ny, nullptr, &twy64)) return r;
        HIPCHK(c, pl.U0.alloc((size_t)S * 2 * M * sizeof(float2)));
        HIPCHK(c, pl.Psi0.alloc((size_t)M * sizeof(float2)));
        double2* v = h.as<double2>();
        HIPCHK(c, msf::launch_scale_f64(v, M, norm2.as<double>(), d.sqrt_nxny, c->stream));
        for (int n = 0; n < S; ++n) {
            HIPCHK(c, msf::launch_to_f32_doubled_rows(v, nx, ny, pl.U0.as<float2>() + (int64_t)n * 2 * M,
                                                      c->stream));
            HIPCHK(c, msf::fft2d_f64(v, nx, ny, 1, false, twx64, twy64, c->stream));
            HIPCHK(c, msf::launch_vacuum_step(v, nx, ny, sp->px64.as<double2>() + (int64_t)n * nx,
                                              sp->py64.as<double2>() + (int64_t)n * ny,
                                              sp->dkeep.as<int2>(), 1.0 / (d.sqrt_nxny * d.sqrt_nxny),
                                              c->stream));
            if (n + 1 < S) HIPCHK(c, msf::fft2d_f64(v, nx, ny, 1, true, twx64, twy64, c->stream));
        }
        HIPCHK(c, msf::launch_to_f32(v, M, pl.Psi0.as<float2>(), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    pl.spec_id = sp->id;
    pl.optics = *o;
    pl.ncp = ncp;
    pl.valid = true;

    const float2 *twx, *twy, *twx_nat, *twy_nat, *twx_alt, *twy_alt;
    if (int r = twiddles(c, nx, &twx, nullptr, &twx_nat)) return r;
    if (int r = twiddles(c, ny, &twy, nullptr, &twy_nat)) return r;
    if (int r = alt_twiddles(c, nx, &twx_alt)) return r;
    if (int r = alt_twiddles(c, ny, &twy_alt)) return r;
    // radix plans as vx + 2 vy (per axis 0 primary, 1 alternate): configured
    // (1 primary, 2 alternate where the length has one) or by the rule
    const int all_alt = (twx_alt ? 1 : 0) + (twy_alt ? 2 : 0);
    const int plan_rule = c->cfg.fft_plan == 1   ? 0
                          : c->cfg.fft_plan == 2 ? all_alt
                                                 : fft_plan_rule(nx) + 2 * fft_plan_rule(ny);
    auto vx_of = [](int pv) { return pv & 1; };
    auto vy_of = [](int pv) { return (pv >> 1) & 1; };

    // batch depth: the fewest batches whose wave buffer stays within 96 MiB,
    // spread evenly over the raster.  The buffer, the t and u0 slices and the
    // partial sums then stay resident in the 256 MB Infinity Cache across the
    // two passes of a slice (512^2: 64 positions per batch, measured best of
    // 64..256).
    int64_t B = c->cfg.batch;
    if (B == 0) {
        const int64_t per_pos = plane * (int64_t)sizeof(float2);
        int64_t bmax = std::max<int64_t>(1, (96ll << 20) / per_pos);
        if (bmax >= 32) bmax -= bmax % 16; // whole groups of 16 positions per launch
        const int64_t nbat = (P + bmax - 1) / bmax;
        B = (P + nbat - 1) / nbat;
        if (B > 16 && B % 16) B = std::min<int64_t>(bmax, B + 16 - B % 16);
    }
    B = std::max<int64_t>(1, std::min<int64_t>(B, P));
    int col_tile = c->cfg.col_tile;
    int chains_cfg = c->cfg.chains;

    // Autotuner (opt-in): with batch, column tile and chains all automatic, time one
    // step through every slice of this job's own specimen and shifts for each
    // candidate (batch depth around the rule's, column tile, one or two
    // chains, and with fft_plan automatic the primary and alternate radix
    // plans) on trial buffers once per shape, and keep the fastest per
    // position-slice (the faster of two timings each) if it beats the rule
    // above by more than 3 %.  Candidates of one radix plan give bit-identical
    // results; the two plans differ by FP32 rounding.
    int tuned_state = 0;
    int plan_v = plan_rule;  // radix plans, vx + 2 vy
    const bool prof0 = c->cfg.profile_kernels != 0;
    if (c->cfg.autotune && c->cfg.batch == 0 && c->cfg.col_tile == 0 && c->cfg.chains == 0 &&
        !prof0 && !out.exit && P >= 2) {
        const std::array<int, 4> key = {nx, ny, ncp, (int)std::min<int64_t>(P, 256)};
        auto it = c->tuned.find(key);
        if (it == c->tuned.end()) {
            const int64_t per_pos = plane * (int64_t)sizeof(float2);
            const int64_t bcap = std::max<int64_t>(1, (384ll << 20) / (2 * per_pos));
            // around the rule's depth: B/2, B, 3B/2, 2B in whole groups of 16
            std::vector<int64_t> bs = {B};
            for (int64_t b : {B / 2, (3 * B) / 2, 2 * B}) {
                if (B >= 16) b = std::max<int64_t>(16, b - b % 16);
                if (b >= 1 && b <= P && b <= bcap && std::find(bs.begin(), bs.end(), b) == bs.end())
                    bs.push_back(b);
            }
            const int64_t bmx = *std::max_element(bs.begin(), bs.end());
            const int64_t ns = std::min<int64_t>(P, 2 * bmx);
            DevBuf tpsi, tsh;
            HIPCHK(c, tpsi.alloc((size_t)2 * bmx * per_pos));
            HIPCHK(c, hipMemsetAsync(tpsi.p, 0, tpsi.bytes, c->stream));
            std::vector<int2> sh(shifts.begin(), shifts.begin() + ns);
            HIPCHK(c, upload(tsh, sh, c->stream));
            if (!c->stream2) {
                HIPCHK(c, hipStreamCreate(&c->stream2));
                for (hipEvent_t* e : {&c->ev_fork, &c->ev_join, &c->ev_fem[0], &c->ev_fem[1]})
                    HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
            }
            hipEvent_t e0 = nullptr, e1 = nullptr;
            HIPCHK(c, hipEventCreate(&e0));
            HIPCHK(c, hipEventCreate(&e1));
            msf::RowPassArgs tr{};
            tr.nx = nx; tr.ncp = ncp; tr.ka = ka; tr.kb = kb;
            tr.inv_s = (float)(1.0 / d.sqrt_nxny);
            tr.prune = msf::row_prunable(ny, ka, kb) ? 1 : 0;
            msf::ColPassArgs tc{};
            tc.ny = ny; tc.ncp = ncp; tc.ka = ka; tc.kb = kb; tc.keep = sp->dkeep.as<int2>();
            tc.psi0 = pl.Psi0.as<float2>(); tc.twx_nat = twx_nat; tc.twy = twy_nat;
            tc.prune = msf::col_prunable(nx, keep.ab.data()) ? 1 : 0;
            // one trial: a steady-state step through every slice of the stack
            // (so t and u0 stream as in the job) on `ch` chains of b positions
            const int T = S;
            auto trial = [&](int64_t b, int w, int ch, int fp, float& ms) -> int {
                hipStream_t st[2] = {c->stream, c->stream2};
                tr.plan = vy_of(fp);
                tr.tw = tr.plan ? twy_alt : twy;
                tc.plan = vx_of(fp);
                tc.tw = tc.plan ? twx_alt : twx;
                auto steps = [&](int count) -> int {
                    if (ch == 2) {
                        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
                        HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
                    }
                    for (int k = 0; k < ch; ++k) {
                        tr.psi = tc.psi = tpsi.as<float2>() + (int64_t)k * bmx * plane;
                        tr.shift = tc.shift = tsh.as<int2>() + k * b;
                        tr.batch = tc.batch = (int)b;
                        tc.tile = w;
                        for (int n = 1; n <= count; ++n) {
                            const int sl = n % S;
                            tr.t = sp->t32.as<float2>() + (int64_t)sl * M;
                            tr.u0 = pl.U0.as<float2>() + (int64_t)sl * 2 * M;
                            tc.px = sp->px.as<float2>() + (int64_t)sl * nx;
                            tc.py = sp->py.as<float2>() + (int64_t)sl * ny;
                            HIPCHK(c, msf::launch_row_pass(ny, false, tmode, tr, st[k]));
                            HIPCHK(c, msf::launch_col_pass(nx, false, tc, st[k]));
                        }
                    }
                    if (ch == 2) {
                        HIPCHK(c, hipEventRecord(c->ev_join, c->stream2));
                        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
                    }
                    return MSF_OK;
                };
                if (int r = steps(1)) return r; // warm-up
                ms = 0.f;
                for (int rep = 0; rep < 2; ++rep) { // the faster of two timings
                    HIPCHK(c, hipEventRecord(e0, c->stream));
                    if (int r = steps(T)) return r;
                    HIPCHK(c, hipEventRecord(e1, c->stream));
                    HIPCHK(c, hipEventSynchronize(e1));
                    float t = 0.f;
                    HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
                    if (rep == 0 || t < ms) ms = t;
                }
                return MSF_OK;
            };
            const int tile0 = msf::col_tile_default(nx, vx_of(plan_rule));
            const int def_ch = (B * plane >= (4ll << 20) && (P + B - 1) / B >= 2) ? 2 : 1;
            double best = 0.0, def_cost = 0.0;
            std::array<int, 4> pick = {(int)B, tile0, def_ch, plan_rule};
            // plans: the rule's, then all primary and all alternate
            std::vector<int> plans = {plan_rule};
            if (c->cfg.fft_plan == 0)
                for (int pv : {0, all_alt})
                    if (std::find(plans.begin(), plans.end(), pv) == plans.end()) plans.push_back(pv);
            int rc = MSF_OK;
            {
                float ms = 0.f; // bring clocks and caches to steady state first
                rc = trial(B, tile0, 1, plan_rule, ms);
            }
            for (int fp : plans)
                for (int64_t b : bs)
                    for (int w : {4, 8, 16})
                        for (int ch : {1, 2}) {
                            if (rc || (ch == 2 && 2 * b > ns)) continue;
                            float ms = 0.f;
                            if ((rc = trial(b, w, ch, fp, ms))) break;
                            const double cost = ms / ((double)b * ch * T);
                            if (b == B && w == tile0 && ch == def_ch && fp == plan_rule) def_cost = cost;
                            if (best == 0.0 || cost < best) {
                                best = cost;
                                pick = {(int)b, w, ch, fp};
                            }
                        }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (rc) return rc;
            if (def_cost > 0.0 && best > def_cost / 1.03) pick = {(int)B, tile0, def_ch, plan_rule};
            it = c->tuned.emplace(key, pick).first;
            tuned_state = 1;
        } else {
            tuned_state = 2;
        }
        B = std::min<int64_t>(it->second[0], P);
        col_tile = it->second[1];
        chains_cfg = it->second[2];
        plan_v = it->second[3];
    }

    // work buffers, grown as needed and kept for the next run
    // Two chains of batches in flight on two streams: the row passes (VALU
    // bound) of one chain overlap the column passes (memory bound) of the
    // other.  Automatic: two chains, launched on the streams, once a launch
    // covers >= 4M pixels (512^2: +4 % over one chain, graph or not; the
    // runtime replays a two-branch graph's branches one after the other);
    // below that one chain replayed as a graph, where launch cost matters.
    // Profiled and exit-wave runs use one chain.
    const bool prof = c->cfg.profile_kernels != 0;
    const int64_t n_batches = (P + B - 1) / B;
    const bool big_launch = B * plane >= (4ll << 20);
    const int chains_req = chains_cfg ? chains_cfg : (big_launch ? 2 : 1);
    const int chains = (chains_req == 2 && n_batches >= 2 && !prof && !out.exit) ? 2 : 1;
    const bool auto_eager = c->cfg.chains == 0 && chains == 2;
    if (B > pl.cap_b || pl.psi.bytes < (size_t)B * plane * sizeof(float2)) {
        HIPCHK(c, pl.psi.alloc((size_t)B * plane * sizeof(float2)));
        HIPCHK(c, pl.part.alloc((size_t)B * G * sizeof(double)));
        HIPCHK(c, pl.segp.alloc((size_t)B * std::max(pl.n_tiles, 1) * 64 * sizeof(double)));
        HIPCHK(c, hipMemsetAsync(pl.psi.p, 0, pl.psi.bytes, c->stream));
        pl.psi2.release();
        pl.mtf_full.release();
        pl.mtf_full2.release();
        pl.cap_b = B;
        pl.exit.release();
    }
    if (chains == 2 && pl.psi2.bytes < pl.psi.bytes) {
        HIPCHK(c, pl.psi2.alloc(pl.psi.bytes));
        HIPCHK(c, pl.part2.alloc(pl.part.bytes));
        HIPCHK(c, pl.segp2.alloc(pl.segp.bytes));
        HIPCHK(c, hipMemsetAsync(pl.psi2.p, 0, pl.psi2.bytes, c->stream));
    }
    if (chains == 2 && !c->stream2) {
        HIPCHK(c, hipStreamCreate(&c->stream2)); // default priority, as the first stream
        for (hipEvent_t* e : {&c->ev_fork, &c->ev_join, &c->ev_fem[0], &c->ev_fem[1]})
            HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    if (P > pl.cap_p) {
        HIPCHK(c, pl.adf.alloc((size_t)P * sizeof(double)));
        HIPCHK(c, pl.shift.alloc((size_t)P * sizeof(int2)));
        pl.cap_p = P;
    }
    if (pl.state.bytes < (size_t)std::max(nbins, 1) * 3 * sizeof(double)) {
        HIPCHK(c, pl.state.alloc((size_t)std::max(nbins, 1) * 3 * sizeof(double)));
        HIPCHK(c, pl.femv.alloc((size_t)std::max(nbins, 1) * sizeof(double)));
        HIPCHK(c, pl.vsum.alloc((size_t)std::max(nbins, 1) * sizeof(double)));
    }
    if (want_mtf && pl.mtf_full.bytes < (size_t)B * M * sizeof(double2))
        HIPCHK(c, pl.mtf_full.alloc((size_t)B * M * sizeof(double2)));
    if (want_mtf && chains == 2 && pl.mtf_full2.bytes < (size_t)B * M * sizeof(double2))
        HIPCHK(c, pl.mtf_full2.alloc((size_t)B * M * sizeof(double2)));
    const double2 *twx64 = nullptr, *twy64 = nullptr;
    if (want_mtf) {
        if (int r = twiddles(c, nx, nullptr, &twx64)) return r;
        if (int r = twiddles(c, ny, nullptr, &twy64)) return r;
    }
    const bool want_each = want_corr && out.corr_each && nbins * nphi > 0;
    if (want_each && pl.ceach.bytes < (size_t)P * nbins * nphi * sizeof(double))
        HIPCHK(c, pl.ceach.alloc((size_t)P * nbins * nphi * sizeof(double)));
    if (want_corr && pl.csum.bytes < ncorr * sizeof(double)) {
        HIPCHK(c, pl.csum.alloc(ncorr * sizeof(double)));
        HIPCHK(c, pl.csqr.alloc(ncorr * sizeof(double)));
    }
    if (want_pix && pl.pm1.bytes < (size_t)plane * sizeof(double)) {
        HIPCHK(c, pl.pm1.alloc((size_t)plane * sizeof(double)));
        HIPCHK(c, pl.pm2.alloc((size_t)plane * sizeof(double)));
    }
    // accumulators start at zero (outside any captured graph)
    HIPCHK(c, hipMemsetAsync(pl.state.p, 0, (size_t)std::max(nbins, 1) * 3 * sizeof(double), c->stream));
    if (want_d2)
        HIPCHK(c, hipMemsetAsync(pl.vsum.p, 0, (size_t)std::max(nbins, 1) * sizeof(double), c->stream));
    if (want_corr) {
        HIPCHK(c, hipMemsetAsync(pl.csum.p, 0, ncorr * sizeof(double), c->stream));
        HIPCHK(c, hipMemsetAsync(pl.csqr.p, 0, ncorr * sizeof(double), c->stream));
    }
    if (want_pix) {
        HIPCHK(c, hipMemsetAsync(pl.pm1.p, 0, (size_t)plane * sizeof(double), c->stream));
        HIPCHK(c, hipMemsetAsync(pl.pm2.p, 0, (size_t)plane * sizeof(double), c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(pl.shift.p, shifts.data(), (size_t)P * sizeof(int2),
                             hipMemcpyHostToDevice, c->stream));
    if (out.exit && pl.exit.bytes < (size_t)B * M * sizeof(float2))
        HIPCHK(c, pl.exit.alloc((size_t)B * M * sizeof(float2)));
    const DevBuf& dkeep = sp->dkeep;
    DevBuf &dshift = pl.shift, &dcnt = pl.cnt, &dzero = pl.zero, &dpsi = pl.psi, &dpart = pl.part,
           &dadf = pl.adf, &dstate = pl.state, &dfemv = pl.femv, &dexit = pl.exit,
           &dU0 = pl.U0, &dPsi0 = pl.Psi0;

    // events for per-kernel timing
    enum { K_ROW0, K_ROW, K_COL, K_COLL, K_RED, K_NUM };
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_kind;
    auto ev_begin = [&](int kind) -> hipError_t {
        if (!prof) return hipSuccess;
        hipEvent_t a, b;
        hipError_t e = hipEventCreate(&a);
        if (e == hipSuccess) e = hipEventCreate(&b);
        if (e != hipSuccess) return e;
        ev.push_back(a);
        ev.push_back(b);
        ev_kind.push_back(kind);
        return hipEventRecord(a, c->stream);
    };
    auto ev_end = [&]() -> hipError_t { return prof ? hipEventRecord(ev.back(), c->stream) : hipSuccess; };
    // an event pair stamped by the kernel's own dispatch (slice passes)
    hipEvent_t ka0 = nullptr, ka1 = nullptr;
    auto ev_kernel = [&](int kind) -> hipError_t {
        ka0 = ka1 = nullptr;
        if (!prof) return hipSuccess;
        hipError_t e = hipEventCreate(&ka0);
        if (e == hipSuccess) e = hipEventCreate(&ka1);
        if (e != hipSuccess) return e;
        ev.push_back(ka0);
        ev.push_back(ka1);
        ev_kind.push_back(kind);
        return hipSuccess;
    };

    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double t_loop = now_ms();
    msf::RowPassArgs ra{};
    ra.nx = nx;
    ra.ncp = ncp;
    ra.ka = ka;
    ra.kb = kb;
    ra.plan = vy_of(plan_v);
    ra.tw = ra.plan ? twy_alt : twy;
    ra.inv_s = (float)(1.0 / d.sqrt_nxny);
    ra.prune = msf::row_prunable(ny, ka, kb) ? 1 : 0;
    msf::ColPassArgs ca{};
    ca.ny = ny;
    ca.ncp = ncp;
    ca.ka = ka;
    ca.kb = kb;
    ca.keep = dkeep.as<int2>();
    ca.plan = vx_of(plan_v);
    ca.tw = ca.plan ? twx_alt : twx;
    ca.psi0 = dPsi0.as<float2>();
    ca.twx_nat = twx_nat;
    ca.twy = twy_nat;
    ca.tile = col_tile;
    ca.prune = msf::col_prunable(nx, keep.ab.data()) ? 1 : 0;
    const float2* t32 = sp->t32.as<float2>();
    const float2* px = sp->px.as<float2>();
    const float2* py = sp->py.as<float2>();
    // every launch of the job (all batches, reductions, FEM statistics);
    // issued directly or captured once into a graph
    auto enqueue = [&]() -> int {
    hipStream_t streams[2] = {c->stream, c->stream2};
    if (chains == 2) { // fork: the second stream starts after everything issued so far
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t b0 = 0; b0 < P; b0 += B, ++k) {
        const int nb = (int)std::min<int64_t>(B, P - b0);
        const int ch = (int)(k % chains);
        hipStream_t s = streams[ch];
        float2* bpsi = (ch ? pl.psi2 : dpsi).as<float2>();
        double* bpart = (ch ? pl.part2 : dpart).as<double>();
        double* bsegp = (ch ? pl.segp2 : pl.segp).as<double>();
        ra.psi = bpsi;
        ra.batch = nb;
        ra.shift = dshift.as<int2>() + b0;
        ca.psi = bpsi;
        ca.batch = nb;
        ca.shift = dshift.as<int2>() + b0;
        for (int n = 0; n < S; ++n) {
            ra.t = t32 + (int64_t)n * M;
            ra.u0 = dU0.as<float2>() + (int64_t)n * 2 * M;
            HIPCHK(c, ev_kernel(n == 0 ? K_ROW0 : K_ROW));
            HIPCHK(c, msf::launch_row_pass(ny, n == 0, tmode, ra, s, ka0, ka1));
            ca.px = px + (int64_t)n * nx;
            ca.py = py + (int64_t)n * ny;
            HIPCHK(c, ev_kernel(n == S - 1 ? K_COLL : K_COL));
            HIPCHK(c, msf::launch_col_pass(nx, n == S - 1, ca, s, ka0, ka1));
        }
        HIPCHK(c, ev_begin(K_RED));
        HIPCHK(c, msf::launch_reduce_segments(bpsi, plane, pl.seg_pix.as<int32_t>(),
                                              pl.tile_off.as<int32_t>(), pl.seg_kind.as<int32_t>(),
                                              pl.n_tiles, nb, bsegp, s));
        HIPCHK(c, msf::launch_combine_segments(bsegp, pl.n_tiles * 64,
                                               pl.seg_of_group.as<int32_t>(), G, nb, bpart, s));
        HIPCHK(c, ev_end());
        if (want_adf)
            HIPCHK(c, msf::launch_collect_adf(bpart, G, nb, dadf.as<double>() + b0, s));
        if (want_mtf) {
            // adfstem.cpp:2272-2385 after the ADF sum: |psi| outside the
            // aperture, BACKWARD FFT, x MTF / s (in-place update), FORWARD
            // FFT, / s, bw_limit; then the group sums again for the FEM
            double2* full = (ch ? pl.mtf_full2 : pl.mtf_full).as<double2>();
            HIPCHK(c, msf::launch_mtf_magnitude(bpsi, nx, ny, ncp, ka, kb,
                                                pl.mtf_out.as<uint8_t>(), nb, full, s));
            HIPCHK(c, msf::fft2d_f64(full, nx, ny, nb, true, twx64, twy64, s));
            HIPCHK(c, msf::launch_transmit_f64(full, pl.mtf2d.as<double2>(), M, nb, 0, d.sqrt_nxny,
                                               s));
            HIPCHK(c, msf::fft2d_f64(full, nx, ny, nb, false, twx64, twy64, s));
            HIPCHK(c, msf::launch_mtf_compact(full, nx, ny, ncp, ka, kb, dkeep.as<int2>(),
                                              d.sqrt_nxny, nb, bpsi, s));
            HIPCHK(c, msf::launch_reduce_segments(bpsi, plane, pl.seg_pix.as<int32_t>(),
                                                  pl.tile_off.as<int32_t>(),
                                                  pl.seg_kind.as<int32_t>(), pl.n_tiles, nb, bsegp,
                                                  s));
            HIPCHK(c, msf::launch_combine_segments(bsegp, pl.n_tiles * 64,
                                                   pl.seg_of_group.as<int32_t>(), G, nb, bpart, s));
        }
        // the FEM accumulators are shared: batch k updates them after batch k-1
        if (chains == 2 && k > 0) HIPCHK(c, hipStreamWaitEvent(s, c->ev_fem[(k - 1) % 2], 0));
        if (want_d1)
            HIPCHK(c, msf::launch_fem_welford(bpart, G, nb, nbins, dcnt.as<int32_t>(),
                                              dzero.as<uint8_t>(), dstate.as<double>(), s));
        if (want_d2)
            HIPCHK(c, msf::launch_fem_vbar(bpart, G, nb, nbins, pl.hg.fem_abs, pl.hg.fem_sqr,
                                           dcnt.as<int32_t>(), dzero.as<uint8_t>(),
                                           pl.vsum.as<double>(), s));
        if (want_corr)
            HIPCHK(c, msf::launch_fem_correlograph(bpart, G, nb, nbins, nphi, pl.hg.fem_abs,
                                                   pl.hg.corr, dcnt.as<int32_t>(),
                                                   dzero.as<uint8_t>(), pl.csum.as<double>(),
                                                   pl.csqr.as<double>(),
                                                   want_each ? pl.ceach.as<double>() +
                                                                   b0 * nbins * nphi
                                                             : nullptr,
                                                   s));
        if (want_pix)
            HIPCHK(c, msf::launch_pixel_moments(bpsi, plane, nb, pl.pm1.as<double>(),
                                                pl.pm2.as<double>(), s));
        if (chains == 2) HIPCHK(c, hipEventRecord(c->ev_fem[k % 2], s));
        if (out.exit) {
            HIPCHK(c, msf::launch_expand_exit(bpsi, nx, ny, ncp, ka, kb, nb, dexit.as<float2>(), s));
            HIPCHK(c, hipMemcpyAsync(out.exit + (size_t)b0 * nx * ny * 2, dexit.p,
                                     (size_t)nb * nx * ny * sizeof(float2), hipMemcpyDeviceToHost,
                                     s));
        }
    }
    if (chains == 2) { // join: the first stream waits for everything on the second
        HIPCHK(c, hipEventRecord(c->ev_join, c->stream2));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    if (want_d1)
        HIPCHK(c, msf::launch_fem_finalize(dstate.as<double>(), nbins, dfemv.as<double>(), c->stream));
    return MSF_OK;
    };

    // Graph replay: the signature lists everything the captured kernel
    // arguments depend on (job shape, device addresses); any change
    // re-captures.  Profiled runs and exit-wave runs (host copies per batch)
    // launch directly.
    int graph_state = 0;
    if (!prof && !out.exit && !c->cfg.no_graph && !auto_eager) {
        const std::vector<uint64_t> sig = {
            sp->id, (uint64_t)P, (uint64_t)B, (uint64_t)tmode, (uint64_t)want_adf,
            (uint64_t)want_d1 | (uint64_t)want_d2 << 1 | (uint64_t)want_pix << 2 |
                (uint64_t)want_corr << 3,
            (uint64_t)(uintptr_t)pl.csum.p, (uint64_t)(uintptr_t)pl.csqr.p, (uint64_t)nphi,
            (uint64_t)pl.hg.corr, (uint64_t)want_mtf, (uint64_t)(uintptr_t)pl.mtf2d.p,
            (uint64_t)want_each, (uint64_t)(uintptr_t)pl.ceach.p,
            (uint64_t)chains, (uint64_t)(uintptr_t)pl.psi2.p, (uint64_t)(uintptr_t)pl.part2.p,
            (uint64_t)(uintptr_t)pl.segp2.p, (uint64_t)(uintptr_t)pl.mtf_full2.p,
            (uint64_t)(uintptr_t)pl.mtf_out.p, (uint64_t)(uintptr_t)pl.mtf_full.p,
            (uint64_t)(uintptr_t)twx64, (uint64_t)(uintptr_t)twy64,
            (uint64_t)(uintptr_t)pl.vsum.p, (uint64_t)(uintptr_t)pl.pm1.p,
            (uint64_t)(uintptr_t)pl.pm2.p, (uint64_t)pl.hg.fem_abs, (uint64_t)pl.hg.fem_sqr,
            (uint64_t)plane, (uint64_t)(uintptr_t)dpsi.p, (uint64_t)(uintptr_t)dpart.p,
            (uint64_t)(uintptr_t)pl.segp.p, (uint64_t)(uintptr_t)dadf.p,
            (uint64_t)(uintptr_t)dshift.p, (uint64_t)(uintptr_t)dstate.p,
            (uint64_t)(uintptr_t)dfemv.p, (uint64_t)(uintptr_t)dU0.p, (uint64_t)(uintptr_t)dPsi0.p,
            (uint64_t)(uintptr_t)pl.seg_pix.p, (uint64_t)(uintptr_t)pl.tile_off.p,
            (uint64_t)(uintptr_t)pl.seg_kind.p, (uint64_t)(uintptr_t)pl.seg_of_group.p,
            (uint64_t)(uintptr_t)dcnt.p, (uint64_t)(uintptr_t)dzero.p,
            (uint64_t)(uintptr_t)ca.tw, (uint64_t)(uintptr_t)ra.tw, (uint64_t)(uintptr_t)twx_nat,
            (uint64_t)(uintptr_t)twy_nat, (uint64_t)ncp, (uint64_t)G, (uint64_t)nbins,
            (uint64_t)pl.n_tiles, (uint64_t)col_tile, (uint64_t)plan_v};
        graph_state = 2;
        bool captured = true;
        if (!pl.graph || pl.graph_sig != sig) {
            pl.drop_graph();
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const int r = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (r != MSF_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                // A capture can be invalidated from outside, e.g. by another
                // thread's allocation while it runs; nothing captured has
                // executed, so one chain simply launches on the stream.
                if (chains != 1) return r != MSF_OK ? r : fail(c, MSF_ERR_DEVICE, hipGetErrorString(e));
                (void)hipGetLastError();
                captured = false;
            } else {
                const hipError_t ei = hipGraphInstantiate(&pl.graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) pl.graph = nullptr;
                HIPCHK(c, ei);
                pl.graph_sig = sig;
                graph_state = 1;
            }
        }
        if (captured) {
            HIPCHK(c, hipGraphLaunch(pl.graph, c->stream));
        } else {
            graph_state = 0;
            if (int r = enqueue()) return r;
        }
    } else {
        if (int r = enqueue()) return r;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double t_end = now_ms();

    if (want_adf && out.adf)
        HIPCHK(c, hipMemcpy(out.adf, dadf.p, (size_t)P * sizeof(double), hipMemcpyDeviceToHost));
    if (want_d1 && nbins > 0)
        HIPCHK(c, hipMemcpy(out.fem, dfemv.p, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
    if (want_d2 && nbins > 0) { // adfstem.cpp:2896-2915: the mean over positions
        std::vector<double> vs(nbins);
        HIPCHK(c, hipMemcpy(vs.data(), pl.vsum.p, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < nbins; ++k) out.vbar_r[k] = vs[k] / (double)P;
    }
    if (want_pix && nbins > 0) {
        std::vector<double> s1((size_t)plane), s2((size_t)plane);
        HIPCHK(c, hipMemcpy(s1.data(), pl.pm1.p, (size_t)plane * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(s2.data(), pl.pm2.p, (size_t)plane * sizeof(double), hipMemcpyDeviceToHost));
        fem_pixel_modes(d, pl.hg, s1, s2, P, o->alpha_max, out);
        if (out.mean_abs || out.mean_sqr || out.var2d)
            fem_pixel_images(d, keep, ncp, s1, s2, P, o->alpha_max, out);
    }
    if (want_corr && nbins * nphi > 0) { // adfstem.cpp:3620-3632: the mean over positions
        const size_t n = (size_t)nbins * nphi;
        std::vector<double> cs(n);
        HIPCHK(c, hipMemcpy(cs.data(), pl.csum.p, n * sizeof(double), hipMemcpyDeviceToHost));
        if (out.corr_avg)
            for (size_t k = 0; k < n; ++k) out.corr_avg[k] = cs[k] / (double)P;
        if (out.corr_sqr)
            HIPCHK(c, hipMemcpy(out.corr_sqr, pl.csqr.p, n * sizeof(double), hipMemcpyDeviceToHost));
        if (want_each)
            HIPCHK(c, hipMemcpy(out.corr_each, pl.ceach.p, (size_t)P * n * sizeof(double),
                                hipMemcpyDeviceToHost));
    }

    msf_stats st{};
    st.setup_ms = t_loop - t_start;
    st.slice_loop_ms = t_end - t_loop;
    st.total_ms = now_ms() - t_start;
    st.probe_slices = (uint64_t)P * (uint64_t)S;
    st.batch = (uint32_t)B;
    st.n_batches = (uint32_t)n_batches;
    // compulsory traffic of a full-batch pass: read + write of the compact wave
    st.row_pass_bytes = 2.0 * (double)B * nx * (ka + kb) * sizeof(float2);
    st.col_pass_bytes = st.row_pass_bytes;
    st.graph = graph_state;
    st.chains = chains;
    st.col_tile = col_tile ? col_tile : msf::col_tile_default(nx, ca.plan);
    st.autotuned = tuned_state;
    st.row_fft_plan = 1 + ra.plan;
    st.col_fft_plan = 1 + ca.plan;
    if (prof) {
        double sum[K_NUM] = {0}, cnt[K_NUM] = {0};
        for (size_t k = 0; k < ev_kind.size(); ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) == hipSuccess) {
                sum[ev_kind[k]] += ms;
                cnt[ev_kind[k]] += 1;
            }
        }
        auto avg = [&](int k) { return cnt[k] > 0 ? 1000.0 * sum[k] / cnt[k] : 0.0; };
        st.first_row_pass_us = avg(K_ROW0);
        st.row_pass_us = avg(K_ROW);
        st.col_pass_us = avg(K_COL);
        st.last_col_pass_us = avg(K_COLL);
        st.reduce_us = avg(K_RED);
        st.row_pass_count = (uint64_t)cnt[K_ROW];
        st.col_pass_count = (uint64_t)cnt[K_COL];
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    c->stats = st;
    return MSF_OK;
}

} // namespace

extern "C" {

int msf_create(msf_ctx** out, int device)
{
    try {
        if (!out) return MSF_ERR_INVALID;
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MSF_ERR_DEVICE;
        if (hipSetDevice(device) != hipSuccess) return MSF_ERR_DEVICE;
        msf_ctx* c = new msf_ctx;
        c->device = device;
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return MSF_ERR_DEVICE;
        }
        *out = c;
        return MSF_OK;
    }
    MSF_CATCH_ALL(nullptr)
}

int msf_destroy(msf_ctx* c)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        std::unique_lock<std::mutex> serve_lk(c->serve_mu);
        if (c->serve) { // finish the queued jobs, then stop the workers
            {
                std::lock_guard<std::mutex> lk(c->serve->m);
                c->serve->stop = true;
            }
            c->serve->work_cv.notify_all();
            for (std::thread& w : c->serve->workers)
                if (w.joinable()) w.join();
            for (msf_ctx* sc : c->serve->slot_ctx) msf_destroy(sc);
            c->serve.reset();
        }
        serve_lk.unlock();
        {
            std::lock_guard<std::recursive_mutex> lk(c->mu);
            (void)hipSetDevice(c->device);
            (void)hipStreamSynchronize(c->stream);
            // an instantiated graph refers to the stream it was launched on:
            // destroy it (and every buffer its kernels use) before the stream
            c->plan.drop_graph();
            // specimens may outlive their context: detach them so that a later
            // msf_specimen_free only releases their own device memory
            for (msf_specimen* sp : c->specimens) sp->ctx = nullptr;
            c->specimens.clear();
            c->tw32.clear();
            c->tw64.clear();
            c->twnat.clear();
            c->tw32alt.clear();
            for (hipEvent_t e : {c->ev_fork, c->ev_join, c->ev_fem[0], c->ev_fem[1]})
                if (e) (void)hipEventDestroy(e);
            if (c->stream2) {
                (void)hipStreamSynchronize(c->stream2);
                (void)hipStreamDestroy(c->stream2);
            }
            (void)hipStreamDestroy(c->stream);
        }
        delete c;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

const char* msf_last_error(const msf_ctx* c) { return c ? c->err.c_str() : "no context"; }

int msf_set_config(msf_ctx* c, const msf_config* cfg)
{
    try {
        if (!c || !cfg) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (cfg->transmit_mode != MSF_TRANSMIT_REFERENCE && cfg->transmit_mode != MSF_TRANSMIT_EXACT)
            return fail(c, MSF_ERR_INVALID, "unknown transmit_mode");
        if (cfg->col_tile != 0 && cfg->col_tile != 4 && cfg->col_tile != 8 && cfg->col_tile != 16)
            return fail(c, MSF_ERR_INVALID, "col_tile must be 0, 4, 8 or 16");
        if (cfg->chains < 0 || cfg->chains > 2) return fail(c, MSF_ERR_INVALID, "chains must be 0, 1 or 2");
        if (cfg->fft_plan < 0 || cfg->fft_plan > 2) return fail(c, MSF_ERR_INVALID, "fft_plan must be 0, 1 or 2");
        c->cfg = *cfg;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_get_stats(const msf_ctx* c, msf_stats* out)
{
    try {
        if (!c || !out) return MSF_ERR_INVALID;
        MSF_LOCK(const_cast<msf_ctx*>(c));
        *out = c->stats;
        return MSF_OK;
    }
    MSF_CATCH_ALL(nullptr)
}

int msf_supported_length(uint32_t n) { return msf::slice_lengths_ok((int)n) ? 1 : 0; }

int msf_wavelength(double voltage, double* lambda, double* gamma)
{
    try {
        if (!(voltage > 0.0) || !lambda || !gamma) return MSF_ERR_INVALID;
        msf::wavelength(voltage, *lambda, *gamma);
        return MSF_OK;
    }
    MSF_CATCH_ALL(nullptr)
}

double msf_scherzer_defocus(int32_t corrected, double lambda, double cs)
{
    if (corrected) return -0.983 * std::pow(msf::ref_pow2(lambda) * cs, 1.0 / 3.0); // scherzer.cpp:141-148
    return 0.87 * std::sqrt(cs * lambda);                                          // scherzer.cpp:61-69
}

double msf_scherzer_alphamax(int32_t corrected, double lambda, double cs)
{
    if (corrected) return 1.513 * std::pow(lambda / cs, 1.0 / 6.0); // scherzer.cpp:150-157
    return 1.34 * std::pow(lambda / cs, 0.25);                      // scherzer.cpp:71-79
}

double msf_scherzer_cs3(double lambda, double cs5)
{
    return -2.289 * std::pow(lambda * msf::ref_pow2(cs5), 1.0 / 3.0); // scherzer.cpp:131-139
}

int msf_specimen_build(msf_ctx* c, const msf_grid* grid, const double* q, const uint32_t* Z,
                       uint32_t n, double min_slice, msf_specimen** out)
{
    try {
        if (!c || !out) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        *out = nullptr;
        if (int r = check_grid(c, grid)) return r;
        if (n > 0 && (!q || !Z)) return fail(c, MSF_ERR_INVALID, "atom arrays are NULL");
        if (!(min_slice >= 0.0)) return fail(c, MSF_ERR_INVALID, "minslice must be >= 0");
        HIPCHK(c, hipSetDevice(c->device));
        const int nx = (int)grid->nx, ny = (int)grid->ny;
        msf_specimen* sp = new msf_specimen;
        sp->ctx = c;
        sp->grid = *grid;
        sp->dom = msf::make_domain(nx, ny, grid->xperiod, grid->yperiod, grid->xmin, grid->ymin,
                                   grid->voltage);
        const msf::Domain& d = sp->dom;
        sp->id = g_specimen_ids.fetch_add(1);
        sp->keep = msf::slice_keep_ranges(d);
        if (!sp->keep.ok) {
            delete sp;
            return fail(c, MSF_ERR_INVALID, "band limit is not a disk in FFT order");
        }
        if (hipError_t e = upload(sp->dkeep, to_int2(sp->keep), c->stream)) {
            delete sp;
            return fail(c, MSF_ERR_DEVICE, hipGetErrorString(e));
        }

        // species present, ascending Z (tem.cpp:944-972)
        std::vector<int> species_of(104, -1), zlist;
        for (uint32_t a = 0; a < n; ++a) {
            if (Z[a] < 1 || Z[a] > 103) {
                delete sp;
                return fail(c, MSF_ERR_INVALID, "atomic number outside 1..103 (tem_io.cpp:4768-4781)");
            }
            species_of[Z[a]] = 0;
        }
        for (int z = 1; z <= 103; ++z)
            if (species_of[z] == 0) {
                species_of[z] = (int)zlist.size();
                zlist.push_back(z);
            }
        const int nsp = (int)zlist.size();

        // auto-slicing and per-atom nearest-pixel shifts
        msf::Slices sl = msf::autoslice(q, n, grid->zmin, grid->zperiod, min_slice);
        const int S = (int)sl.lower.size();
        sp->n_slices = S;
        sp->lower = sl.lower;
        sp->thick = sl.thick;
        sp->counts.resize(S);
        std::vector<int4> atoms; // (slice, species, sx, sy)
        atoms.reserve(n);
        for (int s = 0; s < S; ++s) {
            sp->counts[s] = (uint32_t)sl.atoms[s].size();
            for (int32_t a : sl.atoms[s]) {
                const int64_t sx = msf::shift_index(d.xx, q[3 * a], 0);
                const int64_t sy = msf::shift_index(d.yy, q[3 * a + 1], 0);
                if (sx < 0 || sy < 0) {
                    delete sp;
                    return fail(c, MSF_ERR_DOMAIN, "atom position below the domain minimum");
                }
                atoms.push_back(int4{s, species_of[Z[a]], (int)sx, (int)sy});
            }
        }

        // propagators (slice.cpp:1293-1363), FP64 -> FP32; P_y carries 1/sqrtNxNy
        std::vector<float2> pxs((size_t)S * nx), pys((size_t)S * ny);
        std::vector<double2> pxd((size_t)S * nx), pyd((size_t)S * ny);
        for (int s = 0; s < S; ++s) {
            std::vector<double> p;
            msf::propagator(d.lambda, sl.thick[s], d.kx, d.bw_t, p);
            for (int i = 0; i < nx; ++i) {
                pxd[(size_t)s * nx + i] = double2{p[2 * i], p[2 * i + 1]};
                pxs[(size_t)s * nx + i] = float2{(float)p[2 * i], (float)p[2 * i + 1]};
            }
            msf::propagator(d.lambda, sl.thick[s], d.ky, d.bw_t, p);
            for (int j = 0; j < ny; ++j) {
                pyd[(size_t)s * ny + j] = double2{p[2 * j], p[2 * j + 1]};
                pys[(size_t)s * ny + j] = float2{(float)(p[2 * j] / d.sqrt_nxny), (float)(p[2 * j + 1] / d.sqrt_nxny)};
            }
        }
        int rc = MSF_OK;
        auto bail = [&](int code) {
            (void)hipGetLastError(); // the failure is reported here, not by a later launch
            delete sp;
            return code;
        };
        if (hipError_t e = upload(sp->px, pxs, c->stream)) return bail(fail(c, MSF_ERR_DEVICE, hipGetErrorString(e)));
        if (hipError_t e = upload(sp->py, pys, c->stream)) return bail(fail(c, MSF_ERR_DEVICE, hipGetErrorString(e)));
        if (hipError_t e = upload(sp->px64, pxd, c->stream)) return bail(fail(c, MSF_ERR_DEVICE, hipGetErrorString(e)));
        if (hipError_t e = upload(sp->py64, pyd, c->stream)) return bail(fail(c, MSF_ERR_DEVICE, hipGetErrorString(e)));

        // transmission functions, FP64 on the device, in chunks of slices
        const msf::RowRanges sup = msf::template_keep_ranges(d);
        const msf::RowRanges bw = msf::bwlimit_keep_ranges(d);
        if (!sup.ok || !bw.ok) return bail(fail(c, MSF_ERR_INVALID, "band limit is not a disk in FFT order"));
        std::vector<double> params;
        for (int z : zlist) params.insert(params.end(), kKirkland[z - 1], kKirkland[z - 1] + 12);
        const int64_t M = (int64_t)nx * ny;
        const double2 *twx, *twy;
        if ((rc = twiddles(c, nx, nullptr, &twx))) return bail(rc);
        if ((rc = twiddles(c, ny, nullptr, &twy))) return bail(rc);
        DevBuf dkx, dky, dsup, dbw, dpar, dmaps, datoms;
        hipError_t e = hipSuccess;
        if (!e) e = upload(dkx, d.kx, c->stream);
        if (!e) e = upload(dky, d.ky, c->stream);
        if (!e) e = upload(dsup, to_int2(sup), c->stream);
        if (!e) e = upload(dbw, to_int2(bw), c->stream);
        if (!e) e = upload(dpar, params, c->stream);
        if (!e) e = sp->t64.alloc((size_t)S * M * sizeof(double2));
        if (!e) e = sp->t32.alloc((size_t)S * M * sizeof(float2));
        const int64_t per_slot = (int64_t)std::max(nsp, 1) * M * (int64_t)sizeof(double2);
        const int K = (int)std::max<int64_t>(1, std::min<int64_t>(S, (256ll << 20) / per_slot));
        if (!e) e = dmaps.alloc((size_t)K * per_slot);
        if (!e) e = datoms.alloc(std::max<size_t>(atoms.size(), 1) * sizeof(int4));
        if (!e && !atoms.empty())
            e = hipMemcpyAsync(datoms.p, atoms.data(), atoms.size() * sizeof(int4), hipMemcpyHostToDevice,
                               c->stream);
        if (e) {
            (void)hipGetLastError();
            return bail(fail(c, e == hipErrorOutOfMemory ? MSF_ERR_MEMORY : MSF_ERR_DEVICE, hipGetErrorString(e)));
        }
        const double lgab = d.lambda * d.gamma * (1 / (grid->xperiod * grid->yperiod));
        size_t a0 = 0;
        for (int s0 = 0; s0 < S && !e; s0 += K) {
            const int k = std::min(K, S - s0);
            size_t a1 = a0;
            while (a1 < atoms.size() && atoms[a1].x < s0 + k) ++a1;
            // slots are relative to the chunk
            std::vector<int4> chunk(atoms.begin() + a0, atoms.begin() + a1);
            for (auto& at : chunk) at.x -= s0;
            if (!chunk.empty())
                e = hipMemcpyAsync(datoms.p, chunk.data(), chunk.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c->stream);
            double2* maps = dmaps.as<double2>();
            double2* sig = sp->t64.as<double2>() + (int64_t)s0 * M;
            if (!e) e = msf::launch_fill_f64(maps, (int64_t)k * std::max(nsp, 1) * M, c->stream);
            if (!e) e = msf::launch_scatter_atoms(maps, datoms.as<int4>(), (int)chunk.size(), nx, ny,
                                                  std::max(nsp, 1), c->stream);
            if (!e && nsp > 0) e = msf::fft2d_f64(maps, nx, ny, k * nsp, false, twx, twy, c->stream);
            if (!e) e = msf::launch_combine_species(maps, sig, k, nsp, dpar.as<double>(), dkx.as<double>(),
                                                    dky.as<double>(), nx, ny, dsup.as<int2>(), lgab,
                                                    grid->xmin, grid->ymin, c->stream);
            if (!e) e = msf::fft2d_f64(sig, nx, ny, k, true, twx, twy, c->stream);
            if (!e) e = msf::launch_exp_transmit(sig, (int64_t)k * M, 1.0 / std::sqrt((double)M), c->stream);
            if (!e) e = msf::fft2d_f64(sig, nx, ny, k, false, twx, twy, c->stream);
            if (!e) e = msf::launch_apply_keep(sig, nx, ny, k, dbw.as<int2>(), c->stream);
            if (!e) e = msf::fft2d_f64(sig, nx, ny, k, true, twx, twy, c->stream);
            if (!e) e = msf::launch_finish_transmit(sig, (int64_t)k * M, d.nxny, d.sqrt_nxny,
                                                    sp->t32.as<float2>() + (int64_t)s0 * M, c->stream);
            // the atom list buffer is reused by the next chunk
            if (!e) e = hipStreamSynchronize(c->stream);
            a0 = a1;
        }
        if (!e) e = hipStreamSynchronize(c->stream);
        if (e) {
            (void)hipGetLastError();
            return bail(fail(c, MSF_ERR_DEVICE, std::string("transmission build: ") + hipGetErrorString(e)));
        }
        c->specimens.insert(sp);
        *out = sp;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_specimen_free(msf_specimen* sp)
{
    try {
        if (!sp) return MSF_ERR_INVALID;
        if (sp->ctx) {
            MSF_LOCK(sp->ctx);
            (void)hipSetDevice(sp->ctx->device);
            (void)hipStreamSynchronize(sp->ctx->stream);
            if (sp->ctx->plan.spec_id == sp->id) {
                sp->ctx->plan.valid = false;
                sp->ctx->plan.drop_graph();
            }
            sp->ctx->specimens.erase(sp);
        }
        delete sp;
        return MSF_OK;
    }
    MSF_CATCH_ALL(sp ? sp->ctx : nullptr)
}

uint32_t msf_specimen_slices(const msf_specimen* sp) { return sp ? (uint32_t)sp->n_slices : 0; }

int msf_specimen_slice_info(const msf_specimen* sp, double* lower, double* thick, uint32_t* atoms)
{
    try {
        if (!sp) return MSF_ERR_INVALID;
        for (int s = 0; s < sp->n_slices; ++s) {
            if (lower) lower[s] = sp->lower[s];
            if (thick) thick[s] = sp->thick[s];
            if (atoms) atoms[s] = sp->counts[s];
        }
        return MSF_OK;
    }
    MSF_CATCH_ALL(sp ? sp->ctx : nullptr)
}

int msf_specimen_transmission(const msf_specimen* sp, uint32_t slice, double* t)
{
    try {
        if (!sp || !t || (int)slice >= sp->n_slices) return MSF_ERR_INVALID;
        if (!sp->ctx) return MSF_ERR_INVALID; // its context was destroyed
        MSF_LOCK(sp->ctx);
        const size_t M = (size_t)sp->dom.nx * sp->dom.ny;
        (void)hipSetDevice(sp->ctx->device);
        if (hipMemcpy(t, sp->t64.as<double2>() + slice * M, M * sizeof(double2), hipMemcpyDeviceToHost) !=
            hipSuccess)
            return fail(sp->ctx, MSF_ERR_DEVICE, "copy of the transmission function failed");
        return MSF_OK;
    }
    MSF_CATCH_ALL(sp ? sp->ctx : nullptr)
}

int msf_specimen_propagator(const msf_specimen* sp, uint32_t slice, double* px, double* py)
{
    try {
        if (!sp || (int)slice >= sp->n_slices) return MSF_ERR_INVALID;
        if (!sp->ctx) return MSF_ERR_INVALID;
        MSF_LOCK(sp->ctx);
        (void)hipSetDevice(sp->ctx->device);
        const int nx = sp->dom.nx, ny = sp->dom.ny;
        if (px && hipMemcpy(px, sp->px64.as<double2>() + (size_t)slice * nx, nx * sizeof(double2),
                            hipMemcpyDeviceToHost) != hipSuccess)
            return fail(sp->ctx, MSF_ERR_DEVICE, "copy of the propagator failed");
        if (py && hipMemcpy(py, sp->py64.as<double2>() + (size_t)slice * ny, ny * sizeof(double2),
                            hipMemcpyDeviceToHost) != hipSuccess)
            return fail(sp->ctx, MSF_ERR_DEVICE, "copy of the propagator failed");
        return MSF_OK;
    }
    MSF_CATCH_ALL(sp ? sp->ctx : nullptr)
}

int msf_probe_k(msf_ctx* c, const msf_grid* grid, const msf_optics* o, int32_t sx, int32_t sy,
                double* psi_k)
{
    try {
        if (!c || !psi_k) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (int r = check_grid(c, grid)) return r;
        if (sx < 0 || sy < 0 || sx >= (int32_t)grid->nx || sy >= (int32_t)grid->ny)
            return fail(c, MSF_ERR_DOMAIN, "shift outside the grid");
        const int nx = (int)grid->nx, ny = (int)grid->ny;
        const size_t M = (size_t)nx * ny;
        std::vector<double> probe(2 * M);
        if (int r = msf_probe_realspace(c, grid, o, probe.data())) return r;
        // adfstem.cpp:1590-1671: circular shift by (sx, sy) pixels
        for (int i = 0; i < nx; ++i) {
            const int ix = (i < sx) ? i + nx - sx : i - sx;
            for (int j = 0; j < ny; ++j) {
                const int iy = (j < sy) ? j + ny - sy : j - sy;
                psi_k[2 * ((size_t)i * ny + j)] = probe[2 * ((size_t)ix * ny + iy)];
                psi_k[2 * ((size_t)i * ny + j) + 1] = probe[2 * ((size_t)ix * ny + iy) + 1];
            }
        }
        // :1831-1837: FFT on the device, / sqrt(NxNy)
        const double2 *twx, *twy;
        if (int r = twiddles(c, nx, nullptr, &twx)) return r;
        if (int r = twiddles(c, ny, nullptr, &twy)) return r;
        DevBuf d;
        HIPCHK(c, d.alloc(M * sizeof(double2)));
        HIPCHK(c, hipMemcpyAsync(d.p, psi_k, M * sizeof(double2), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, msf::fft2d_f64(d.as<double2>(), nx, ny, 1, false, twx, twy, c->stream));
        HIPCHK(c, hipMemcpyAsync(psi_k, d.p, M * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double s = (nx == ny) ? (double)nx : std::sqrt((double)M);
        for (size_t k = 0; k < 2 * M; ++k) psi_k[k] = psi_k[k] / s;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_slice_step_batch(msf_ctx* c, const msf_specimen* sp, uint32_t slice, double* psi_k,
                         uint32_t count)
{
    try {
        if (!c || !sp || !psi_k || count == 0) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (sp->ctx != c) return fail(c, MSF_ERR_INVALID, "specimen was built on another context");
        if ((int)slice >= sp->n_slices) return fail(c, MSF_ERR_INVALID, "slice index out of range");
        HIPCHK(c, hipSetDevice(c->device));
        const msf::Domain& d = sp->dom;
        const int nx = d.nx, ny = d.ny;
        const size_t M = (size_t)nx * ny;
        const double2 *twx, *twy;
        if (int r = twiddles(c, nx, nullptr, &twx)) return r;
        if (int r = twiddles(c, ny, nullptr, &twy)) return r;
        DevBuf w;
        HIPCHK(c, w.alloc(M * count * sizeof(double2)));
        double2* v = w.as<double2>();
        HIPCHK(c, hipMemcpyAsync(v, psi_k, M * count * sizeof(double2), hipMemcpyHostToDevice, c->stream));
        // adfstem.cpp:1891 IFFT; :1922-1952 transmission / s; :1986-1992 FFT / s;
        // bw_limit.cpp:30-47 and slice.cpp:1366-1421 mask and propagator
        HIPCHK(c, msf::fft2d_f64(v, nx, ny, (int)count, true, twx, twy, c->stream));
        HIPCHK(c, msf::launch_transmit_f64(v, sp->t64.as<double2>() + (size_t)slice * M, (int64_t)M,
                                           (int)count, c->cfg.transmit_mode, d.sqrt_nxny, c->stream));
        HIPCHK(c, msf::fft2d_f64(v, nx, ny, (int)count, false, twx, twy, c->stream));
        for (uint32_t k = 0; k < count; ++k)
            HIPCHK(c, msf::launch_vacuum_step(v + (size_t)k * M, nx, ny,
                                              sp->px64.as<double2>() + (size_t)slice * nx,
                                              sp->py64.as<double2>() + (size_t)slice * ny,
                                              sp->dkeep.as<int2>(), 1.0 / d.sqrt_nxny, c->stream));
        HIPCHK(c, hipMemcpyAsync(psi_k, v, M * count * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_probe_realspace(msf_ctx* c, const msf_grid* grid, const msf_optics* o, double* probe)
{
    try {
        if (!c || !probe) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (int r = check_grid(c, grid)) return r;
        if (int r = check_optics(c, o)) return r;
        HIPCHK(c, hipSetDevice(c->device));
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        DevBuf h, norm2;
        if (int r = probe_device(c, d, o, h, norm2)) return r;
        const size_t M = (size_t)d.nx * d.ny;
        HIPCHK(c, hipMemcpy(probe, h.p, M * sizeof(double2), hipMemcpyDeviceToHost));
        // adfstem.cpp:699-704 (/ sqrtNxNy), probe_wavefunction.cpp:300-321 (A_p), :813-817
        for (size_t k = 0; k < 2 * M; ++k) probe[k] = probe[k] / d.sqrt_nxny;
        double ap = 0.0;
        for (size_t k = 0; k < M; ++k) ap += probe[2 * k] * probe[2 * k] + probe[2 * k + 1] * probe[2 * k + 1];
        ap = 1.0 / std::sqrt(ap);
        for (size_t k = 0; k < 2 * M; ++k) probe[k] = ap * probe[k];
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_fem_bins(const msf_grid* grid, double dr, uint32_t* n_bins, double* lower_edges)
{
    try {
        if (!grid || !n_bins || !(dr > 0.0) || grid->nx < 2 || grid->ny < 2) return MSF_ERR_INVALID;
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        double dk = 0.0;
        const int nb = msf::fem_bin_count(d, dr, dk);
        *n_bins = (uint32_t)nb;
        if (lower_edges)
            for (int i = 0; i < nb; ++i) lower_edges[i] = (unsigned int)i * dk;
        return MSF_OK;
    }
    MSF_CATCH_ALL(nullptr)
}

} // extern "C"

namespace {

// Every pixel of the grid kept: the full [nx][ny] layout as its own compact plane.
msf::RowRanges full_ranges(int nx, int ny)
{
    msf::RowRanges r;
    r.ab.assign(2 * (size_t)nx, 0);
    for (int i = 0; i < nx; ++i) r.ab[2 * i] = ny;
    r.amax = ny;
    r.bmax = 0;
    return r;
}

// Group sums of `count` FP64 k-space waves on the device: out[count][G]
int group_sums_host(msf_ctx* c, const msf::Groups& g, const msf::Domain& d, const double* psi_k,
                    uint32_t count, std::vector<double>& out)
{
    const int G = (int)g.kind.size();
    const size_t M = (size_t)d.nx * d.ny;
    DevBuf w, off, pix, kind, part;
    HIPCHK(c, w.alloc(M * count * sizeof(double2)));
    HIPCHK(c, hipMemcpyAsync(w.p, psi_k, M * count * sizeof(double2), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, upload(off, g.offsets, c->stream));
    HIPCHK(c, upload(pix, g.pixels, c->stream));
    HIPCHK(c, upload(kind, g.kind, c->stream));
    HIPCHK(c, part.alloc((size_t)count * G * sizeof(double)));
    HIPCHK(c, msf::launch_group_sums_f64(w.as<double2>(), (int64_t)M, off.as<int32_t>(),
                                         pix.as<int32_t>(), kind.as<int32_t>(), G, (int)count,
                                         part.as<double>(), c->stream));
    out.resize((size_t)count * G);
    HIPCHK(c, hipMemcpyAsync(out.data(), part.p, out.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MSF_OK;
}

} // namespace
extern "C" {

int msf_adf_reduce(msf_ctx* c, const msf_grid* grid, const msf_detector* det,
                   const double* psi_k, uint32_t count, double* adf)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (!det || !psi_k || !adf) return fail(c, MSF_ERR_INVALID, "NULL argument");
        if (count == 0 || count > 65535) return fail(c, MSF_ERR_INVALID, "count must be 1..65535");
        if (int r = check_grid(c, grid)) return r;
        if (det->inner < 0.0 || det->outer < 0.0)
            return fail(c, MSF_ERR_INVALID, "detector angles must be >= 0 (tem.cpp:710-720)");
        HIPCHK(c, hipSetDevice(c->device));
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        double k1 = det->inner / d.lambda, k2 = det->outer / d.lambda; // adfstem.cpp:2251-2252
        if (det->threshold_mode == MSF_ADF_PHYSICAL) {
            k1 = k1 * k1;
            k2 = k2 * k2;
        }
        const msf::Groups g = msf::build_groups(d, full_ranges(d.nx, d.ny), d.ny, true, k1, k2,
                                                false, 1.0, 0.0);
        std::vector<double> part;
        if (int r = group_sums_host(c, g, d, psi_k, count, part)) return r;
        const size_t G = g.kind.size();
        for (uint32_t p = 0; p < count; ++p) adf[p] = part[p * G];
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_fem_accumulate(msf_ctx* c, const msf_grid* grid, const msf_fem* fem, double alpha_max,
                       const double* psi_k, uint32_t count, double* radial, int32_t* counts,
                       double* v_omega)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (!fem || !psi_k) return fail(c, MSF_ERR_INVALID, "NULL argument");
        if (!(fem->dr > 0.0)) return fail(c, MSF_ERR_INVALID, "dr must be positive");
        if (count == 0 || count > 65535) return fail(c, MSF_ERR_INVALID, "count must be 1..65535");
        if (int r = check_grid(c, grid)) return r;
        HIPCHK(c, hipSetDevice(c->device));
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        const msf::Groups g = msf::build_groups(d, full_ranges(d.nx, d.ny), d.ny, false, 0.0, 0.0,
                                                true, fem->dr, alpha_max);
        std::vector<double> part;
        if (int r = group_sums_host(c, g, d, psi_k, count, part)) return r;
        const int nb = g.nbins, G = (int)g.kind.size();
        if (radial)
            for (uint32_t p = 0; p < count; ++p)
                for (int k = 0; k < nb; ++k)
                    radial[(size_t)p * nb + k] = g.zeroed[k] ? 0.0 : part[(size_t)p * G + g.fem_abs + k];
        if (counts)
            for (int k = 0; k < nb; ++k) counts[k] = g.counts[k];
        if (v_omega && nb > 0) { // the production kernels on these sums (adfstem.cpp:2591-2599, :2837-2856)
            DevBuf dp, dc, dz, st, v;
            HIPCHK(c, upload(dp, part, c->stream));
            HIPCHK(c, upload(dc, g.counts, c->stream));
            HIPCHK(c, upload(dz, g.zeroed, c->stream));
            HIPCHK(c, st.alloc((size_t)nb * 3 * sizeof(double)));
            HIPCHK(c, v.alloc((size_t)nb * sizeof(double)));
            HIPCHK(c, hipMemsetAsync(st.p, 0, (size_t)nb * 3 * sizeof(double), c->stream));
            HIPCHK(c, msf::launch_fem_welford(dp.as<double>(), G, (int)count, nb, dc.as<int32_t>(),
                                              dz.as<uint8_t>(), st.as<double>(), c->stream));
            HIPCHK(c, msf::launch_fem_finalize(st.as<double>(), nb, v.as<double>(), c->stream));
            HIPCHK(c, hipMemcpyAsync(v_omega, v.p, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_mtf_2d(msf_ctx* c, const msf_grid* grid, const msf_fem* fem, double* mtf2d)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (!mtf2d) return fail(c, MSF_ERR_INVALID, "NULL argument");
        if (int r = check_grid(c, grid)) return r;
        if (int r = check_mtf(c, fem)) return r;
        HIPCHK(c, hipSetDevice(c->device));
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        DevBuf m;
        if (int r = build_mtf2d(c, d, grid->xperiod, grid->yperiod, fem, m)) return r;
        HIPCHK(c, hipMemcpy(mtf2d, m.p, (size_t)d.nx * d.ny * sizeof(double2), hipMemcpyDeviceToHost));
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_apply_mtf(msf_ctx* c, const msf_grid* grid, const msf_fem* fem, double alpha_max,
                  double* psi_k, uint32_t count)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        if (!psi_k) return fail(c, MSF_ERR_INVALID, "NULL argument");
        if (count == 0 || count > 65535) return fail(c, MSF_ERR_INVALID, "count must be 1..65535");
        if (int r = check_grid(c, grid)) return r;
        if (int r = check_mtf(c, fem)) return r;
        HIPCHK(c, hipSetDevice(c->device));
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        const int nx = d.nx, ny = d.ny;
        const size_t M = (size_t)nx * ny;
        DevBuf m, w, out, keep;
        if (int r = build_mtf2d(c, d, grid->xperiod, grid->yperiod, fem, m)) return r;
        const double2 *twx, *twy;
        if (int r = twiddles(c, nx, nullptr, &twx)) return r;
        if (int r = twiddles(c, ny, nullptr, &twy)) return r;
        HIPCHK(c, upload(out, msf::outside_aperture(d, alpha_max), c->stream));
        HIPCHK(c, upload(keep, to_int2(msf::bwlimit_keep_ranges(d)), c->stream));
        HIPCHK(c, w.alloc(M * count * sizeof(double2)));
        double2* v = w.as<double2>();
        HIPCHK(c, hipMemcpyAsync(v, psi_k, M * count * sizeof(double2), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, msf::launch_mtf_magnitude_f64(v, (int64_t)M, (int)count, out.as<uint8_t>(), c->stream));
        HIPCHK(c, msf::fft2d_f64(v, nx, ny, (int)count, true, twx, twy, c->stream));
        HIPCHK(c, msf::launch_transmit_f64(v, m.as<double2>(), (int64_t)M, (int)count, 0, d.sqrt_nxny,
                                           c->stream));
        HIPCHK(c, msf::fft2d_f64(v, nx, ny, (int)count, false, twx, twy, c->stream));
        HIPCHK(c, msf::launch_div_f64(v, (int64_t)(M * count), d.sqrt_nxny, c->stream));
        HIPCHK(c, msf::launch_apply_keep(v, nx, ny, (int)count, keep.as<int2>(), c->stream));
        HIPCHK(c, hipMemcpyAsync(psi_k, v, M * count * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_correlograph_bins(const msf_grid* grid, double dr, uint32_t* n_phi, double* k_edges,
                          double* phi_edges)
{
    try {
        if (!grid || !n_phi || !(dr > 0.0) || grid->nx < 2 || grid->ny < 2) return MSF_ERR_INVALID;
        const msf::Domain d = msf::make_domain((int)grid->nx, (int)grid->ny, grid->xperiod,
                                               grid->yperiod, grid->xmin, grid->ymin, grid->voltage);
        double dk = 0.0;
        const int nb = msf::fem_bin_count(d, dr, dk);
        const int np = msf::correlograph_nphi(dk, nb);
        *n_phi = (uint32_t)np;
        if (k_edges)
            for (unsigned int ii = 0; ii < (unsigned int)nb + 1; ++ii) k_edges[ii] = ii * dk;
        if (phi_edges && np > 0) {
            const double dphi = 2 * msf::kRefPi / np;
            for (unsigned int ii = 0; ii < (unsigned int)np + 1; ++ii) phi_edges[ii] = ii * dphi;
        }
        return MSF_OK;
    }
    MSF_CATCH_ALL(nullptr)
}

int msf_run_stem(msf_ctx* c, const msf_specimen* sp, const msf_optics* o, const msf_detector* det,
                 const msf_fem* fem, const double* x_p, uint32_t n_xp, const double* y_p,
                 uint32_t n_yp, double* adf_image, double* fem_v)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        HIPCHK(c, hipSetDevice(c->device));
        if (det && det->enabled && !adf_image) return fail(c, MSF_ERR_INVALID, "adf_image is NULL");
        if (fem && fem->enabled && !fem_v) return fail(c, MSF_ERR_INVALID, "fem_v is NULL");
        if (!sp) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
        if (!x_p || !y_p || n_xp == 0 || n_yp == 0) return fail(c, MSF_ERR_INVALID, "empty raster");
        std::vector<int2> shifts;
        if (!raster_shifts(sp->dom, x_p, n_xp, y_p, n_yp, shifts))
            return fail(c, MSF_ERR_DOMAIN, "probe position below the domain minimum");
        RunOut out;
        out.adf = adf_image;
        out.fem = fem_v;
        return run_internal(c, sp, o, det, fem, shifts, out);
    }
    MSF_CATCH_ALL(c)
}

int msf_run_stem_fem(msf_ctx* c, const msf_specimen* sp, const msf_optics* o,
                     const msf_detector* det, const msf_fem* fem, const double* x_p,
                     uint32_t n_xp, const double* y_p, uint32_t n_yp, double* adf_image,
                     const msf_fem_outputs* fo)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        HIPCHK(c, hipSetDevice(c->device));
        if (det && det->enabled && !adf_image) return fail(c, MSF_ERR_INVALID, "adf_image is NULL");
        if (!sp) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
        if (!x_p || !y_p || n_xp == 0 || n_yp == 0) return fail(c, MSF_ERR_INVALID, "empty raster");
        std::vector<int2> shifts;
        if (!raster_shifts(sp->dom, x_p, n_xp, y_p, n_yp, shifts))
            return fail(c, MSF_ERR_DOMAIN, "probe position below the domain minimum");
        RunOut out;
        out.adf = adf_image;
        if (fo) {
            out.fem = fo->v_omega;
            out.vbar_r = fo->vbar_r;
            out.v_re = fo->v_re;
            out.omega_vimage = fo->omega_vimage;
            out.v_gt17 = fo->v_gt17;
            out.v_rbar = fo->v_rbar;
            out.corr_avg = fo->correlograph_avg;
            out.corr_sqr = fo->correlograph_sqr_sum;
            out.corr_each = fo->correlograph_each;
            out.mean_abs = fo->diffraction_mean_abs;
            out.mean_sqr = fo->diffraction_mean_sqr;
            out.var2d = fo->variance_2d;
        }
        return run_internal(c, sp, o, det, fem, shifts, out);
    }
    MSF_CATCH_ALL(c)
}

int msf_probe_shifts(const msf_grid* g, const double* x_p, uint32_t n_xp, const double* y_p,
                     uint32_t n_yp, int32_t* shifts)
{
    try {
        if (!g || !x_p || !y_p || !shifts || g->nx < 2 || g->ny < 2) return MSF_ERR_INVALID;
        const msf::Domain d =
            msf::make_domain((int)g->nx, (int)g->ny, g->xperiod, g->yperiod, g->xmin, g->ymin, g->voltage);
        std::vector<int2> sh;
        if (!raster_shifts(d, x_p, n_xp, y_p, n_yp, sh)) return MSF_ERR_DOMAIN;
        for (size_t k = 0; k < sh.size(); ++k) {
            shifts[2 * k] = sh[k].x;
            shifts[2 * k + 1] = sh[k].y;
        }
        return MSF_OK;
    }
    MSF_CATCH_ALL(nullptr)
}

int msf_exit_waves_shifted(msf_ctx* c, const msf_specimen* sp, const msf_optics* o,
                           const int32_t* shifts, uint32_t n, float* psi)
{
    try {
        if (!c || !psi || !shifts) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        HIPCHK(c, hipSetDevice(c->device));
        if (n == 0) return fail(c, MSF_ERR_INVALID, "empty shift list");
        std::vector<int2> sh(n);
        for (uint32_t k = 0; k < n; ++k) sh[k] = int2{shifts[2 * k], shifts[2 * k + 1]};
        RunOut out;
        out.exit = psi;
        return run_internal(c, sp, o, nullptr, nullptr, sh, out);
    }
    MSF_CATCH_ALL(c)
}

int msf_exit_waves(msf_ctx* c, const msf_specimen* sp, const msf_optics* o, const double* x_p,
                   uint32_t n_xp, const double* y_p, uint32_t n_yp, float* psi)
{
    try {
        if (!c || !psi) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        HIPCHK(c, hipSetDevice(c->device));
        if (!sp) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
        if (!x_p || !y_p || n_xp == 0 || n_yp == 0) return fail(c, MSF_ERR_INVALID, "empty raster");
        std::vector<int2> shifts;
        if (!raster_shifts(sp->dom, x_p, n_xp, y_p, n_yp, shifts))
            return fail(c, MSF_ERR_DOMAIN, "probe position below the domain minimum");
        RunOut out;
        out.exit = psi;
        return run_internal(c, sp, o, nullptr, nullptr, shifts, out);
    }
    MSF_CATCH_ALL(c)
}

// ---------------------------------------------------------------- serving
} // extern "C"

namespace {

double ms_since(double t0) { return now_ms() - t0; }

// Device bytes a job claims while it runs (specimen, vacuum waves, batch
// buffers of both chains, set-up temporaries), over-estimated by 10 %.
uint64_t job_footprint(const ServeJob& j)
{
    const uint64_t M = (uint64_t)j.grid.nx * j.grid.ny;
    const msf::Slices sl = msf::autoslice(j.q.data(), (uint32_t)j.Z.size(), j.grid.zmin,
                                          j.grid.zperiod, j.min_slice);
    const uint64_t S = std::max<uint64_t>(1, sl.lower.size());
    const uint64_t P = (uint64_t)j.xp.size() * j.yp.size();
    uint64_t wave = j.cfg.batch ? (uint64_t)j.cfg.batch * M * 8 : (112ull << 20);
    wave = std::min<uint64_t>(wave, P * M * 8 + (16ull << 20));
    uint64_t b = 40 * M * S                     // t (FP32 + FP64) and u0 stacks
                 + 64 * M                       // probe, vacuum wave, per-pixel sums
                 + 2 * wave                     // two chains of wave buffers
                 + (256ull << 20);              // species maps of the set-up
    if (j.mtf.size() && j.fem.mtf_sampling_rate != 0.0)
        b += 2 * std::min<uint64_t>(P, (wave / 8 / std::max<uint64_t>(M, 1)) + 16) * M * 16;
    return b + b / 10 + (16ull << 20);
}

// a served job's work buffers, released between jobs; the pixel groups stay
// for the next job on the same grid and settings
void reset_plan(msf_ctx* c) { c->plan.release_work(); }

void serve_worker(msf_ctx* c, uint32_t slot)
{
    ServeState& st = *c->serve;
    msf_ctx* sc = st.slot_ctx[slot];
    for (;;) {
        ServeJob job;
        double t_start = 0;
        {
            std::unique_lock<std::mutex> lk(st.m);
            st.work_cv.wait(lk, [&] {
                if (st.queue.empty()) return st.stop;
                return st.running == 0 || st.admitted + st.queue.front().footprint <= st.limit;
            });
            if (st.queue.empty()) return; // stop requested and nothing left
            job = std::move(st.queue.front());
            st.queue.pop_front();
            st.admitted += job.footprint;
            st.running += 1;
            st.max_running = std::max(st.max_running, st.running);
            st.peak_admitted = std::max(st.peak_admitted, st.admitted);
            t_start = now_ms();
        }
        st.work_cv.notify_all(); // the next head job may fit as well
        ServeDone d;
        {
            msf_specimen* sp = nullptr;
            d.rc = msf_set_config(sc, &job.cfg);
            if (d.rc == MSF_OK)
                d.rc = msf_specimen_build(sc, &job.grid, job.q.data(), job.Z.data(),
                                          (uint32_t)job.Z.size(), job.min_slice, &sp);
            if (!job.mtf.empty()) {
                job.fem.mtf = job.mtf.data();
                job.fem.mtf_domain = job.mtf_dom.data();
            }
            if (d.rc == MSF_OK)
                d.rc = msf_run_stem_fem(sc, sp, &job.optics, &job.det, &job.fem, job.xp.data(),
                                        (uint32_t)job.xp.size(), job.yp.data(),
                                        (uint32_t)job.yp.size(), job.adf, &job.fo);
            if (d.rc != MSF_OK) d.msg = sc->err;
            if (sp) msf_specimen_free(sp);
            std::lock_guard<std::recursive_mutex> lk(sc->mu);
            reset_plan(sc);
        }
        d.info.queue_ms = t_start - job.t_submit;
        d.info.run_ms = ms_since(t_start);
        d.info.footprint = job.footprint;
        d.info.slot = slot;
        {
            std::lock_guard<std::mutex> lk(st.m);
            st.admitted -= job.footprint;
            st.running -= 1;
            st.jobs_done += 1;
            st.pending.erase(job.id);
            st.done[job.id] = std::move(d);
        }
        st.work_cv.notify_all();
        st.done_cv.notify_all();
    }
}

// starts the engine on first use: slot contexts on the same device, workers
int start_serving(msf_ctx* c)
{
    std::lock_guard<std::mutex> lk(c->serve_mu);
    if (c->serve) return MSF_OK;
    std::unique_ptr<ServeState> st(new ServeState);
    st->slots = c->serve_slots ? c->serve_slots : 3;
    st->limit = c->serve_limit;
    if (!st->limit) {
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(c->device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            return MSF_ERR_DEVICE;
        st->limit = (uint64_t)(0.8 * (double)free_b);
    }
    for (uint32_t k = 0; k < st->slots; ++k) {
        msf_ctx* sc = nullptr;
        if (int r = msf_create(&sc, c->device)) {
            for (msf_ctx* x : st->slot_ctx) msf_destroy(x);
            return r;
        }
        st->slot_ctx.push_back(sc);
    }
    c->serve = std::move(st);
    for (uint32_t k = 0; k < c->serve->slots; ++k)
        c->serve->workers.emplace_back(serve_worker, c, k);
    return MSF_OK;
}

} // namespace
extern "C" {

int msf_submit(msf_ctx* c, const msf_job* j, uint64_t* id)
{
    try {
        if (!c || !j || !id) return MSF_ERR_INVALID;
        if (!j->x_p || !j->y_p || j->n_xp == 0 || j->n_yp == 0 || (j->n_atoms && (!j->q_xyz || !j->Z)))
            return MSF_ERR_INVALID;
        if (j->detector.enabled && !j->adf_image) return MSF_ERR_INVALID;
        ServeJob job;
        job.grid = j->grid;
        job.q.assign(j->q_xyz, j->q_xyz + 3 * (size_t)j->n_atoms);
        job.Z.assign(j->Z, j->Z + j->n_atoms);
        job.min_slice = j->min_slice;
        job.optics = j->optics;
        job.xp.assign(j->x_p, j->x_p + j->n_xp);
        job.yp.assign(j->y_p, j->y_p + j->n_yp);
        job.det = j->detector;
        job.fem = j->fem;
        if (j->fem.mtf && j->fem.mtf_domain && j->fem.mtf_n) {
            job.mtf.assign(j->fem.mtf, j->fem.mtf + j->fem.mtf_n);
            job.mtf_dom.assign(j->fem.mtf_domain, j->fem.mtf_domain + j->fem.mtf_n);
        }
        job.fem.mtf = nullptr;
        job.fem.mtf_domain = nullptr;
        job.adf = j->adf_image;
        job.fo = j->fem_out;
        {
            MSF_LOCK(c);
            job.cfg = c->cfg;
            job.cfg.profile_kernels = 0;
            // served jobs launch on their streams: a graph capture on one worker
            // would be invalidated by another worker's allocations
            job.cfg.no_graph = 1;
        }
        job.footprint = job_footprint(job);
        job.t_submit = now_ms();
        if (int r = start_serving(c)) return r;
        ServeState& st = *c->serve;
        {
            std::lock_guard<std::mutex> lk(st.m);
            job.id = st.next_id++;
            *id = job.id;
            st.pending.insert(job.id);
            st.queue.push_back(std::move(job));
        }
        st.work_cv.notify_one();
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_poll(msf_ctx* c, uint64_t id)
{
    try {
        if (!c || !c->serve) return MSF_ERR_INVALID;
        ServeState& st = *c->serve;
        std::lock_guard<std::mutex> lk(st.m);
        auto it = st.done.find(id);
        if (it != st.done.end()) return it->second.rc;
        return st.pending.count(id) ? MSF_PENDING : MSF_ERR_INVALID;
    }
    MSF_CATCH_ALL(c)
}

int msf_wait_info(msf_ctx* c, uint64_t id, msf_job_info* info)
{
    try {
        if (!c || !c->serve) return MSF_ERR_INVALID;
        ServeState& st = *c->serve;
        ServeDone r;
        {
            std::unique_lock<std::mutex> lk(st.m);
            if (!st.done.count(id) && !st.pending.count(id)) return MSF_ERR_INVALID;
            st.done_cv.wait(lk, [&] { return st.done.count(id) != 0; });
            r = st.done[id];
            st.done.erase(id);
        }
        if (info) *info = r.info;
        if (r.rc != MSF_OK) {
            MSF_LOCK(c);
            c->err = "job " + std::to_string(id) + ": " + r.msg;
        }
        return r.rc;
    }
    MSF_CATCH_ALL(c)
}

int msf_wait(msf_ctx* c, uint64_t id) { return msf_wait_info(c, id, nullptr); }

int msf_serve_configure(msf_ctx* c, uint32_t slots, uint64_t mem_limit)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        std::lock_guard<std::mutex> lk(c->serve_mu);
        if (c->serve) return fail(c, MSF_ERR_INVALID, "the serving engine has already started");
        if (slots > 16) return fail(c, MSF_ERR_INVALID, "slots must be 0..16");
        c->serve_slots = slots;
        c->serve_limit = mem_limit;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

int msf_serve_get_stats(msf_ctx* c, msf_serve_stats* out)
{
    try {
        if (!c || !out) return MSF_ERR_INVALID;
        *out = msf_serve_stats{};
        std::lock_guard<std::mutex> lk0(c->serve_mu);
        if (!c->serve) return MSF_OK;
        ServeState& st = *c->serve;
        std::lock_guard<std::mutex> lk(st.m);
        out->slots = st.slots;
        out->running = st.running;
        out->max_running = st.max_running;
        out->queued = st.queue.size();
        out->jobs_done = st.jobs_done;
        out->mem_limit = st.limit;
        out->admitted_bytes = st.admitted;
        out->peak_admitted_bytes = st.peak_admitted;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

uint64_t msf_job_footprint(const msf_job* j)
{
    if (!j || (j->n_atoms && (!j->q_xyz || !j->Z))) return 0;
    ServeJob job;
    job.grid = j->grid;
    job.q.assign(j->q_xyz, j->q_xyz + 3 * (size_t)j->n_atoms);
    job.Z.assign(j->Z, j->Z + j->n_atoms);
    job.min_slice = j->min_slice;
    job.xp.assign(j->x_p, j->x_p + j->n_xp);
    job.yp.assign(j->y_p, j->y_p + j->n_yp);
    job.fem = j->fem;
    if (j->fem.mtf && j->fem.mtf_n) job.mtf.assign(j->fem.mtf, j->fem.mtf + j->fem.mtf_n);
    return job_footprint(job);
}

} // extern "C"
