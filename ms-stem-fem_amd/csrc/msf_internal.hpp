// msf_internal.hpp -- state shared by the units of libmsf.so behind the C
// ABI: device buffers and where their memory comes from, the run plan cached
// on a context, the serving queue, the context and the specimen.
#pragma once

#include "../../include/msf.h"

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
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

namespace msf_impl {

// A device allocation owned by the library (hipMalloc / hipFree).
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

// Where the device memory of one call comes from.  Ordinary calls grow
// buffers owned by the context and keep them for the next call.  The
// allocation-free run carves every buffer out of one caller-independent
// region (the library's static scratch array) with a bump pointer that
// starts over at every call, and never allocates.
struct Mem {
    char* base = nullptr;    // region of the allocation-free run; null: ordinary
    size_t cap = 0, used = 0;
    bool scratch() const { return base != nullptr; }
    size_t left() const { return cap - used; }
};

double now_ms();

// Everything a raster run needs besides the positions, cached on the context
// and rebuilt only when the specimen, optics, detector or FEM settings change.
// The raw pointers are the views the launch sequence uses: into the owning
// DevBufs for ordinary runs, into the scratch region for the allocation-free
// run (whose plan is rebuilt at every call).
struct RunPlan {
    uint64_t spec_id = 0;
    msf_optics optics{};
    bool valid = false;      // vacuum waves of (spec_id, optics) are current
    int ncp = 0, G = 0, nbins = 0;
    // vacuum waves: u0 of every slice [S][nx][ny] pairs (s * real space; for
    // every shift r of a row, u0[r] and u0[(r + ny/2) mod ny]: pair_pos) and
    // the k-space vacuum exit wave, column-major [ny][nx], FP32
    DevBuf U0, Psi0;
    float2* u0 = nullptr;
    float2* psi0 = nullptr;
    // reduction groups (host tables and their device copies); they depend on
    // the grid, the aperture and the detector/FEM settings only and are kept
    // across specimens on the same grid
    bool groups_ok = false;  // host tables and their keys are complete
    bool groups_dev = false; // device copies match the host tables
    msf_grid ggrid{};
    double galpha = 0.0;
    msf_detector det{};      // keys of the tables: detector and FEM settings
    msf_fem fem{};
    bool fem_sqr = false;    // groups carry the |psi|^2 FEM rings (Vbar_r)
    bool corr = false;       // groups carry the correlograph sectors
    msf::Groups hg;          // host groups (also for the per-pixel FEM modes)
    msf::SegmentTiles st;    // host segment tables
    int n_tiles = 0;
    DevBuf seg_pix_b, tile_off_b, seg_kind_b, seg_of_group_b, cnt_b, zero_b, grp_small_b,
        grp_large_b;
    int32_t *seg_pix = nullptr, *tile_off = nullptr, *seg_kind = nullptr,
            *seg_of_group = nullptr, *cnt = nullptr, *grp_small = nullptr, *grp_large = nullptr;
    std::vector<int32_t> small_ids, large_ids; // groups summed by one lane / one wavefront
    uint8_t* zero = nullptr;
    // detector MTF: the table it was built from, the 2-D MTF [nx][ny] and the
    // outside-aperture map
    bool mtf_on = false;
    std::vector<double> mtf_v, mtf_d;
    double mtf_rate = 0.0;
    DevBuf mtf2d_b, mtf_out_b;
    double2* mtf2d = nullptr;
    uint8_t* mtf_out = nullptr;
    // work buffers of the two batch chains and the run's accumulators
    int64_t cap_b = 0;       // batch depth the chain buffers were sized for
    DevBuf psi_b[2], part_b[2], segp_b[2], mtf_full_b[2];
    DevBuf state_b, femv_b, adf_b, shift_b, exit_b, vsum_b, pm1_b, pm2_b, csum_b, csqr_b,
        ceach_b;
    // the job's launch sequence as an instantiated HIP graph, replayed while
    // every pointer and size it was captured with is unchanged
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_sig;
    uint64_t graph_fused = 0; // fused pass launches in the captured graph
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
        for (DevBuf* b : {&U0, &Psi0, &state_b, &femv_b, &adf_b, &shift_b, &exit_b, &vsum_b,
                          &pm1_b, &pm2_b, &csum_b, &csqr_b, &ceach_b})
            b->release();
        for (int k = 0; k < 2; ++k)
            for (DevBuf* b : {&psi_b[k], &part_b[k], &segp_b[k], &mtf_full_b[k]}) b->release();
        u0 = psi0 = nullptr;
        cap_b = 0;
    }
};

// Stage twiddle tables of one transform length, built in FP64 on the host
// and rounded once: FP64 and FP32 tables of the primary plan, the natural
// roots W_n^e in FP32, the FP32 table of the alternate plan (when the length
// has one).
struct TwiddleHost {
    std::vector<double2> f64;
    std::vector<float2> f32, nat, alt;
};
struct TwiddleDev {
    DevBuf f64, f32, nat, alt;
};
// device addresses of one length's tables for one call
struct Twiddles {
    const double2* f64 = nullptr;
    const float2* f32 = nullptr;
    const float2* nat = nullptr;
    const float2* alt = nullptr; // null when the length has no alternate plan
};

// A job of the serving queue: inputs copied at submission, outputs written
// to the caller's arrays.
struct ServeJob {
    uint64_t id = 0;
    msf_specimen* spec = nullptr;     // resident specimen (msf_submit_run), or null: build one
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
    // resident specimens referred to by queued or running jobs (by id): their
    // device bytes count against the limit once, not per job
    struct Resident {
        uint64_t bytes = 0;
        uint32_t jobs = 0;
    };
    std::map<uint64_t, Resident> resident;
    uint64_t resident_bytes = 0;
    uint64_t specimen_builds = 0;
    // grouping (msf_serve_group): a worker takes up to group_max queued jobs
    // of one grid size and settings and runs them with shared launches
    uint32_t group_max = 1;
    uint64_t groups = 0, grouped_jobs = 0, group_launches = 0;
};

// Launch shape of a run: positions per batch, column-pass tile, batch
// chains in flight and the radix plans (vx + 2 vy, per axis 0 primary,
// 1 alternate).
using TunedShape = std::array<int, 4>;       // batch, col_tile, chains, plan
// msf_config.autotune of served jobs: the tuning table's choices, no trials
constexpr int32_t kAutotuneLookup = 2;

} // namespace msf_impl

struct msf_ctx {
    std::recursive_mutex mu;    // serialises device work of the context
    std::mutex serve_mu;        // guards the creation of the serving engine
    std::unique_ptr<msf_impl::ServeState> serve;
    uint32_t serve_slots = 0;   // msf_serve_configure (0: automatic)
    uint64_t serve_limit = 0;
    uint32_t serve_group = 0;   // msf_serve_group (0, 1: each job alone)
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream and the events that order two batch chains (created on first use)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fem[2] = {nullptr, nullptr};
    std::string err;
    msf_config cfg{0, MSF_TRANSMIT_REFERENCE, 0, 0, 0, 0, 0, 0};
    msf_stats stats{};
    std::map<int, msf_impl::TwiddleHost> tw_host;
    std::map<int, msf_impl::TwiddleDev> tw_dev;
    std::string arch;           // device key of the tuning table
    msf_impl::RunPlan plan;     // ordinary runs
    msf_impl::RunPlan plan_na;  // the allocation-free run (device views rebuilt per call)
    std::vector<std::unique_ptr<msf_impl::RunPlan>> gplans; // runs of a group, one each
    char* scratch = nullptr;    // this device's copy of the allocation-free run's array
    msf_ctx* owner = nullptr;   // serving slot: the context whose specimens it may run on
    std::set<msf_specimen*> specimens; // live specimens built on this context
};

struct msf_specimen {
    msf_ctx* ctx = nullptr;
    msf_grid grid{};
    msf::Domain dom;
    int n_slices = 0;
    std::vector<double> lower, thick;
    std::vector<uint32_t> counts;
    msf_impl::DevBuf t32;  // [S][nx][ny] float2, t / sqrtNxNy, rows in the pair layout (pair_pos)
    msf_impl::DevBuf t64;  // [S][nx][ny] double2, t as built
    msf_impl::DevBuf px;   // [S][nx] float2
    msf_impl::DevBuf py;   // [S][ny] float2, P_y / sqrtNxNy
    msf_impl::DevBuf px64; // [S][nx] double2
    msf_impl::DevBuf py64; // [S][ny] double2
    uint64_t id = 0;           // unique per build, keys the run-plan cache
    msf::RowRanges keep;       // columns the slice step keeps (bw_limit + propagate)
    msf_impl::DevBuf dkeep;    // keep as int2 [nx]
    msf_impl::DevBuf dcolkeep; // the same set by compact column, int2 [ncp] (column pass)
    std::atomic<int> served{0}; // queued or running served jobs that run on it
};

namespace msf_impl {

int fail(msf_ctx* c, int code, const std::string& m);

// device work of a context is serialised (direct calls and the serving worker)
#define MSF_LOCK(c) std::lock_guard<std::recursive_mutex> msf_lock_((c)->mu)

// A failed call is reported once: the runtime's last-error state is cleared,
// so the next launch's hipGetLastError() check reports only its own status.
#define HIPCHK(ctx, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (void)hipGetLastError();                                                          \
            return ::msf_impl::fail((ctx),                                                    \
                                    e_ == hipErrorOutOfMemory ? MSF_ERR_MEMORY : MSF_ERR_DEVICE, \
                                    std::string(#expr) + ": " + hipGetErrorString(e_));       \
        }                                                                                     \
    } while (0)

// Every entry point turns a C++ exception (a failed host allocation, a length
// error from a malformed size) into an error code: none may cross the C ABI.
#define MSF_CATCH_ALL(c)                                                                      \
    catch (const std::bad_alloc&)                                                             \
    {                                                                                         \
        return ::msf_impl::fail((c), MSF_ERR_MEMORY, "host allocation failed");              \
    }                                                                                         \
    catch (const std::exception& e_)                                                          \
    {                                                                                         \
        return ::msf_impl::fail((c), MSF_ERR_INVALID,                                        \
                                std::string("unexpected error: ") + e_.what());              \
    }                                                                                         \
    catch (...)                                                                               \
    {                                                                                         \
        return ::msf_impl::fail((c), MSF_ERR_INVALID, "unexpected error");                   \
    }

// ---- device memory (msf_run.cpp)
// n bytes of device memory for buffer `b`: a fresh 256-byte aligned slice of
// the scratch region, or b grown to at least n bytes (a grown buffer's old
// contents are lost).  *fresh tells whether the memory is new to this buffer.
int take(msf_ctx* c, Mem& m, DevBuf& b, size_t n, void** out, bool* fresh = nullptr);
template <typename T> int take_as(msf_ctx* c, Mem& m, DevBuf& b, size_t count, T** out,
                                  bool* fresh = nullptr)
{
    void* p = nullptr;
    const int r = take(c, m, b, count * sizeof(T), &p, fresh);
    *out = static_cast<T*>(p);
    return r;
}
// v copied to the device (into b, or a scratch slice)
template <typename T> int put(msf_ctx* c, Mem& m, DevBuf& b, const std::vector<T>& v, T** out)
{
    if (int r = take_as<T>(c, m, b, v.size(), out)) return r;
    if (!v.empty())
        HIPCHK(c, hipMemcpyAsync(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                                 c->stream));
    return MSF_OK;
}
// copy of v into a fresh allocation (ordinary, uncached temporaries)
template <typename T> hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t s)
{
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

// ---- shared set-up (msf_run.cpp)
int twiddles(msf_ctx* c, Mem& m, int n, Twiddles& out);
int check_grid(msf_ctx* c, const msf_grid* g);
int check_optics(msf_ctx* c, const msf_optics* o);
int check_mtf(msf_ctx* c, const msf_fem* fem);
std::vector<int2> to_int2(const msf::RowRanges& r);
int fft_plan_rule(int n);
// radix plans of a run, vx + 2 vy (per axis 0 primary, 1 alternate), from
// msf_config.fft_plan and the lengths' alternate tables
int plan_choice(const msf_config& cfg, int nx, int ny, const Twiddles& twx, const Twiddles& twy);
// H(k) -> BACKWARD FFT: h = IFFT(H) (unnormalised, [nx][ny] double2) and its
// squared norm, in memory from m
int probe_device(msf_ctx* c, Mem& m, const msf::Domain& d, const msf_optics* o, DevBuf& hb,
                 double2** h, DevBuf& nb, double** norm2);
int build_mtf2d(msf_ctx* c, Mem& m, const msf::Domain& d, double xperiod, double yperiod,
                const msf_fem* fem, DevBuf& ob, double2** out);
bool raster_shifts(const msf::Domain& d, const double* x_p, uint32_t n_xp, const double* y_p,
                   uint32_t n_yp, std::vector<int2>& shifts);

// Caller buffers of one run (null: not wanted).
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

// What one run computes, from its arguments.
struct RunFlags {
    bool want_adf = false, fem_on = false, want_d1 = false, want_d2 = false, want_pix = false,
         want_corr = false, want_fem = false, want_mtf = false, want_each = false;
    int tmode = 0;
    int64_t P = 0;
    int nx = 0, ny = 0, S = 0, ka = 0, kb = 0, ncp = 0;
    int64_t M = 0, plane = 0;
    int nphi = 0;
    size_t ncorr = 1;
};

// Launch shape of one run.
struct Shape {
    int64_t B = 1, n_batches = 1;
    int col_tile = 0;        // configured tile, 0: the plan's default
    int chains = 1;
    int plan_v = 0;          // radix plans, vx + 2 vy
    int tuned_state = 0;
    bool auto_eager = false; // two automatic chains launch on the streams
};

// Device addresses of one run's buffers.
struct RunViews {
    float2* psi[2] = {nullptr, nullptr};
    double* part[2] = {nullptr, nullptr};
    double* segp[2] = {nullptr, nullptr};
    double2* mtf_full[2] = {nullptr, nullptr};
    double *state = nullptr, *femv = nullptr, *adf = nullptr, *vsum = nullptr, *pm1 = nullptr,
           *pm2 = nullptr, *csum = nullptr, *csqr = nullptr, *ceach = nullptr;
    int2* shift = nullptr;
    float2* exit = nullptr;
    Twiddles twx, twy;
};


// Per-kernel timing with HIP events (profile_kernels).
enum { K_ROW0, K_ROW, K_COL, K_COLL, K_RED, K_NUM };
struct Prof {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    hipEvent_t ka0 = nullptr, ka1 = nullptr; // pair stamped by a slice pass's own dispatch
    uint64_t fused = 0;                      // fused pass launches enqueued
    ~Prof()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    hipError_t begin(int k)
    {
        if (!on) return hipSuccess;
        hipEvent_t a, b;
        hipError_t e = hipEventCreate(&a);
        if (e == hipSuccess) e = hipEventCreate(&b);
        if (e != hipSuccess) return e;
        ev.push_back(a);
        ev.push_back(b);
        kind.push_back(k);
        return hipEventRecord(a, s);
    }
    hipError_t end() { return on ? hipEventRecord(ev.back(), s) : hipSuccess; }
    hipError_t kernel(int k)
    {
        ka0 = ka1 = nullptr;
        if (!on) return hipSuccess;
        hipError_t e = hipEventCreate(&ka0);
        if (e == hipSuccess) e = hipEventCreate(&ka1);
        if (e != hipSuccess) return e;
        ev.push_back(ka0);
        ev.push_back(ka1);
        kind.push_back(k);
        return hipSuccess;
    }
};

// Arguments of both slice passes of one run before a batch and slice are
// bound (msf_run.cpp).
struct PassArgs {
    msf::RowPassArgs ra{};
    msf::ColPassArgs ca{};
    const float2 *t32 = nullptr, *px = nullptr, *py = nullptr, *u0 = nullptr;
    int64_t M = 0;
    int nx = 0, ny = 0;
    // a batch of nb positions in wave buffer psi, shifts from `shift`
    void bind(float2* psi, const int2* shift, int nb, msf::RowPassArgs& r,
              msf::ColPassArgs& c) const;
    void row_at(msf::RowPassArgs& r, int n) const; // transmission, vacuum wave of slice n
    void col_at(msf::ColPassArgs& c, int n) const; // propagators of slice n
};

// ---- the units of a raster run (msf_run.cpp), in the order a run calls them
int check_run(msf_ctx* c, const msf_specimen* sp, const msf_optics* o, const msf_detector* det,
              const msf_fem* fem, const std::vector<int2>& shifts, const RunOut& out, RunFlags& f);
int prepare_groups(msf_ctx* c, Mem& m, RunPlan& pl, const msf_specimen* sp, const msf_optics* o,
                   const msf_detector* det, const msf_fem* fem, RunFlags& f);
int prepare_vacuum(msf_ctx* c, Mem& m, RunPlan& pl, const msf_specimen* sp, const msf_optics* o,
                   const RunFlags& f);
// positions per batch by the rule: the fewest batches whose wave buffer stays
// within 96 MiB, in whole groups of 16 positions
int64_t batch_rule(const RunFlags& f);
// the context's second stream and the events that order two batch chains
// (created on first use)
int ensure_stream2(msf_ctx* c);
int prepare_buffers(msf_ctx* c, Mem& m, RunPlan& pl, const RunFlags& f, const Shape& sh,
                    const std::vector<int2>& shifts, const RunOut& out, RunViews& v);
PassArgs pass_args(const RunPlan& pl, const msf_specimen* sp, const RunFlags& f, const Shape& sh,
                   const RunViews& v);
// everything after the slice loop of batch k (positions b0 .. b0+nb-1, wave
// buffer `ch`): reductions, ADF, MTF, FEM statistics (after batch k-1's when
// chains == 2), exit waves
int enqueue_batch_tail(msf_ctx* c, const RunPlan& pl, const msf_specimen* sp, const RunFlags& f,
                       const RunViews& v, const RunOut& out, Prof& prof, int chains, int64_t k,
                       int64_t b0, int nb, int ch, hipStream_t s);
int finish_outputs(msf_ctx* c, const RunPlan& pl, const msf_specimen* sp, const msf_optics* o,
                   const RunFlags& f, const RunViews& v, const RunOut& out);

// The raster run behind msf_run_stem, msf_run_stem_fem, msf_exit_waves(_shifted)
// and the allocation-free run (m.scratch()).
int run_internal(msf_ctx* c, Mem& m, const msf_specimen* sp, const msf_optics* o,
                 const msf_detector* det, const msf_fem* fem, const std::vector<int2>& shifts,
                 RunOut out);

// ---- grouped runs (msf_group.cpp)
// the caller buffers of msf_run_stem_fem's arguments
RunOut run_out(double* adf, const msf_fem_outputs* fo);
// One run of a group: arguments, outputs and, after run_group, its status.
struct GroupRun {
    // set-up before the checks, on the set-up stream (may be empty): e.g.
    // builds the run's specimen and its shifts
    std::function<int(msf_ctx*, GroupRun&)> prepare;
    const msf_specimen* sp = nullptr;
    const msf_optics* o = nullptr;
    const msf_detector* det = nullptr;
    const msf_fem* fem = nullptr;
    std::vector<int2> shifts;
    RunOut out;
    int rc = MSF_OK;
    std::string err;
};
struct GroupInfo {
    uint32_t runs = 0;     // runs that took part
    uint64_t launches = 0; // grouped slice-pass launches
    uint64_t steps = 0;    // slice steps (a row and a column launch per flag class)
};
// Runs specimens of one grid size together, their slice passes sharing
// launches; each run's rc/err tells its outcome.  Returns MSF_OK unless a
// device error ended the group.
int run_group(msf_ctx* c, std::vector<GroupRun>& runs, GroupInfo* info);
// frees the work buffers of the group's run plans (pixel groups stay)
void release_group_plans(msf_ctx* c);

// ---- serving (msf_serve.cpp)
// Blocks until no queued or running served job refers to the specimen.
void wait_unserved(msf_specimen* sp);

// ---- the allocation-free run (msf_noalloc.cpp)
// With `device` current: the address of its copy of the static scratch
// array (looked up once per device).
int scratch_attach(int device, char** out);

// ---- autotuner (msf_tune.cpp)
// "<gcnArchName>-<CUs>cu" of a device: the tuning table's device key
std::string device_arch(int device);
// Times one candidate: seconds per position-slice of a whole batch step.
using TrialFn = std::function<int(const TunedShape&, double&)>;
// The tuned shape of `key` (process-wide), timing the candidates (and the
// rule's shape) on first use.  state: 1 tuned in this call, 2 tuned earlier.
// With measure false, a shape not in the table takes the rule's (state 0).
int autotune_shape(msf_ctx* c, const std::string& key, const TunedShape& rule,
                   const std::vector<TunedShape>& cands, const TrialFn& trial, bool measure,
                   TunedShape& pick, int& state);

} // namespace msf_impl
