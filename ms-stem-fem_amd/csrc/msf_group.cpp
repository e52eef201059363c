// msf_group.cpp -- grouped runs (msf_run_group and the serving engine's job
// groups): several raster runs on specimens of one grid size share their
// slice-pass launches.  Each run keeps its own specimen, vacuum waves, pixel
// groups, batch buffers and accumulators; every step launches the row pass
// over the current batch of every run, then the column pass, with up to
// kGroupMax runs per launch (kernels_group.hip).  A run's batches, reductions
// and FEM statistics follow in raster order as in a run on its own, so its
// results equal that run's bit for bit.  No reference counterpart: the
// reference runs one job per process (adfstem.cpp:1556-2793).
#include "msf_internal.hpp"

#include <algorithm>
#include <functional>

namespace msf_impl {

namespace {

// One run's state in the step loop.
struct GroupState {
    RunFlags f;
    Shape sh;
    RunViews v;
    PassArgs pa;
    bool ok = false;
};

// A run's current batch.
struct Active {
    int r;          // run
    int64_t k, b0;  // batch index, first position
    int nb;         // positions
    int n;          // next slice
    msf::RowPassArgs ra;
    msf::ColPassArgs ca;
};

int pad8(int b) { return (b + 7) & ~7; }

} // namespace

int run_group(msf_ctx* c, std::vector<GroupRun>& runs, GroupInfo* info)
{
    const int R = (int)runs.size();
    if (R == 0) return MSF_OK;
    std::vector<GroupState> gs((size_t)R);
    while ((int)c->gplans.size() < R) c->gplans.emplace_back(new RunPlan);
    Mem heap;
    // Runs are set up one at a time on a second stream, each joining the
    // step loop as soon as it is ready: the host work of a set-up (specimen
    // build, pixel-group tables) and its device work overlap the steps of
    // the runs already going.  An event orders each run's first launch after
    // its set-up.
    hipStream_t st = c->stream; // the steps
    if (int r = ensure_stream2(c)) return r;
    hipStream_t su = c->stream2; // the set-ups
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev()
        {
            if (e) (void)hipEventDestroy(e);
        }
    } ready;
    HIPCHK(c, hipEventCreateWithFlags(&ready.e, hipEventDisableTiming));
    int nx = 0, ny = 0;
    Twiddles twx, twy;
    auto failed = [&](int r, int rc) {
        runs[(size_t)r].rc = rc;
        runs[(size_t)r].err = c->err;
    };
    // set-up of run r on the set-up stream: its specimen (if it builds one),
    // checks, pixel groups, vacuum waves, buffers
    auto setup_on = [&](int r) -> int {
        GroupRun& g = runs[(size_t)r];
        GroupState& s = gs[(size_t)r];
        RunPlan& pl = *c->gplans[(size_t)r];
        if (g.prepare)
            if (int rc = g.prepare(c, g)) return rc;
        if (int rc = check_run(c, g.sp, g.o, g.det, g.fem, g.shifts, g.out, s.f)) return rc;
        if (nx == 0) { // the first run fixes the group's lengths
            if (int rc = twiddles(c, heap, s.f.nx, twx)) return rc;
            if (int rc = twiddles(c, heap, s.f.ny, twy)) return rc;
            nx = s.f.nx;
            ny = s.f.ny;
        }
        if (s.f.nx != nx || s.f.ny != ny || g.out.exit)
            return fail(c, MSF_ERR_INVALID,
                        "the runs of a group share nx and ny and write no exit waves");
        if (int rc = prepare_groups(c, heap, pl, g.sp, g.o, g.det, g.fem, s.f)) return rc;
        s.f.want_each = s.f.want_corr && g.out.corr_each && pl.nbins * s.f.nphi > 0;
        if (int rc = prepare_vacuum(c, heap, pl, g.sp, g.o, s.f)) return rc;
        pl.ncp = s.f.ncp;
        s.v.twx = twx;
        s.v.twy = twy;
        // launch shape: the rules (the radix plans of a run on its own), one chain
        s.sh.plan_v = plan_choice(c->cfg, nx, ny, twx, twy);
        s.sh.B = std::min<int64_t>(c->cfg.batch ? (int64_t)c->cfg.batch : batch_rule(s.f), s.f.P);
        s.sh.n_batches = (s.f.P + s.sh.B - 1) / s.sh.B;
        s.sh.col_tile = c->cfg.col_tile;
        s.sh.chains = 1;
        if (int rc = prepare_buffers(c, heap, pl, s.f, s.sh, g.shifts, g.out, s.v)) return rc;
        s.pa = pass_args(pl, g.sp, s.f, s.sh, s.v);
        s.ok = true;
        return MSF_OK;
    };

    // the step loop: per step one row launch and one column launch for each
    // class of shared flags (first slice / last slice, pruning)
    Prof prof; // off
    std::vector<Active> act;
    auto start = [&](int r, int64_t k) {
        const GroupState& s = gs[(size_t)r];
        Active a{};
        a.r = r;
        a.k = k;
        a.b0 = k * s.sh.B;
        a.nb = (int)std::min<int64_t>(s.sh.B, s.f.P - a.b0);
        a.n = 0;
        s.pa.bind(s.v.psi[0], s.v.shift + a.b0, a.nb, a.ra, a.ca);
        act.push_back(a);
    };
    // sets up run r on the set-up stream and lets it join the steps
    auto join = [&](int r) -> int {
        struct OnSetup { // the context's stream is the set-up stream meanwhile
            msf_ctx* c;
            hipStream_t st;
            ~OnSetup() { c->stream = st; }
        } on{c, st};
        c->stream = su;
        const int rc = setup_on(r);
        c->stream = st;
        if (rc != MSF_OK) {
            failed(r, rc);
            gs[(size_t)r].ok = false;
            return MSF_OK; // the run fails alone
        }
        HIPCHK(c, hipEventRecord(ready.e, su));
        HIPCHK(c, hipStreamWaitEvent(st, ready.e, 0));
        start(r, 0);
        return MSF_OK;
    };
    uint64_t launches = 0, steps = 0;
    int rc = MSF_OK;
    int next = 0; // the next run to set up
    while (rc == MSF_OK) {
        // one set-up per step: its host work runs while the GPU works through
        // the steps already queued
        if (next < R) rc = join(next++);
        if (rc != MSF_OK) break;
        if (act.empty()) {
            if (next < R) continue;
            break;
        }
        ++steps;
        // row passes, classes (first slice, pruned)
        for (int cls = 0; cls < 4 && rc == MSF_OK; ++cls) {
            const bool first = (cls & 1) != 0, pr = (cls & 2) != 0;
            msf::RowGroupArgs g{};
            auto flush = [&]() -> int {
                if (g.n == 0) return MSF_OK;
                HIPCHK(c, msf::launch_row_group(ny, first, c->cfg.transmit_mode, g, st));
                ++launches;
                g = msf::RowGroupArgs{};
                return MSF_OK;
            };
            for (Active& a : act) {
                if ((a.n == 0) != first || (a.ra.prune != 0) != pr) continue;
                gs[(size_t)a.r].pa.row_at(a.ra, a.n);
                g.a[g.n] = a.ra;
                g.off[g.n + 1] = g.off[g.n] + pad8(msf::row_group_blocks(ny, a.ra));
                if (++g.n == msf::kGroupMax)
                    if ((rc = flush()) != MSF_OK) break;
            }
            if (rc == MSF_OK) rc = flush();
        }
        // column passes, classes (last slice, pruned)
        for (int cls = 0; cls < 4 && rc == MSF_OK; ++cls) {
            const bool last = (cls & 1) != 0, pr = (cls & 2) != 0;
            msf::ColGroupArgs g{};
            auto flush = [&]() -> int {
                if (g.n == 0) return MSF_OK;
                HIPCHK(c, msf::launch_col_group(nx, last, g, st));
                ++launches;
                g = msf::ColGroupArgs{};
                return MSF_OK;
            };
            for (Active& a : act) {
                const bool is_last = a.n == gs[(size_t)a.r].f.S - 1;
                if (is_last != last || (a.ca.prune != 0) != pr) continue;
                gs[(size_t)a.r].pa.col_at(a.ca, a.n);
                g.a[g.n] = a.ca;
                g.off[g.n + 1] = g.off[g.n] + pad8(msf::col_group_blocks(nx, a.ca));
                if (++g.n == msf::kGroupMax)
                    if ((rc = flush()) != MSF_OK) break;
            }
            if (rc == MSF_OK) rc = flush();
        }
        if (rc != MSF_OK) break;
        // batches past their last slice: reductions and statistics, then the
        // run's next batch
        std::vector<Active> cur;
        cur.swap(act);
        for (Active& a : cur) {
            GroupState& s = gs[(size_t)a.r];
            if (a.n + 1 < s.f.S) {
                ++a.n;
                act.push_back(a);
                continue;
            }
            GroupRun& g = runs[(size_t)a.r];
            rc = enqueue_batch_tail(c, *c->gplans[(size_t)a.r], g.sp, s.f, s.v, g.out, prof, 1, a.k,
                                    a.b0, a.nb, 0, st);
            if (rc != MSF_OK) break;
            if (a.k + 1 < s.sh.n_batches) start(a.r, a.k + 1);
        }
    }
    for (int r = 0; r < R && rc == MSF_OK; ++r) {
        const GroupState& s = gs[(size_t)r];
        if (s.ok && s.f.want_d1)
            rc = msf::launch_fem_finalize(s.v.state, c->gplans[(size_t)r]->nbins, s.v.femv, st) ==
                         hipSuccess
                     ? MSF_OK
                     : fail(c, MSF_ERR_DEVICE, "fem_finalize launch failed");
    }
    for (hipStream_t q : {su, st}) {
        const hipError_t e = hipStreamSynchronize(q);
        if (e != hipSuccess && rc == MSF_OK) {
            (void)hipGetLastError();
            rc = fail(c, MSF_ERR_DEVICE, std::string("group run: ") + hipGetErrorString(e));
        }
    }
    // a device error ends every run that was still going or not yet set up
    for (int r = 0; r < R; ++r) {
        GroupState& s = gs[(size_t)r];
        if (!s.ok) {
            if (r >= next && rc != MSF_OK) failed(r, rc);
            continue;
        }
        if (rc != MSF_OK) {
            failed(r, rc);
            continue;
        }
        GroupRun& g = runs[(size_t)r];
        if (int rf = finish_outputs(c, *c->gplans[(size_t)r], g.sp, g.o, s.f, s.v, g.out))
            failed(r, rf);
    }
    if (info) {
        info->runs = 0;
        for (const GroupState& s : gs) info->runs += s.ok ? 1 : 0;
        info->launches = launches;
        info->steps = steps;
    }
    return rc;
}

RunOut run_out(double* adf, const msf_fem_outputs* fo)
{
    RunOut out;
    out.adf = adf;
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
    return out;
}

void release_group_plans(msf_ctx* c)
{
    for (auto& p : c->gplans) p->release_work();
}

} // namespace msf_impl

using namespace msf_impl;

extern "C" {

int msf_run_group(msf_ctx* c, const msf_run_job* runs, uint32_t n_runs, int32_t* status)
{
    try {
        if (!c || (n_runs && !runs)) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        HIPCHK(c, hipSetDevice(c->device));
        std::vector<GroupRun> g(n_runs);
        std::vector<bool> pre(n_runs, false); // failed before the group started
        std::string first_err;
        for (uint32_t r = 0; r < n_runs; ++r) {
            const msf_run_job& j = runs[r];
            GroupRun& x = g[r];
            x.sp = j.specimen;
            x.o = &j.optics;
            x.det = &j.detector;
            x.fem = &j.fem;
            x.out = run_out(j.adf_image, &j.fem_out);
            int rc = MSF_OK;
            if (!j.specimen) rc = fail(c, MSF_ERR_INVALID, "specimen is NULL");
            else if (j.detector.enabled && !j.adf_image)
                rc = fail(c, MSF_ERR_INVALID, "adf_image is NULL");
            else if (!j.x_p || !j.y_p || j.n_xp == 0 || j.n_yp == 0)
                rc = fail(c, MSF_ERR_INVALID, "empty raster");
            else if (!raster_shifts(j.specimen->dom, j.x_p, j.n_xp, j.y_p, j.n_yp, x.shifts))
                rc = fail(c, MSF_ERR_DOMAIN, "probe position below the domain minimum");
            if (rc != MSF_OK) {
                x.rc = rc;
                x.err = c->err;
                x.sp = nullptr;
                pre[r] = true;
            }
        }
        // runs that failed their argument checks take no part
        std::vector<GroupRun> go;
        std::vector<uint32_t> idx;
        for (uint32_t r = 0; r < n_runs; ++r)
            if (!pre[r]) {
                go.push_back(g[r]);
                idx.push_back(r);
            }
        GroupInfo info{};
        const double t0 = now_ms();
        (void)run_group(c, go, &info);
        release_group_plans(c);
        for (size_t k = 0; k < go.size(); ++k) g[idx[k]] = go[k];
        int ret = MSF_OK;
        for (uint32_t r = 0; r < n_runs; ++r) {
            if (status) status[r] = g[r].rc;
            if (g[r].rc != MSF_OK && ret == MSF_OK) {
                ret = g[r].rc;
                first_err = "run " + std::to_string(r) + ": " + g[r].err;
            }
        }
        msf_stats st{};
        st.slice_loop_ms = now_ms() - t0;
        st.total_ms = st.slice_loop_ms;
        for (const GroupRun& x : go)
            if (x.rc == MSF_OK) st.probe_slices += (uint64_t)x.shifts.size() * x.sp->n_slices;
        st.group_runs = info.runs;
        st.group_launches = info.launches;
        st.graph = 0;
        st.chains = 1;
        c->stats = st;
        if (ret != MSF_OK) c->err = first_err;
        return ret;
    }
    MSF_CATCH_ALL(c)
}

} // extern "C"
