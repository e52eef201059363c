// msf_serve.cpp -- the serving queue (msf_submit, msf_submit_run, msf_poll,
// msf_wait, msf_serve_*): continuous batching of whole jobs on worker slots.
// No reference counterpart: the reference runs one job per process.
#include "msf_internal.hpp"

#include <algorithm>

using namespace msf_impl;

namespace {

double ms_since(double t0) { return now_ms() - t0; }

// Device bytes of a resident specimen's stacks and tables.
uint64_t specimen_bytes(const msf_specimen* sp)
{
    uint64_t b = 0;
    for (const DevBuf* d : {&sp->t32, &sp->t64, &sp->px, &sp->py, &sp->px64, &sp->py64, &sp->dkeep,
                            &sp->dcolkeep})
        b += d->bytes;
    return b;
}

// Device bytes a job claims while it runs, over-estimated by 10 %: vacuum
// waves, batch buffers of both chains and set-up temporaries, and for a job
// that builds its own specimen that specimen's stacks and the species maps
// of its build.  A job on a resident specimen claims no stacks: they are
// counted once, as resident bytes, while jobs refer to them.
uint64_t job_footprint(const ServeJob& j)
{
    const uint64_t M = (uint64_t)j.grid.nx * j.grid.ny;
    uint64_t S = 1;
    if (j.spec) {
        S = std::max<uint64_t>(1, (uint64_t)j.spec->n_slices);
    } else {
        const msf::Slices sl = msf::autoslice(j.q.data(), (uint32_t)j.Z.size(), j.grid.zmin,
                                              j.grid.zperiod, j.min_slice);
        S = std::max<uint64_t>(1, sl.lower.size());
    }
    const uint64_t P = (uint64_t)j.xp.size() * j.yp.size();
    uint64_t wave = j.cfg.batch ? (uint64_t)j.cfg.batch * M * 8 : (112ull << 20);
    wave = std::min<uint64_t>(wave, P * M * 8 + (16ull << 20));
    uint64_t b = 16 * M * S                     // u0 stack (rows doubled, FP32)
                 + 64 * M                       // probe, vacuum wave, per-pixel sums
                 + 2 * wave;                    // two chains of wave buffers
    if (!j.spec)
        b += 24 * M * S                         // t (FP32 + FP64) stacks
             + (256ull << 20);                  // species maps of the set-up
    if (j.mtf.size() && j.fem.mtf_sampling_rate != 0.0)
        b += 2 * std::min<uint64_t>(P, (wave / 8 / std::max<uint64_t>(M, 1)) + 16) * M * 16;
    return b + b / 10 + (16ull << 20);
}

// a served job's work buffers, released between jobs; the pixel groups stay
// for the next job on the same grid and settings
void reset_plan(msf_ctx* c) { c->plan.release_work(); }

// Jobs that may share slice-pass launches: one grid size, the same settings.
bool groupable(const ServeJob& a, const ServeJob& b)
{
    const msf_grid& ga = a.spec ? a.spec->grid : a.grid;
    const msf_grid& gb = b.spec ? b.spec->grid : b.grid;
    return ga.nx == gb.nx && ga.ny == gb.ny && a.cfg.transmit_mode == b.cfg.transmit_mode &&
           a.cfg.fft_plan == b.cfg.fft_plan && a.cfg.col_tile == b.cfg.col_tile &&
           a.cfg.batch == b.cfg.batch;
}

// One job on the slot's context: its specimen (unless resident), the run,
// the specimen freed.
void run_one(msf_ctx* sc, ServeJob& job, ServeDone& d, bool& built)
{
    msf_specimen* sp = nullptr;
    d.rc = msf_set_config(sc, &job.cfg);
    if (d.rc == MSF_OK && !job.spec) {
        d.rc = msf_specimen_build(sc, &job.grid, job.q.data(), job.Z.data(),
                                  (uint32_t)job.Z.size(), job.min_slice, &sp);
        built = d.rc == MSF_OK;
    }
    if (!job.mtf.empty()) {
        job.fem.mtf = job.mtf.data();
        job.fem.mtf_domain = job.mtf_dom.data();
    }
    const msf_specimen* use = job.spec ? job.spec : sp;
    if (d.rc == MSF_OK)
        d.rc = msf_run_stem_fem(sc, use, &job.optics, &job.det, &job.fem, job.xp.data(),
                                (uint32_t)job.xp.size(), job.yp.data(), (uint32_t)job.yp.size(),
                                job.adf, &job.fo);
    if (d.rc != MSF_OK) d.msg = sc->err;
    if (sp) msf_specimen_free(sp);
    std::lock_guard<std::recursive_mutex> lk(sc->mu);
    reset_plan(sc);
}

// Several jobs of one grid size as one group (run_group): each job's
// specimen is built as part of its set-up, which overlaps the slice passes of
// the jobs set up before it.
void run_grouped(msf_ctx* sc, std::vector<ServeJob>& jobs, std::vector<ServeDone>& d,
                 std::vector<bool>& built, GroupInfo& info)
{
    const size_t n = jobs.size();
    std::vector<msf_specimen*> own(n, nullptr);
    std::vector<GroupRun> runs(n);
    const int rc0 = msf_set_config(sc, &jobs[0].cfg);
    for (size_t k = 0; k < n; ++k) {
        ServeJob& job = jobs[k];
        if (!job.mtf.empty()) {
            job.fem.mtf = job.mtf.data();
            job.fem.mtf_domain = job.mtf_dom.data();
        }
        GroupRun& g = runs[k];
        g.o = &job.optics;
        g.det = &job.det;
        g.fem = &job.fem;
        g.out = run_out(job.adf, &job.fo);
        msf_specimen** slot = &own[k];
        g.prepare = [&job, slot](msf_ctx* c, GroupRun& r) -> int {
            const msf_specimen* sp = job.spec;
            if (!sp) {
                if (int e = msf_specimen_build(c, &job.grid, job.q.data(), job.Z.data(),
                                               (uint32_t)job.Z.size(), job.min_slice, slot))
                    return e;
                sp = *slot;
            }
            r.sp = sp;
            if (!raster_shifts(sp->dom, job.xp.data(), (uint32_t)job.xp.size(), job.yp.data(),
                               (uint32_t)job.yp.size(), r.shifts))
                return fail(c, MSF_ERR_DOMAIN, "probe position below the domain minimum");
            return MSF_OK;
        };
    }
    {
        std::lock_guard<std::recursive_mutex> lk(sc->mu);
        if (rc0 != MSF_OK) {
            for (GroupRun& g : runs) {
                g.rc = rc0;
                g.err = sc->err;
            }
        } else if (hipSetDevice(sc->device) == hipSuccess) {
            (void)run_group(sc, runs, &info);
        } else {
            (void)hipGetLastError();
            for (GroupRun& g : runs) {
                g.rc = MSF_ERR_DEVICE;
                g.err = "hipSetDevice failed";
            }
        }
        release_group_plans(sc);
        reset_plan(sc);
    }
    for (size_t k = 0; k < n; ++k) {
        d[k].rc = runs[k].rc;
        if (runs[k].rc != MSF_OK) d[k].msg = runs[k].err;
        built[k] = own[k] != nullptr;
    }
    for (msf_specimen* sp : own)
        if (sp) msf_specimen_free(sp);
}

void serve_worker(msf_ctx* c, uint32_t slot)
{
    ServeState& st = *c->serve;
    msf_ctx* sc = st.slot_ctx[slot];
    for (;;) {
        std::vector<ServeJob> jobs;
        uint64_t fp = 0;
        double t_start = 0;
        {
            std::unique_lock<std::mutex> lk(st.m);
            st.work_cv.wait(lk, [&] {
                if (st.queue.empty()) return st.stop;
                return st.running == 0 ||
                       st.admitted + st.resident_bytes + st.queue.front().footprint <= st.limit;
            });
            if (st.queue.empty()) return; // stop requested and nothing left
            jobs.push_back(std::move(st.queue.front()));
            st.queue.pop_front();
            fp = jobs[0].footprint;
            // grouping: further queued jobs of the head job's grid size and
            // settings that fit beside it, in queue order
            for (auto it = st.queue.begin(); it != st.queue.end() && jobs.size() < st.group_max;) {
                if (groupable(jobs[0], *it) &&
                    st.admitted + st.resident_bytes + fp + it->footprint <= st.limit) {
                    fp += it->footprint;
                    jobs.push_back(std::move(*it));
                    it = st.queue.erase(it);
                } else {
                    ++it;
                }
            }
            st.admitted += fp;
            st.running += (uint32_t)jobs.size();
            st.max_running = std::max(st.max_running, st.running);
            st.peak_admitted = std::max(st.peak_admitted, st.admitted + st.resident_bytes);
            t_start = now_ms();
        }
        st.work_cv.notify_all(); // the next head job may fit as well
        const size_t n = jobs.size();
        std::vector<ServeDone> d(n);
        std::vector<bool> built(n, false);
        GroupInfo info{};
        if (n == 1) {
            bool b = false;
            run_one(sc, jobs[0], d[0], b);
            built[0] = b;
        } else {
            run_grouped(sc, jobs, d, built, info);
        }
        const double t_end = now_ms();
        {
            std::lock_guard<std::mutex> lk(st.m);
            st.admitted -= fp;
            st.running -= (uint32_t)n;
            if (n > 1) {
                st.groups += 1;
                st.grouped_jobs += n;
                st.group_launches += info.launches;
            }
            for (size_t k = 0; k < n; ++k) {
                ServeJob& job = jobs[k];
                d[k].info.queue_ms = t_start - job.t_submit;
                d[k].info.run_ms = t_end - t_start;
                d[k].info.footprint = job.footprint;
                d[k].info.slot = slot;
                st.jobs_done += 1;
                if (built[k]) st.specimen_builds += 1;
                if (job.spec) { // the specimen stays resident while jobs refer to it
                    auto it = st.resident.find(job.spec->id);
                    if (it != st.resident.end() && --it->second.jobs == 0) {
                        st.resident_bytes -= it->second.bytes;
                        st.resident.erase(it);
                    }
                    job.spec->served.fetch_sub(1);
                }
                st.pending.erase(job.id);
                st.done[job.id] = std::move(d[k]);
            }
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
    st->group_max = std::max<uint32_t>(1, c->serve_group);
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
        sc->owner = c; // runs on the owner's resident specimens
        st->slot_ctx.push_back(sc);
    }
    c->serve = std::move(st);
    for (uint32_t k = 0; k < c->serve->slots; ++k)
        c->serve->workers.emplace_back(serve_worker, c, k);
    return MSF_OK;
}

// Settings and outputs shared by both kinds of jobs, copied at submission.
void fill_run(ServeJob& job, const msf_optics& optics, const double* x_p, uint32_t n_xp,
              const double* y_p, uint32_t n_yp, const msf_detector& det, const msf_fem& fem,
              double* adf, const msf_fem_outputs& fo)
{
    job.optics = optics;
    job.xp.assign(x_p, x_p + n_xp);
    job.yp.assign(y_p, y_p + n_yp);
    job.det = det;
    job.fem = fem;
    if (fem.mtf && fem.mtf_domain && fem.mtf_n) {
        job.mtf.assign(fem.mtf, fem.mtf + fem.mtf_n);
        job.mtf_dom.assign(fem.mtf_domain, fem.mtf_domain + fem.mtf_n);
    }
    job.fem.mtf = nullptr;
    job.fem.mtf_domain = nullptr;
    job.adf = adf;
    job.fo = fo;
}

// Takes the submitting context's settings, estimates the footprint and
// queues the job.
int queue_job(msf_ctx* c, ServeJob& job, uint64_t* id)
{
    {
        MSF_LOCK(c);
        job.cfg = c->cfg;
        job.cfg.profile_kernels = 0;
        // served jobs launch on their streams: a graph capture on one worker
        // would be invalidated by another worker's allocations
        job.cfg.no_graph = 1;
        // and measure no launch shapes (trials would add seconds to a job's
        // latency): they take the tuning table's choices or the rules
        if (job.cfg.autotune == 0) job.cfg.autotune = kAutotuneLookup;
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
        if (job.spec) {
            job.spec->served.fetch_add(1);
            auto& r = st.resident[job.spec->id];
            if (r.jobs++ == 0) {
                r.bytes = specimen_bytes(job.spec);
                st.resident_bytes += r.bytes;
            }
        }
        st.queue.push_back(std::move(job));
    }
    st.work_cv.notify_one();
    return MSF_OK;
}

} // namespace

namespace msf_impl {

// Blocks until no queued or running job of the context's serving engine
// refers to the specimen (msf_specimen_free).
void wait_unserved(msf_specimen* sp)
{
    msf_ctx* c = sp->ctx;
    if (!c || sp->served.load() == 0) return;
    ServeState* st = nullptr;
    {
        std::lock_guard<std::mutex> lk(c->serve_mu);
        st = c->serve.get();
    }
    if (!st) return;
    std::unique_lock<std::mutex> lk(st->m);
    st->done_cv.wait(lk, [&] { return sp->served.load() == 0; });
}

} // namespace msf_impl

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
        fill_run(job, j->optics, j->x_p, j->n_xp, j->y_p, j->n_yp, j->detector, j->fem,
                 j->adf_image, j->fem_out);
        return queue_job(c, job, id);
    }
    MSF_CATCH_ALL(c)
}

int msf_submit_run(msf_ctx* c, const msf_run_job* j, uint64_t* id)
{
    try {
        if (!c || !j || !id) return MSF_ERR_INVALID;
        if (!j->specimen) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
        if (j->specimen->ctx != c)
            return fail(c, MSF_ERR_INVALID, "specimen was built on another context");
        if (!j->x_p || !j->y_p || j->n_xp == 0 || j->n_yp == 0)
            return fail(c, MSF_ERR_INVALID, "empty raster");
        if (j->detector.enabled && !j->adf_image) return fail(c, MSF_ERR_INVALID, "adf_image is NULL");
        ServeJob job;
        job.spec = const_cast<msf_specimen*>(j->specimen);
        job.grid = j->specimen->grid;
        fill_run(job, j->optics, j->x_p, j->n_xp, j->y_p, j->n_yp, j->detector, j->fem,
                 j->adf_image, j->fem_out);
        return queue_job(c, job, id);
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

int msf_serve_group(msf_ctx* c, uint32_t max_jobs)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        std::lock_guard<std::mutex> lk(c->serve_mu);
        if (c->serve) return fail(c, MSF_ERR_INVALID, "the serving engine has already started");
        if (max_jobs > (uint32_t)msf::kGroupMax)
            return fail(c, MSF_ERR_INVALID,
                        "a group holds at most " + std::to_string(msf::kGroupMax) + " jobs");
        c->serve_group = max_jobs;
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
        out->specimen_builds = st.specimen_builds;
        out->resident_bytes = st.resident_bytes;
        out->groups = st.groups;
        out->grouped_jobs = st.grouped_jobs;
        out->group_launches = st.group_launches;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

uint64_t msf_job_footprint(const msf_job* j)
{
    try {
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
    } catch (...) {
        return 0;
    }
}

} // extern "C"
