// msf_api.cpp -- C ABI (include/msf.h): contexts, specimen set-up, probe and
// single-stage parity entry points, the raster-run entry points and the
// serving queue.
//
// Reference seams replaced:
//   tem.cpp:944-1276 (species templates, auto-slicing, propagators,
//   transmission functions)  -> msf_specimen_build
//   TEM_NS::adfstem (adfstem.hpp:66-121)  -> msf_run_stem (msf_run.cpp)
#include "msf_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

using namespace msf_impl;

namespace {

const double kKirkland[103][12] = {
#include "kirkland_params.inc"
};

std::atomic<uint64_t> g_specimen_ids{1};

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
        if (scratch_attach(device, &c->scratch) != MSF_OK) {
            (void)hipStreamDestroy(c->stream);
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
            c->plan_na.drop_graph();
            // specimens may outlive their context: detach them so that a later
            // msf_specimen_free only releases their own device memory
            for (msf_specimen* sp : c->specimens) sp->ctx = nullptr;
            c->specimens.clear();
            c->tw_dev.clear();
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
        if (cfg->chains < 0 || cfg->chains > 3)
            return fail(c, MSF_ERR_INVALID, "chains must be 0, 1, 2 or 3");
        if (cfg->fft_plan < 0 || cfg->fft_plan > 4)
            return fail(c, MSF_ERR_INVALID, "fft_plan must be 0 .. 4");
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

int msf_copy_rate(msf_ctx* c, uint64_t bytes, uint32_t reps, double* gbs)
{
    try {
        if (!c || !gbs || reps == 0) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        const uint64_t unit = 16ull * 4 * 256;
        if (bytes < unit) return fail(c, MSF_ERR_INVALID, "copy size below 16 KiB");
        bytes -= bytes % unit;
        HIPCHK(c, hipSetDevice(c->device));
        DevBuf a, b;
        HIPCHK(c, a.alloc(bytes));
        HIPCHK(c, b.alloc(bytes));
        HIPCHK(c, hipMemsetAsync(a.p, 1, bytes, c->stream));
        HIPCHK(c, hipMemsetAsync(b.p, 0, bytes, c->stream));
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev() { if (e) (void)hipEventDestroy(e); }
        } e0, e1;
        HIPCHK(c, hipEventCreate(&e0.e));
        HIPCHK(c, hipEventCreate(&e1.e));
        const int64_t n16 = (int64_t)(bytes / 16);
        HIPCHK(c, msf::launch_stream_copy(a.as<float4>(), b.as<float4>(), n16, c->stream)); // warm-up
        float best = 0.f;
        for (uint32_t r = 0; r < reps; ++r) {
            HIPCHK(c, hipEventRecord(e0.e, c->stream));
            HIPCHK(c, msf::launch_stream_copy(a.as<float4>(), b.as<float4>(), n16, c->stream));
            HIPCHK(c, hipEventRecord(e1.e, c->stream));
            HIPCHK(c, hipEventSynchronize(e1.e));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, e0.e, e1.e));
            if (r == 0 || ms < best) best = ms;
        }
        *gbs = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
        return MSF_OK;
    }
    MSF_CATCH_ALL(c)
}

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
        {   // the column pass's view of the same set, per compact column
            const int W = msf::col_tile(nx);
            const int ncp = std::max(W, ((sp->keep.amax + sp->keep.bmax + W - 1) / W) * W);
            const msf::RowRanges ck = msf::column_keep_ranges(sp->keep, nx, ny, ncp);
            if (!ck.ok) {
                delete sp;
                return fail(c, MSF_ERR_UNSUPPORTED, "band limit is not a disk by column");
            }
            if (hipError_t e = upload(sp->dcolkeep, to_int2(ck), c->stream)) {
                (void)hipGetLastError();
                delete sp;
                return fail(c, MSF_ERR_DEVICE, hipGetErrorString(e));
            }
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
        Mem heap;
        Twiddles tx, ty;
        if ((rc = twiddles(c, heap, nx, tx))) return bail(rc);
        if ((rc = twiddles(c, heap, ny, ty))) return bail(rc);
        const double2 *twx = tx.f64, *twy = ty.f64;
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
            if (!e) e = msf::launch_finish_transmit(sig, (int64_t)k * M, ny, d.nxny, d.sqrt_nxny,
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
        wait_unserved(sp); // served runs on the specimen finish first
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
        Mem heap;
        Twiddles tx, ty;
        if (int r = twiddles(c, heap, nx, tx)) return r;
        if (int r = twiddles(c, heap, ny, ty)) return r;
        const double2 *twx = tx.f64, *twy = ty.f64;
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
        Mem heap;
        Twiddles tx, ty;
        if (int r = twiddles(c, heap, nx, tx)) return r;
        if (int r = twiddles(c, heap, ny, ty)) return r;
        const double2 *twx = tx.f64, *twy = ty.f64;
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
        Mem heap;
        DevBuf hb, nb;
        double2* h = nullptr;
        double* norm2 = nullptr;
        if (int r = probe_device(c, heap, d, o, hb, &h, nb, &norm2)) return r;
        const size_t M = (size_t)d.nx * d.ny;
        HIPCHK(c, hipMemcpy(probe, h, M * sizeof(double2), hipMemcpyDeviceToHost));
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
        Mem heap;
        DevBuf mb;
        double2* m = nullptr;
        if (int r = build_mtf2d(c, heap, d, grid->xperiod, grid->yperiod, fem, mb, &m)) return r;
        HIPCHK(c, hipMemcpy(mtf2d, m, (size_t)d.nx * d.ny * sizeof(double2), hipMemcpyDeviceToHost));
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
        DevBuf mb, w, out, keep;
        Mem heap;
        double2* m = nullptr;
        if (int r = build_mtf2d(c, heap, d, grid->xperiod, grid->yperiod, fem, mb, &m)) return r;
        Twiddles tx, ty;
        if (int r = twiddles(c, heap, nx, tx)) return r;
        if (int r = twiddles(c, heap, ny, ty)) return r;
        const double2 *twx = tx.f64, *twy = ty.f64;
        HIPCHK(c, upload(out, msf::outside_aperture(d, alpha_max), c->stream));
        HIPCHK(c, upload(keep, to_int2(msf::bwlimit_keep_ranges(d)), c->stream));
        HIPCHK(c, w.alloc(M * count * sizeof(double2)));
        double2* v = w.as<double2>();
        HIPCHK(c, hipMemcpyAsync(v, psi_k, M * count * sizeof(double2), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, msf::launch_mtf_magnitude_f64(v, (int64_t)M, (int)count, out.as<uint8_t>(), c->stream));
        HIPCHK(c, msf::fft2d_f64(v, nx, ny, (int)count, true, twx, twy, c->stream));
        HIPCHK(c, msf::launch_transmit_f64(v, m, (int64_t)M, (int)count, 0, d.sqrt_nxny,
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
        Mem heap;
        return run_internal(c, heap, sp, o, det, fem, shifts, out);
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
        const RunOut out = run_out(adf_image, fo);
        Mem heap;
        return run_internal(c, heap, sp, o, det, fem, shifts, out);
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
        Mem heap;
        return run_internal(c, heap, sp, o, nullptr, nullptr, sh, out);
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
        Mem heap;
        return run_internal(c, heap, sp, o, nullptr, nullptr, shifts, out);
    }
    MSF_CATCH_ALL(c)
}

} // extern "C"
