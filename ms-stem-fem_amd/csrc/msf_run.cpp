// msf_run.cpp -- the raster run behind msf_run_stem (TEM_NS::adfstem,
// adfstem.hpp:66-121, body adfstem.cpp:37-4096) in separate units:
//   prepare_groups   detector / FEM pixel groups (host FP64 decisions, device copies)
//   prepare_vacuum   probe and vacuum waves of every slice (FP64 -> FP32)
//   launch_shape     batch depth, column tile, chains, radix plans (rules or the tuner)
//   prepare_buffers  batch-chain buffers and accumulators
//   enqueue_job      every launch of the job, on the streams or captured as a graph
//   finish_outputs   copies and host-side finishing of the FEM modes
// Device memory comes from a Mem: the context's cached buffers, or the
// scratch region of the allocation-free run.
#include "msf_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

namespace msf_impl {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int fail(msf_ctx* c, int code, const std::string& m)
{
    if (c) c->err = m;
    return code;
}

int take(msf_ctx* c, Mem& m, DevBuf& b, size_t n, void** out, bool* fresh)
{
    if (n == 0) n = 16;
    if (m.scratch()) {
        const size_t at = (m.used + 255) & ~size_t(255);
        if (at > m.cap || n > m.cap - at)
            return fail(c, MSF_ERR_INVALID,
                        "the run needs more than the " + std::to_string(m.cap >> 20) +
                            " MiB scratch region of the allocation-free entry point (" +
                            std::to_string((at + n) >> 20) + " MiB asked for so far); use "
                            "msf_run_stem, or fewer positions, slices or a smaller grid");
        *out = m.base + at;
        m.used = at + n;
        if (fresh) *fresh = true;
        return MSF_OK;
    }
    const bool grow = b.bytes < n || !b.p;
    if (grow) HIPCHK(c, b.alloc(n));
    *out = b.p;
    if (fresh) *fresh = grow;
    return MSF_OK;
}

int twiddles(msf_ctx* c, Mem& m, int n, Twiddles& out)
{
    auto hit = c->tw_host.find(n);
    if (hit == c->tw_host.end()) {
        std::vector<double> st;
        if (!msf::stage_twiddles(n, st)) return fail(c, MSF_ERR_UNSUPPORTED, "no plan for this length");
        TwiddleHost h;
        h.f64.resize(st.size() / 2);
        h.f32.resize(st.size() / 2);
        h.nat.resize(n);
        for (size_t e = 0; e < h.f64.size(); ++e) {
            h.f64[e] = double2{st[2 * e], st[2 * e + 1]};
            h.f32[e] = float2{(float)st[2 * e], (float)st[2 * e + 1]};
        }
        for (int e = 0; e < n; ++e) {
            const double a = 2.0 * M_PI * (double)e / (double)n;
            h.nat[e] = float2{(float)std::cos(a), (float)-std::sin(a)};
        }
        if (msf::has_alt_plan(n)) {
            std::vector<double> sa;
            if (!msf::stage_twiddles(n, sa, 1))
                return fail(c, MSF_ERR_UNSUPPORTED, "alternate plan table");
            h.alt.resize(sa.size() / 2);
            for (size_t e = 0; e < h.alt.size(); ++e)
                h.alt[e] = float2{(float)sa[2 * e], (float)sa[2 * e + 1]};
        }
        hit = c->tw_host.emplace(n, std::move(h)).first;
    }
    const TwiddleHost& h = hit->second;
    if (m.scratch()) { // copied into the region at every call (a few KB)
        DevBuf none;
        double2* f64 = nullptr;
        float2 *f32 = nullptr, *nat = nullptr, *alt = nullptr;
        if (int r = put(c, m, none, h.f64, &f64)) return r;
        if (int r = put(c, m, none, h.f32, &f32)) return r;
        if (int r = put(c, m, none, h.nat, &nat)) return r;
        if (!h.alt.empty())
            if (int r = put(c, m, none, h.alt, &alt)) return r;
        out.f64 = f64;
        out.f32 = f32;
        out.nat = nat;
        out.alt = alt;
        return MSF_OK;
    }
    auto dit = c->tw_dev.find(n);
    if (dit == c->tw_dev.end()) {
        TwiddleDev& dv = c->tw_dev[n];
        hipError_t e = upload(dv.f64, h.f64, c->stream);
        if (e == hipSuccess) e = upload(dv.f32, h.f32, c->stream);
        if (e == hipSuccess) e = upload(dv.nat, h.nat, c->stream);
        if (e == hipSuccess && !h.alt.empty()) e = upload(dv.alt, h.alt, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            c->tw_dev.erase(n);
            HIPCHK(c, e);
        }
        dit = c->tw_dev.find(n);
    }
    const TwiddleDev& dv = dit->second;
    out.f64 = dv.f64.as<double2>();
    out.f32 = dv.f32.as<float2>();
    out.nat = dv.nat.as<float2>();
    out.alt = h.alt.empty() ? nullptr : dv.alt.as<float2>();
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

int plan_choice(const msf_config& cfg, int nx, int ny, const Twiddles& twx, const Twiddles& twy)
{
    const int ax = twx.alt ? 1 : 0, ay = twy.alt ? 2 : 0;
    switch (cfg.fft_plan) {
    case 1: return 0;
    case 2: return ax + ay;
    case 3: return ay;  // alternate rows, primary columns
    case 4: return ax;  // primary rows, alternate columns
    default: return fft_plan_rule(nx) + 2 * fft_plan_rule(ny);
    }
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
int build_mtf2d(msf_ctx* c, Mem& m, const msf::Domain& d, double xperiod, double yperiod,
                const msf_fem* fem, DevBuf& ob, double2** out)
{
    const std::vector<double> t = msf::mtf_table_2d(d, xperiod, yperiod, fem->mtf,
                                                    fem->mtf_domain, (int)fem->mtf_n,
                                                    fem->mtf_sampling_rate);
    std::vector<double2> h(t.size());
    for (size_t k = 0; k < t.size(); ++k) h[k] = double2{t[k], 0.0};
    Twiddles twx, twy;
    if (int r = twiddles(c, m, d.nx, twx)) return r;
    if (int r = twiddles(c, m, d.ny, twy)) return r;
    DevBuf keep_b;
    int2* keep = nullptr;
    if (int r = put(c, m, ob, h, out)) return r;
    if (int r = put(c, m, keep_b, to_int2(msf::bwlimit_keep_ranges(d)), &keep)) return r;
    double2* v = *out;
    HIPCHK(c, msf::fft2d_f64(v, d.nx, d.ny, 1, true, twx.f64, twy.f64, c->stream));
    HIPCHK(c, msf::launch_apply_keep(v, d.nx, d.ny, 1, keep, c->stream));
    HIPCHK(c, msf::fft2d_f64(v, d.nx, d.ny, 1, false, twx.f64, twy.f64, c->stream));
    HIPCHK(c, msf::launch_div_f64(v, (int64_t)t.size(), d.nxny, c->stream));
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

int probe_device(msf_ctx* c, Mem& m, const msf::Domain& d, const msf_optics* o, DevBuf& hb,
                 double2** h, DevBuf& nb, double** norm2)
{
    const int nx = d.nx, ny = d.ny;
    const msf::RowRanges ap = msf::aperture_keep_ranges(d, o->alpha_max);
    if (!ap.ok) return fail(c, MSF_ERR_INVALID, "aperture is not a disk in FFT order");
    Twiddles twx, twy;
    if (int r = twiddles(c, m, nx, twx)) return r;
    if (int r = twiddles(c, m, ny, twy)) return r;
    DevBuf kx_b, ky_b, ap_b;
    double *kx = nullptr, *ky = nullptr;
    int2* apr = nullptr;
    if (int r = put(c, m, kx_b, d.kx, &kx)) return r;
    if (int r = put(c, m, ky_b, d.ky, &ky)) return r;
    if (int r = put(c, m, ap_b, to_int2(ap), &apr)) return r;
    if (int r = take_as<double2>(c, m, hb, (size_t)nx * ny, h)) return r;
    if (int r = take_as<double>(c, m, nb, 1, norm2)) return r;
    HIPCHK(c, msf::launch_probe_k(*h, kx, ky, nx, ny, apr, o->corrected, o->cs3, o->cs5,
                                  o->defocus, d.lambda, d.xx[0], d.yy[0], c->stream));
    HIPCHK(c, msf::fft2d_f64(*h, nx, ny, 1, true, twx.f64, twy.f64, c->stream));
    HIPCHK(c, msf::launch_norm2(*h, (int64_t)nx * ny, *norm2, c->stream));
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

bool raster_shifts(const msf::Domain& d, const double* x_p, uint32_t n_xp, const double* y_p,
                   uint32_t n_yp, std::vector<int2>& shifts)
{
    // Probe shifts in raster order (x outer, y inner), each index persisting
    // from the previous position when no grid point qualifies
    // (adfstem.cpp:1476-1477, :1590-1606).  False if a position lies below
    // the domain, where the reference's size_t index would wrap.
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

int check_run(msf_ctx* c, const msf_specimen* sp, const msf_optics* o, const msf_detector* det,
              const msf_fem* fem, const std::vector<int2>& shifts, const RunOut& out, RunFlags& f)
{
    if (!sp) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
    // a serving slot runs on its owner's resident specimens
    if (sp->ctx != c && !(c->owner && sp->ctx == c->owner))
        return fail(c, MSF_ERR_INVALID, "specimen was built on another context");
    if (int r = check_optics(c, o)) return r;
    if (shifts.empty()) return fail(c, MSF_ERR_INVALID, "empty raster");
    const msf::Domain& d = sp->dom;
    f.nx = d.nx;
    f.ny = d.ny;
    f.S = sp->n_slices;
    f.want_adf = det && det->enabled;
    f.fem_on = fem && fem->enabled;
    f.want_d1 = f.fem_on && out.fem;
    f.want_d2 = f.fem_on && out.vbar_r;
    f.want_pix = f.fem_on && (out.v_re || out.omega_vimage || out.v_gt17 || out.v_rbar ||
                              out.mean_abs || out.mean_sqr || out.var2d);
    f.want_corr = f.fem_on && (out.corr_avg || out.corr_sqr || out.corr_each);
    f.want_fem = f.want_d1 || f.want_d2 || f.want_pix || f.want_corr;
    // the detector MTF changes only what the FEM statistics see
    f.want_mtf = f.want_fem && fem->mtf && fem->mtf_domain && fem->mtf_n >= 1 &&
                 fem->mtf_sampling_rate != 0.0;
    if (f.want_adf) {
        if (det->inner < 0.0 || det->outer < 0.0)
            return fail(c, MSF_ERR_INVALID, "detector angles must be >= 0 (tem.cpp:710-720)");
        if (det->inner > d.bw_t * d.lambda)
            return fail(c, MSF_ERR_INVALID,
                        "detector inner angle falls beyond bandwidth limits (tem.cpp:721-732)");
    }
    if (f.want_fem && !(fem->dr > 0.0)) return fail(c, MSF_ERR_INVALID, "dr must be positive");
    f.P = (int64_t)shifts.size();
    f.tmode = c->cfg.transmit_mode;
    for (const int2& sh : shifts)
        if (sh.x < 0 || sh.x >= f.nx || sh.y < 0 || sh.y >= f.ny)
            return fail(c, MSF_ERR_DOMAIN, "probe shift outside the grid");
    const msf::RowRanges& keep = sp->keep;
    const int W = msf::col_tile(f.nx);
    f.ka = keep.amax;
    f.kb = keep.bmax;
    f.ncp = std::max(W, ((f.ka + f.kb + W - 1) / W) * W);
    f.M = (int64_t)f.nx * f.ny;
    f.plane = (int64_t)f.nx * f.ncp;
    return MSF_OK;
}

// Detector and FEM pixel groups.  The host tables are rebuilt when grid,
// aperture or detector/FEM settings change; their device copies (and the
// 2-D MTF) are uploaded when the tables change, and at every allocation-free
// run, whose region other calls overwrite.
int prepare_groups(msf_ctx* c, Mem& m, RunPlan& pl, const msf_specimen* sp, const msf_optics* o,
                   const msf_detector* det, const msf_fem* fem, RunFlags& f)
{
    const msf::Domain& d = sp->dom;
    // cache keys built member by member (no padding bytes from the caller);
    // the MTF table is compared by content below
    msf_detector dkey;
    msf_fem fkey;
    std::memset(&dkey, 0, sizeof dkey);
    std::memset(&fkey, 0, sizeof fkey);
    if (f.want_adf) {
        dkey.enabled = det->enabled;
        dkey.inner = det->inner;
        dkey.outer = det->outer;
        dkey.threshold_mode = det->threshold_mode;
    }
    if (f.want_fem) {
        fkey.enabled = fem->enabled;
        fkey.dr = fem->dr;
    }
    if (f.want_mtf) {
        fkey.mtf_n = fem->mtf_n;
        fkey.mtf_sampling_rate = fem->mtf_sampling_rate;
    }
    bool same_mtf = pl.mtf_on == f.want_mtf;
    if (same_mtf && f.want_mtf)
        same_mtf = std::equal(pl.mtf_v.begin(), pl.mtf_v.end(), fem->mtf) &&
                   std::equal(pl.mtf_d.begin(), pl.mtf_d.end(), fem->mtf_domain) &&
                   pl.mtf_v.size() == fem->mtf_n;
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
                             pl.fem_sqr == f.want_d2 && pl.corr == f.want_corr && same_mtf;
    if (!same_groups) {
        // the tables below are replaced one by one: until all of them and
        // their keys are written, no later call may take them as valid
        pl.groups_ok = false;
        pl.groups_dev = false;
        pl.valid = false;
        if (f.want_mtf && msf::bwlimit_keep_ranges(d).ab != sp->keep.ab) // the MTF's bw_limit must
            return fail(c, MSF_ERR_UNSUPPORTED,                           // keep the compact pixels
                        "MTF band limit differs from the slice band limit on this grid");
        double k1 = 0.0, k2 = 0.0;
        if (f.want_adf) {
            k1 = det->inner / d.lambda;
            k2 = det->outer / d.lambda;
            if (det->threshold_mode == MSF_ADF_PHYSICAL) {
                k1 = k1 * k1;
                k2 = k2 * k2;
            }
        }
        pl.hg = msf::build_groups(d, sp->keep, f.ncp, f.want_adf, k1, k2, f.want_fem,
                                  f.want_fem ? fem->dr : 1.0, o->alpha_max, f.want_d2, f.want_corr);
        pl.st = msf::build_segments(pl.hg, 32);
        pl.G = (int)pl.hg.kind.size();
        pl.nbins = pl.hg.nbins;
        pl.n_tiles = pl.st.n_tiles;
        msf::combine_split(pl.st.seg_of_group, pl.small_ids, pl.large_ids);
        pl.fem_sqr = f.want_d2;
        pl.corr = f.want_corr;
        pl.cap_b = 0; // partial-sum buffers depend on G
        pl.mtf_on = f.want_mtf;
        pl.mtf_v.clear();
        pl.mtf_d.clear();
        if (f.want_mtf) {
            pl.mtf_v.assign(fem->mtf, fem->mtf + fem->mtf_n);
            pl.mtf_d.assign(fem->mtf_domain, fem->mtf_domain + fem->mtf_n);
            pl.mtf_rate = fem->mtf_sampling_rate;
        }
        pl.ggrid = gkey;
        pl.galpha = o->alpha_max;
        pl.det = dkey;
        pl.fem = fkey;
        pl.groups_ok = true;
    }
    if (!pl.groups_dev || m.scratch()) {
        pl.groups_dev = false;
        if (int r = put(c, m, pl.seg_pix_b, pl.st.pix, &pl.seg_pix)) return r;
        if (int r = put(c, m, pl.tile_off_b, pl.st.tile_off, &pl.tile_off)) return r;
        if (int r = put(c, m, pl.seg_kind_b, pl.st.seg_kind, &pl.seg_kind)) return r;
        if (int r = put(c, m, pl.seg_of_group_b, pl.st.seg_of_group, &pl.seg_of_group)) return r;
        if (int r = put(c, m, pl.grp_small_b, pl.small_ids, &pl.grp_small)) return r;
        if (int r = put(c, m, pl.grp_large_b, pl.large_ids, &pl.grp_large)) return r;
        if (int r = put(c, m, pl.cnt_b, pl.hg.counts, &pl.cnt)) return r;
        if (int r = put(c, m, pl.zero_b, pl.hg.zeroed, &pl.zero)) return r;
        if (f.want_mtf) {
            if (int r = build_mtf2d(c, m, d, sp->grid.xperiod, sp->grid.yperiod, fem, pl.mtf2d_b,
                                    &pl.mtf2d))
                return r;
            if (int r = put(c, m, pl.mtf_out_b, msf::outside_aperture(d, o->alpha_max), &pl.mtf_out))
                return r;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream)); // the host tables may change next call
        pl.groups_dev = true;
    }
    f.nphi = f.want_corr ? pl.hg.nphi : 0;
    f.ncorr = (size_t)std::max(pl.nbins * f.nphi, 1);
    return MSF_OK;
}

// Vacuum wave (t = 1) of every slice, FP64, from the probe of
// adfstem.cpp:586-817: u0_0 = s * A_p * IFFT(H)/s; then per slice
// w = FFT(u0_n) / s^2 * M P_n (adfstem.cpp:1986-2043), u0_{n+1} = IFFT(w).
// Kept while specimen and optics are unchanged (ordinary runs).
int prepare_vacuum(msf_ctx* c, Mem& m, RunPlan& pl, const msf_specimen* sp, const msf_optics* o,
                   const RunFlags& f)
{
    const bool same_optics = pl.valid && pl.spec_id == sp->id &&
                             std::memcmp(&pl.optics, o, sizeof(msf_optics)) == 0 && !m.scratch();
    if (same_optics) return MSF_OK;
    pl.valid = false;
    const msf::Domain& d = sp->dom;
    const int nx = f.nx, ny = f.ny, S = f.S;
    const int64_t M = f.M;
    DevBuf hb, nb;
    double2* v = nullptr;
    double* norm2 = nullptr;
    if (int r = probe_device(c, m, d, o, hb, &v, nb, &norm2)) return r;
    Twiddles twx, twy;
    if (int r = twiddles(c, m, nx, twx)) return r;
    if (int r = twiddles(c, m, ny, twy)) return r;
    if (int r = take_as<float2>(c, m, pl.U0, (size_t)S * 2 * M, &pl.u0)) return r;
    if (int r = take_as<float2>(c, m, pl.Psi0, (size_t)M, &pl.psi0)) return r;
    HIPCHK(c, msf::launch_scale_f64(v, M, norm2, d.sqrt_nxny, c->stream));
    for (int n = 0; n < S; ++n) {
        HIPCHK(c, msf::launch_to_f32_shift_pairs(v, nx, ny, pl.u0 + (int64_t)n * 2 * M, c->stream));
        HIPCHK(c, msf::fft2d_f64(v, nx, ny, 1, false, twx.f64, twy.f64, c->stream));
        HIPCHK(c, msf::launch_vacuum_step(v, nx, ny, sp->px64.as<double2>() + (int64_t)n * nx,
                                          sp->py64.as<double2>() + (int64_t)n * ny,
                                          sp->dkeep.as<int2>(), 1.0 / (d.sqrt_nxny * d.sqrt_nxny),
                                          c->stream));
        if (n + 1 < S) HIPCHK(c, msf::fft2d_f64(v, nx, ny, 1, true, twx.f64, twy.f64, c->stream));
    }
    HIPCHK(c, msf::launch_to_f32_transposed(v, nx, ny, pl.psi0, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    pl.spec_id = sp->id;
    pl.optics = *o;
    pl.valid = !m.scratch();
    return MSF_OK;
}

// Device bytes one position of one batch chain needs.
int64_t chain_bytes_per_position(const RunPlan& pl, const RunFlags& f)
{
    int64_t b = f.plane * (int64_t)sizeof(float2) + (int64_t)pl.G * 8 +
                (int64_t)std::max(pl.n_tiles, 1) * 64 * 8;
    if (f.want_mtf) b += f.M * (int64_t)sizeof(double2);
    return b;
}

// Runs smaller than this many probe-slices use the rules unless the
// autotuner is asked for (autotune = 1): tuning a shape costs about 20 trial
// batch steps.
constexpr int64_t kTuneMinWork = 16384;

int tune_shape(msf_ctx* c, RunPlan& pl, const msf_specimen* sp, const std::vector<int2>& shifts,
               const RunFlags& f, const Twiddles& twx, const Twiddles& twy, int64_t B,
               int plan_rule, int all_alt, TunedShape& pick, int& state);

int64_t batch_rule(const RunFlags& f)
{
    const int64_t P = f.P;
    const int64_t per_pos = f.plane * (int64_t)sizeof(float2);
    int64_t bmax = std::max<int64_t>(1, (96ll << 20) / per_pos);
    if (bmax >= 32) bmax -= bmax % 16; // whole groups of 16 positions per launch
    const int64_t nbat = (P + bmax - 1) / bmax;
    int64_t B = (P + nbat - 1) / nbat;
    if (B > 16 && B % 16) B = std::min<int64_t>(bmax, B + 16 - B % 16);
    return std::max<int64_t>(1, std::min<int64_t>(B, P));
}

// Batch depth, column tile, chains and radix plans: configured, by the rules
// below or by the autotuner.
int launch_shape(msf_ctx* c, Mem& m, RunPlan& pl, const msf_specimen* sp,
                 const std::vector<int2>& shifts, const RunOut& out, const RunFlags& f,
                 const Twiddles& twx, const Twiddles& twy, Shape& sh)
{
    const int64_t P = f.P;
    // radix plans as vx + 2 vy (per axis 0 primary, 1 alternate): configured
    // (1 primary, 2 alternate where the length has one) or by the rule
    const int all_alt = (twx.alt ? 1 : 0) + (twy.alt ? 2 : 0);
    const int plan_rule = plan_choice(c->cfg, f.nx, f.ny, twx, twy);

    // batch depth: the fewest batches whose wave buffer stays within 96 MiB,
    // spread evenly over the raster.  The buffer, the t and u0 slices and the
    // partial sums then stay resident in the 256 MB Infinity Cache across the
    // two passes of a slice (512^2: 64 positions per batch, measured best of
    // 64..256).
    int64_t B = c->cfg.batch ? (int64_t)c->cfg.batch : batch_rule(f);
    B = std::max<int64_t>(1, std::min<int64_t>(B, P));
    sh.col_tile = c->cfg.col_tile;
    int chains_cfg = c->cfg.chains;
    sh.plan_v = plan_rule;
    sh.tuned_state = 0;
    const bool prof = c->cfg.profile_kernels != 0 && !m.scratch();
    // Autotuner: with batch, column tile and chains automatic, the launch
    // shape of this job's shape comes from the process-wide tuning table,
    // measured on first use (autotune 0: runs of at least kTuneMinWork
    // probe-slices, shapes only, results bit-identical to the rules;
    // 1: every run of two or more positions, and with fft_plan 0 the radix
    // plans too; -1: the rules).
    // (kAutotuneLookup, served jobs: the table's choice, no trials)
    const int amode = c->cfg.autotune;
    if (amode >= 0 && c->cfg.batch == 0 && c->cfg.col_tile == 0 && c->cfg.chains == 0 && !prof &&
        !out.exit && P >= 2 && !m.scratch() &&
        (amode == 1 || P * (int64_t)f.S >= kTuneMinWork)) {
        TunedShape pick{};
        if (int r = tune_shape(c, pl, sp, shifts, f, twx, twy, B, plan_rule, all_alt, pick,
                               sh.tuned_state))
            return r;
        B = std::min<int64_t>(pick[0], P);
        sh.col_tile = pick[1];
        chains_cfg = pick[2];
        sh.plan_v = pick[3];
    }
    // Two chains of batches in flight on two streams: the row passes (VALU
    // bound) of one chain overlap the column passes (memory bound) of the
    // other.  Automatic: two chains, launched on the streams, once a launch
    // covers >= 4M pixels (512^2: +4 % over one chain, graph or not; the
    // runtime replays a two-branch graph's branches one after the other);
    // below that one chain replayed as a graph, where launch cost matters.
    // Profiled, exit-wave and allocation-free runs use one chain.
    if (m.scratch()) {
        // what is left of the region after the per-position and fixed buffers
        int64_t fixed = P * (int64_t)(sizeof(double) + sizeof(int2)) +
                        (int64_t)std::max(pl.nbins, 1) * 5 * 8 + (int64_t)f.ncorr * 16 + 4096;
        if (f.want_pix) fixed += 2 * f.plane * 8;
        if (f.want_each) fixed += P * (int64_t)pl.nbins * f.nphi * 8;
        const int64_t per = chain_bytes_per_position(pl, f);
        const int64_t room = (int64_t)m.left() - fixed - 8 * 256;
        const int64_t bfit = room > 0 ? room / per : 0;
        if (bfit < 1)
            return fail(c, MSF_ERR_INVALID,
                        "the run needs more than the " + std::to_string(m.cap >> 20) +
                            " MiB scratch region of the allocation-free entry point; use "
                            "msf_run_stem, or fewer positions, slices or a smaller grid");
        if (B > bfit) B = bfit;
    }
    sh.B = B;
    sh.n_batches = (P + B - 1) / B;
    const bool big_launch = B * f.plane >= (4ll << 20);
    const int chains_req = chains_cfg ? chains_cfg : (big_launch ? 2 : 1);
    // chains 3: batches in pairs on one stream, row and column passes of the
    // pair's two batches fused into one launch per step
    sh.chains = (chains_req >= 2 && sh.n_batches >= 2 && !prof && !out.exit && !m.scratch())
                    ? chains_req
                    : 1;
    sh.auto_eager = c->cfg.chains == 0 && sh.chains == 2;
    return MSF_OK;
}

int ensure_stream2(msf_ctx* c)
{
    if (!c->stream2) HIPCHK(c, hipStreamCreate(&c->stream2)); // default priority, as the first
    for (hipEvent_t* e : {&c->ev_fork, &c->ev_join, &c->ev_fem[0], &c->ev_fem[1]})
        if (!*e) HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    return MSF_OK;
}

// Batch-chain buffers (grown as needed and kept for the next ordinary run)
// and the run's accumulators, zeroed outside any captured graph.
int prepare_buffers(msf_ctx* c, Mem& m, RunPlan& pl, const RunFlags& f, const Shape& sh,
                    const std::vector<int2>& shifts, const RunOut& out, RunViews& v)
{
    const int64_t B = sh.B, P = f.P;
    const int nbins = pl.nbins;
    if (B > pl.cap_b) { // sizes depend on G, which changes with the groups
        for (int k = 0; k < 2; ++k)
            for (DevBuf* b : {&pl.psi_b[k], &pl.part_b[k], &pl.segp_b[k], &pl.mtf_full_b[k]})
                b->release();
        pl.exit_b.release();
        pl.cap_b = B;
    }
    for (int k = 0; k < std::min(sh.chains, 2); ++k) {
        bool fresh = false;
        if (int r = take_as<float2>(c, m, pl.psi_b[k], (size_t)B * f.plane, &v.psi[k], &fresh))
            return r;
        if (fresh) HIPCHK(c, hipMemsetAsync(v.psi[k], 0, (size_t)B * f.plane * sizeof(float2), c->stream));
        if (int r = take_as<double>(c, m, pl.part_b[k], (size_t)B * pl.G, &v.part[k])) return r;
        if (int r = take_as<double>(c, m, pl.segp_b[k], (size_t)B * std::max(pl.n_tiles, 1) * 64,
                                    &v.segp[k]))
            return r;
        if (f.want_mtf)
            if (int r = take_as<double2>(c, m, pl.mtf_full_b[k], (size_t)B * f.M, &v.mtf_full[k]))
                return r;
    }
    if (int r = take_as<double>(c, m, pl.adf_b, (size_t)P, &v.adf)) return r;
    if (int r = take_as<int2>(c, m, pl.shift_b, (size_t)P, &v.shift)) return r;
    const size_t nb1 = (size_t)std::max(nbins, 1);
    if (int r = take_as<double>(c, m, pl.state_b, nb1 * 3, &v.state)) return r;
    if (int r = take_as<double>(c, m, pl.femv_b, nb1, &v.femv)) return r;
    if (int r = take_as<double>(c, m, pl.vsum_b, nb1, &v.vsum)) return r;
    if (f.want_each)
        if (int r = take_as<double>(c, m, pl.ceach_b, (size_t)P * nbins * f.nphi, &v.ceach))
            return r;
    if (f.want_corr) {
        if (int r = take_as<double>(c, m, pl.csum_b, f.ncorr, &v.csum)) return r;
        if (int r = take_as<double>(c, m, pl.csqr_b, f.ncorr, &v.csqr)) return r;
    }
    if (f.want_pix) {
        if (int r = take_as<double>(c, m, pl.pm1_b, (size_t)f.plane, &v.pm1)) return r;
        if (int r = take_as<double>(c, m, pl.pm2_b, (size_t)f.plane, &v.pm2)) return r;
    }
    if (out.exit)
        if (int r = take_as<float2>(c, m, pl.exit_b, (size_t)B * f.M, &v.exit)) return r;
    // accumulators start at zero (outside any captured graph)
    HIPCHK(c, hipMemsetAsync(v.state, 0, nb1 * 3 * sizeof(double), c->stream));
    if (f.want_d2) HIPCHK(c, hipMemsetAsync(v.vsum, 0, nb1 * sizeof(double), c->stream));
    if (f.want_corr) {
        HIPCHK(c, hipMemsetAsync(v.csum, 0, f.ncorr * sizeof(double), c->stream));
        HIPCHK(c, hipMemsetAsync(v.csqr, 0, f.ncorr * sizeof(double), c->stream));
    }
    if (f.want_pix) {
        HIPCHK(c, hipMemsetAsync(v.pm1, 0, (size_t)f.plane * sizeof(double), c->stream));
        HIPCHK(c, hipMemsetAsync(v.pm2, 0, (size_t)f.plane * sizeof(double), c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(v.shift, shifts.data(), (size_t)P * sizeof(int2),
                             hipMemcpyHostToDevice, c->stream));
    if (sh.chains == 2)
        if (int r = ensure_stream2(c)) return r;
    return MSF_OK;
}

PassArgs pass_args(const RunPlan& pl, const msf_specimen* sp, const RunFlags& f, const Shape& sh,
                   const RunViews& v)
{
    const msf::Domain& d = sp->dom;
    PassArgs p;
    msf::RowPassArgs& ra = p.ra;
    ra.nx = f.nx;
    ra.ncp = f.ncp;
    ra.ka = f.ka;
    ra.kb = f.kb;
    ra.plan = (sh.plan_v >> 1) & 1;
    ra.tw = ra.plan ? v.twy.alt : v.twy.f32;
    ra.inv_s = (float)(1.0 / d.sqrt_nxny);
    ra.prune = msf::row_prunable(f.ny, f.ka, f.kb) ? 1 : 0;
    msf::ColPassArgs& ca = p.ca;
    ca.ny = f.ny;
    ca.ncp = f.ncp;
    ca.ka = f.ka;
    ca.kb = f.kb;
    ca.colkeep = sp->dcolkeep.as<int2>();
    ca.plan = sh.plan_v & 1;
    ca.tw = ca.plan ? v.twx.alt : v.twx.f32;
    ca.psi0 = pl.psi0;
    ca.twx_nat = v.twx.nat;
    ca.twy = v.twy.nat;
    ca.tile = sh.col_tile;
    ca.prune = msf::col_prunable(f.nx, sp->keep.ab.data()) ? 1 : 0;
    p.t32 = sp->t32.as<float2>();
    p.px = sp->px.as<float2>();
    p.py = sp->py.as<float2>();
    p.u0 = pl.u0;
    p.M = f.M;
    p.nx = f.nx;
    p.ny = f.ny;
    return p;
}

void PassArgs::bind(float2* psi, const int2* shift, int nb, msf::RowPassArgs& r,
                    msf::ColPassArgs& c) const
{
    r = ra;
    c = ca;
    r.psi = c.psi = psi;
    r.batch = c.batch = nb;
    r.shift = c.shift = shift;
}

void PassArgs::row_at(msf::RowPassArgs& r, int n) const
{
    r.t = t32 + (int64_t)n * M;
    r.u0 = u0 + (int64_t)n * 2 * M;
}

void PassArgs::col_at(msf::ColPassArgs& c, int n) const
{
    c.px = px + (int64_t)n * nx;
    c.py = py + (int64_t)n * ny;
}

int enqueue_batch_tail(msf_ctx* c, const RunPlan& pl, const msf_specimen* sp, const RunFlags& f,
                       const RunViews& v, const RunOut& out, Prof& prof, int chains, int64_t k,
                       int64_t b0, int nb, int ch, hipStream_t s)
{
    const msf::Domain& d = sp->dom;
    const int nx = f.nx, ny = f.ny, ka = f.ka, kb = f.kb, ncp = f.ncp;
    const int64_t M = f.M, plane = f.plane;
    const int G = pl.G, nbins = pl.nbins, nphi = f.nphi;
    float2* bpsi = v.psi[ch];
    double* bpart = v.part[ch];
    double* bsegp = v.segp[ch];
    auto combine = [&](const double* sg, int n, double* out, hipStream_t st) {
        return msf::launch_combine_segments(sg, pl.n_tiles * 64, pl.seg_of_group, pl.grp_small,
                                            (int)pl.small_ids.size(), pl.grp_large,
                                            (int)pl.large_ids.size(), G, n, out, st);
    };
    HIPCHK(c, prof.begin(K_RED));
    HIPCHK(c, msf::launch_reduce_segments(bpsi, plane, pl.seg_pix, pl.tile_off, pl.seg_kind,
                                          pl.n_tiles, nb, bsegp, s));
    HIPCHK(c, combine(bsegp, nb, bpart, s));
    HIPCHK(c, prof.end());
    if (f.want_adf) HIPCHK(c, msf::launch_collect_adf(bpart, G, nb, v.adf + b0, s));
    if (f.want_mtf) {
        // adfstem.cpp:2272-2385 after the ADF sum: |psi| outside the
        // aperture, BACKWARD FFT, x MTF / s (in-place update), FORWARD
        // FFT, / s, bw_limit; then the group sums again for the FEM
        double2* full = v.mtf_full[ch];
        HIPCHK(c, msf::launch_mtf_magnitude(bpsi, nx, ny, ncp, ka, kb, pl.mtf_out, nb, full, s));
        HIPCHK(c, msf::fft2d_f64(full, nx, ny, nb, true, v.twx.f64, v.twy.f64, s));
        HIPCHK(c, msf::launch_transmit_f64(full, pl.mtf2d, M, nb, 0, d.sqrt_nxny, s));
        HIPCHK(c, msf::fft2d_f64(full, nx, ny, nb, false, v.twx.f64, v.twy.f64, s));
        HIPCHK(c, msf::launch_mtf_compact(full, nx, ny, ncp, ka, kb, sp->dkeep.as<int2>(),
                                          d.sqrt_nxny, nb, bpsi, s));
        HIPCHK(c, msf::launch_reduce_segments(bpsi, plane, pl.seg_pix, pl.tile_off,
                                              pl.seg_kind, pl.n_tiles, nb, bsegp, s));
        HIPCHK(c, combine(bsegp, nb, bpart, s));
    }
    // the FEM accumulators are shared: batch k updates them after batch k-1
    if (chains == 2 && k > 0) HIPCHK(c, hipStreamWaitEvent(s, c->ev_fem[(k - 1) % 2], 0));
    if (f.want_d1)
        HIPCHK(c, msf::launch_fem_welford(bpart, G, nb, nbins, pl.cnt, pl.zero, v.state, s));
    if (f.want_d2)
        HIPCHK(c, msf::launch_fem_vbar(bpart, G, nb, nbins, pl.hg.fem_abs, pl.hg.fem_sqr, pl.cnt,
                                       pl.zero, v.vsum, s));
    if (f.want_corr)
        HIPCHK(c, msf::launch_fem_correlograph(
                      bpart, G, nb, nbins, nphi, pl.hg.fem_abs, pl.hg.corr, pl.cnt, pl.zero,
                      v.csum, v.csqr, f.want_each ? v.ceach + b0 * nbins * nphi : nullptr,
                      s));
    if (f.want_pix) HIPCHK(c, msf::launch_pixel_moments(bpsi, plane, nb, v.pm1, v.pm2, s));
    if (chains == 2) HIPCHK(c, hipEventRecord(c->ev_fem[k % 2], s));
    if (out.exit) {
        HIPCHK(c, msf::launch_expand_exit(bpsi, nx, ny, ncp, ka, kb, nb, v.exit, s));
        HIPCHK(c, hipMemcpyAsync(out.exit + (size_t)b0 * nx * ny * 2, v.exit,
                                 (size_t)nb * nx * ny * sizeof(float2), hipMemcpyDeviceToHost,
                                 s));
    }
    return MSF_OK;
}

// Every launch of the job (all batches, reductions, FEM statistics), issued
// on the streams in raster order.
int enqueue_job(msf_ctx* c, const RunPlan& pl, const msf_specimen* sp, const RunFlags& f,
                const Shape& sh, const RunViews& v, const RunOut& out, Prof& prof)
{
    const int nx = f.nx, ny = f.ny, S = f.S;
    const int64_t P = f.P, B = sh.B;
    const int chains = sh.chains;
    const PassArgs pa = pass_args(pl, sp, f, sh, v);
    hipStream_t streams[2] = {c->stream, c->stream2};
    auto bind = [&](int64_t b0, int nb, int ch, msf::RowPassArgs& r, msf::ColPassArgs& cc) {
        pa.bind(v.psi[ch], v.shift + b0, nb, r, cc);
    };
    auto row_at = [&](msf::RowPassArgs& r, int n) { pa.row_at(r, n); };
    auto col_at = [&](msf::ColPassArgs& cc, int n) { pa.col_at(cc, n); };
    auto finish_batch = [&](int64_t k, int64_t b0, int nb, int ch, hipStream_t s) -> int {
        return enqueue_batch_tail(c, pl, sp, f, v, out, prof, chains, k, b0, nb, ch, s);
    };
    // the slice loop of one batch alone
    auto slices_alone = [&](msf::RowPassArgs& r, msf::ColPassArgs& cc, hipStream_t s) -> int {
        for (int n = 0; n < S; ++n) {
            row_at(r, n);
            HIPCHK(c, prof.kernel(n == 0 ? K_ROW0 : K_ROW));
            HIPCHK(c, msf::launch_row_pass(ny, n == 0, f.tmode, r, s, prof.ka0, prof.ka1));
            col_at(cc, n);
            HIPCHK(c, prof.kernel(n == S - 1 ? K_COLL : K_COL));
            HIPCHK(c, msf::launch_col_pass(nx, n == S - 1, cc, s, prof.ka0, prof.ka1));
        }
        return MSF_OK;
    };
    if (chains == 3) {
        // Paired batches on one stream: the row pass of one batch and the
        // column pass of the other share a launch (fused_pass_kernel), so the
        // VALU-bound row workgroups and the memory-bound column workgroups of
        // the two batches run side by side on every CU.  Per pair (A, B):
        // A.row0, then (A.col n + B.row n), (A.row n+1 + B.col n), ...,
        // B.col S-1; each batch still sees its passes in slice order, and the
        // reductions and FEM statistics follow in raster order.
        hipStream_t s = c->stream;
        auto pair = [&](bool first, bool last, const msf::RowPassArgs& r,
                        const msf::ColPassArgs& cc) -> int {
            if (nx == ny && msf::fused_supported(nx, r, cc)) {
                HIPCHK(c, msf::launch_fused_pass(nx, first, last, f.tmode, r, cc, s));
                prof.fused += 1;
            } else {
                HIPCHK(c, msf::launch_row_pass(ny, first, f.tmode, r, s));
                HIPCHK(c, msf::launch_col_pass(nx, last, cc, s));
            }
            return MSF_OK;
        };
        int64_t k = 0;
        for (int64_t b0 = 0; b0 < P;) {
            const int nbA = (int)std::min<int64_t>(B, P - b0);
            const int64_t b1 = b0 + nbA;
            msf::RowPassArgs rA, rB;
            msf::ColPassArgs cA, cB;
            bind(b0, nbA, 0, rA, cA);
            if (b1 >= P) { // an odd last batch
                if (int r = slices_alone(rA, cA, s)) return r;
                if (int r = finish_batch(k, b0, nbA, 0, s)) return r;
                break;
            }
            const int nbB = (int)std::min<int64_t>(B, P - b1);
            bind(b1, nbB, 1, rB, cB);
            row_at(rA, 0);
            HIPCHK(c, msf::launch_row_pass(ny, true, f.tmode, rA, s));
            for (int n = 0; n < S; ++n) {
                row_at(rB, n);
                col_at(cA, n);
                if (int r = pair(n == 0, n == S - 1, rB, cA)) return r;
                col_at(cB, n);
                if (n + 1 < S) {
                    row_at(rA, n + 1);
                    if (int r = pair(false, false, rA, cB)) return r;
                } else {
                    HIPCHK(c, msf::launch_col_pass(nx, true, cB, s));
                }
            }
            if (int r = finish_batch(k, b0, nbA, 0, s)) return r;
            if (int r = finish_batch(k + 1, b1, nbB, 1, s)) return r;
            k += 2;
            b0 = b1 + nbB;
        }
        if (f.want_d1) HIPCHK(c, msf::launch_fem_finalize(v.state, pl.nbins, v.femv, c->stream));
        return MSF_OK;
    }
    if (chains == 2) { // fork: the second stream starts after everything issued so far
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t b0 = 0; b0 < P; b0 += B, ++k) {
        const int nb = (int)std::min<int64_t>(B, P - b0);
        const int ch = (int)(k % chains);
        hipStream_t s = streams[ch];
        msf::RowPassArgs r;
        msf::ColPassArgs cc;
        bind(b0, nb, ch, r, cc);
        if (int e = slices_alone(r, cc, s)) return e;
        if (int e = finish_batch(k, b0, nb, ch, s)) return e;
    }
    if (chains == 2) { // join: the first stream waits for everything on the second
        HIPCHK(c, hipEventRecord(c->ev_join, c->stream2));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    if (f.want_d1) HIPCHK(c, msf::launch_fem_finalize(v.state, pl.nbins, v.femv, c->stream));
    return MSF_OK;
}

// Graph replay: the signature lists everything the captured kernel
// arguments depend on (job shape, device addresses); any change re-captures.
std::vector<uint64_t> graph_signature(const RunPlan& pl, const msf_specimen* sp, const RunFlags& f,
                                      const Shape& sh, const RunViews& v, const RunOut& out)
{
    auto u = [](const void* p) { return (uint64_t)(uintptr_t)p; };
    return {sp->id, (uint64_t)f.P, (uint64_t)sh.B, (uint64_t)f.tmode, (uint64_t)f.want_adf,
            (uint64_t)f.want_d1 | (uint64_t)f.want_d2 << 1 | (uint64_t)f.want_pix << 2 |
                (uint64_t)f.want_corr << 3,
            u(v.csum), u(v.csqr), (uint64_t)f.nphi, (uint64_t)pl.hg.corr, (uint64_t)f.want_mtf,
            u(pl.mtf2d), (uint64_t)f.want_each, u(v.ceach),
            (uint64_t)sh.chains, u(v.psi[1]), u(v.part[1]), u(v.segp[1]), u(v.mtf_full[1]),
            u(pl.mtf_out), u(v.mtf_full[0]), u(v.twx.f64), u(v.twy.f64), u(v.vsum), u(v.pm1),
            u(v.pm2), (uint64_t)pl.hg.fem_abs, (uint64_t)pl.hg.fem_sqr, (uint64_t)f.plane,
            u(v.psi[0]), u(v.part[0]), u(v.segp[0]), u(v.adf), u(v.shift), u(v.state), u(v.femv),
            u(pl.u0), u(pl.psi0), u(pl.seg_pix), u(pl.tile_off), u(pl.seg_kind), u(pl.grp_small),
            u(pl.grp_large), (uint64_t)pl.small_ids.size(),
            u(pl.seg_of_group), u(pl.cnt), u(pl.zero), u(v.twx.f32), u(v.twx.alt), u(v.twy.f32),
            u(v.twy.alt), u(v.twx.nat), u(v.twy.nat), (uint64_t)f.ncp, (uint64_t)pl.G,
            (uint64_t)pl.nbins, (uint64_t)pl.n_tiles, (uint64_t)sh.col_tile, (uint64_t)sh.plan_v};
}

// square grids whose slice passes can run as fused pass pairs (chains 3)
static bool fused_shape(const RunFlags& f, int tile, int plan_v)
{
    return f.nx == f.ny && msf::fused_shape_ok(f.nx, (plan_v >> 1) & 1, plan_v & 1, tile) &&
           msf::row_prunable(f.ny, f.ka, f.kb);
}

// The autotuner's candidates and trials.  A trial runs the job's own launch
// sequence (enqueue_job: both slice passes of every slice, the reductions and
// FEM statistics, with the candidate's chains) over the first four batches
// of the job's positions on trial buffers, once untimed and once timed; its
// cost is seconds per position-slice.  Candidates: the rule's batch depth,
// half and twice it; column tiles 4, 8 and 16; one chain, two chains, or
// batch pairs with fused pass launches where the grid has them; with
// autotune = 1 and fft_plan = 0 also the all-primary and all-alternate radix
// plans.  Every candidate of one radix plan gives bit-identical results.

int tune_shape(msf_ctx* c, RunPlan& pl, const msf_specimen* sp, const std::vector<int2>& shifts,
               const RunFlags& f, const Twiddles& twx, const Twiddles& twy, int64_t B,
               int plan_rule, int all_alt, TunedShape& pick, int& state)
{
    const int64_t P = f.P;
    if (c->arch.empty()) c->arch = device_arch(c->device);
    const bool plans = c->cfg.autotune == 1 && c->cfg.fft_plan == 0;
    const bool measure = c->cfg.autotune != kAutotuneLookup;
    std::ostringstream key;
    key << c->arch << " " << f.nx << " " << f.ny << " " << f.ncp << " " << std::min<int64_t>(P, 256)
        << " " << (f.want_adf ? "a" : "") << (f.want_d1 ? "v" : "") << (f.want_d2 ? "b" : "")
        << (f.want_pix ? "p" : "") << (f.want_corr ? "c" : "") << (f.want_mtf ? "m" : "") << "-"
        << " " << (plans ? "plans" : "plan" + std::to_string(plan_rule));
    const int tile0 = msf::col_tile_default(f.nx, plan_rule & 1);
    const int ch0 = (B * f.plane >= (4ll << 20) && (P + B - 1) / B >= 2) ? 2 : 1;
    const TunedShape rule = {(int)B, tile0, ch0, plan_rule};
    std::vector<int64_t> bs = {B};
    for (int64_t b : {B / 2, 2 * B}) {
        if (B >= 32) b -= b % 16;
        if (b >= 1 && b <= P && std::find(bs.begin(), bs.end(), b) == bs.end()) bs.push_back(b);
    }
    std::vector<int> pvs = {plan_rule};
    if (plans)
        for (int pv : {0, all_alt})
            if (std::find(pvs.begin(), pvs.end(), pv) == pvs.end()) pvs.push_back(pv);
    std::vector<TunedShape> cands;
    for (int pv : pvs)
        for (int64_t b : bs)
            for (int w : {4, 8, 16})
                for (int ch : {1, 2, 3})
                    if (ch == 1 || ((P + b - 1) / b >= 2 && (ch == 2 || fused_shape(f, w, pv))))
                        cands.push_back({(int)b, w, ch, pv});
    RunPlan tb; // trial buffers, released on return
    tb.G = pl.G;
    tb.nbins = pl.nbins;
    tb.n_tiles = pl.n_tiles;
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev()
        {
            if (e) (void)hipEventDestroy(e);
        }
    } e0, e1;
    HIPCHK(c, hipEventCreate(&e0.e));
    HIPCHK(c, hipEventCreate(&e1.e));
    const TrialFn trial = [&](const TunedShape& cand, double& cost) -> int {
        Shape ts;
        ts.B = cand[0];
        ts.col_tile = cand[1];
        ts.chains = cand[2];
        ts.plan_v = cand[3];
        RunFlags ft = f;
        ft.P = std::min<int64_t>(P, 4 * ts.B);
        ft.want_each = false;
        ts.n_batches = (ft.P + ts.B - 1) / ts.B;
        const std::vector<int2> sub(shifts.begin(), shifts.begin() + ft.P);
        RunViews tv;
        tv.twx = twx;
        tv.twy = twy;
        Mem heap;
        const RunOut none;
        if (int r = prepare_buffers(c, heap, tb, ft, ts, sub, none, tv)) return r;
        Prof off;
        if (int r = enqueue_job(c, pl, sp, ft, ts, tv, none, off)) return r;
        HIPCHK(c, hipEventRecord(e0.e, c->stream));
        if (int r = enqueue_job(c, pl, sp, ft, ts, tv, none, off)) return r;
        HIPCHK(c, hipEventRecord(e1.e, c->stream));
        HIPCHK(c, hipEventSynchronize(e1.e));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, e0.e, e1.e));
        cost = 1e-3 * ms / ((double)ft.P * f.S);
        return MSF_OK;
    };
    const int r = autotune_shape(c, key.str(), rule, cands, trial, measure, pick, state);
    (void)hipStreamSynchronize(c->stream);
    return r;
}

// Copies the results to the caller's buffers and finishes the FEM modes on
// the host.
int finish_outputs(msf_ctx* c, const RunPlan& pl, const msf_specimen* sp, const msf_optics* o,
                   const RunFlags& f, const RunViews& v, const RunOut& out)
{
    const int64_t P = f.P;
    const int nbins = pl.nbins;
    if (f.want_adf && out.adf)
        HIPCHK(c, hipMemcpy(out.adf, v.adf, (size_t)P * sizeof(double), hipMemcpyDeviceToHost));
    if (f.want_d1 && nbins > 0)
        HIPCHK(c, hipMemcpy(out.fem, v.femv, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
    if (f.want_d2 && nbins > 0) { // adfstem.cpp:2896-2915: the mean over positions
        std::vector<double> vs(nbins);
        HIPCHK(c, hipMemcpy(vs.data(), v.vsum, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < nbins; ++k) out.vbar_r[k] = vs[k] / (double)P;
    }
    if (f.want_pix && nbins > 0) {
        std::vector<double> s1((size_t)f.plane), s2((size_t)f.plane);
        HIPCHK(c, hipMemcpy(s1.data(), v.pm1, (size_t)f.plane * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(s2.data(), v.pm2, (size_t)f.plane * sizeof(double), hipMemcpyDeviceToHost));
        fem_pixel_modes(sp->dom, pl.hg, s1, s2, P, o->alpha_max, out);
        if (out.mean_abs || out.mean_sqr || out.var2d)
            fem_pixel_images(sp->dom, sp->keep, f.ncp, s1, s2, P, o->alpha_max, out);
    }
    if (f.want_corr && nbins * f.nphi > 0) { // adfstem.cpp:3620-3632: the mean over positions
        const size_t n = (size_t)nbins * f.nphi;
        std::vector<double> cs(n);
        HIPCHK(c, hipMemcpy(cs.data(), v.csum, n * sizeof(double), hipMemcpyDeviceToHost));
        if (out.corr_avg)
            for (size_t k = 0; k < n; ++k) out.corr_avg[k] = cs[k] / (double)P;
        if (out.corr_sqr)
            HIPCHK(c, hipMemcpy(out.corr_sqr, v.csqr, n * sizeof(double), hipMemcpyDeviceToHost));
        if (f.want_each)
            HIPCHK(c, hipMemcpy(out.corr_each, v.ceach, (size_t)P * n * sizeof(double),
                                hipMemcpyDeviceToHost));
    }
    return MSF_OK;
}

int run_internal(msf_ctx* c, Mem& m, const msf_specimen* sp, const msf_optics* o,
                 const msf_detector* det, const msf_fem* fem, const std::vector<int2>& shifts,
                 RunOut out)
{
    const double t_start = now_ms();
    RunFlags f;
    if (int r = check_run(c, sp, o, det, fem, shifts, out, f)) return r;
    RunPlan& pl = m.scratch() ? c->plan_na : c->plan;
    if (int r = prepare_groups(c, m, pl, sp, o, det, fem, f)) return r;
    f.want_each = f.want_corr && out.corr_each && pl.nbins * f.nphi > 0;
    if (int r = prepare_vacuum(c, m, pl, sp, o, f)) return r;
    pl.ncp = f.ncp;
    RunViews v;
    if (int r = twiddles(c, m, f.nx, v.twx)) return r;
    if (int r = twiddles(c, m, f.ny, v.twy)) return r;
    Shape sh;
    if (int r = launch_shape(c, m, pl, sp, shifts, out, f, v.twx, v.twy, sh)) return r;
    if (int r = prepare_buffers(c, m, pl, f, sh, shifts, out, v)) return r;

    Prof prof;
    prof.on = c->cfg.profile_kernels != 0 && !m.scratch();
    prof.s = c->stream;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double t_loop = now_ms();
    // Profiled, exit-wave (host copies per batch) and allocation-free runs,
    // and two automatic chains, launch directly; the rest replay a graph.
    int graph_state = 0;
    if (!prof.on && !out.exit && !c->cfg.no_graph && !sh.auto_eager && !m.scratch()) {
        const std::vector<uint64_t> sig = graph_signature(pl, sp, f, sh, v, out);
        graph_state = 2;
        bool captured = true;
        if (!pl.graph || pl.graph_sig != sig) {
            pl.drop_graph();
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const int r = enqueue_job(c, pl, sp, f, sh, v, out, prof);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (r != MSF_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                // A capture can be invalidated from outside, e.g. by another
                // thread's allocation while it runs; nothing captured has
                // executed, so one chain simply launches on the stream.
                if (sh.chains == 2) return r != MSF_OK ? r : fail(c, MSF_ERR_DEVICE, hipGetErrorString(e));
                (void)hipGetLastError();
                captured = false;
            } else {
                const hipError_t ei = hipGraphInstantiate(&pl.graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) pl.graph = nullptr;
                HIPCHK(c, ei);
                pl.graph_sig = sig;
                pl.graph_fused = prof.fused;
                graph_state = 1;
            }
        }
        if (captured) {
            HIPCHK(c, hipGraphLaunch(pl.graph, c->stream));
            prof.fused = pl.graph_fused;
        } else {
            graph_state = 0;
            if (int r = enqueue_job(c, pl, sp, f, sh, v, out, prof)) return r;
        }
    } else {
        if (int r = enqueue_job(c, pl, sp, f, sh, v, out, prof)) return r;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double t_end = now_ms();
    if (int r = finish_outputs(c, pl, sp, o, f, v, out)) return r;

    msf_stats st{};
    st.setup_ms = t_loop - t_start;
    st.slice_loop_ms = t_end - t_loop;
    st.total_ms = now_ms() - t_start;
    st.probe_slices = (uint64_t)f.P * (uint64_t)f.S;
    st.batch = (uint32_t)sh.B;
    st.n_batches = (uint32_t)sh.n_batches;
    // compulsory traffic of a full-batch pass: read + write of the compact wave
    st.row_pass_bytes = 2.0 * (double)sh.B * f.nx * (f.ka + f.kb) * sizeof(float2);
    st.col_pass_bytes = st.row_pass_bytes;
    st.graph = graph_state;
    st.chains = sh.chains;
    st.fused_launches = prof.fused;
    st.col_tile = sh.col_tile ? sh.col_tile
                              : msf::col_tile_auto(f.nx, sh.plan_v & 1, (int)sh.B, f.ncp);
    st.autotuned = sh.tuned_state;
    st.row_fft_plan = 1 + ((sh.plan_v >> 1) & 1);
    st.col_fft_plan = 1 + (sh.plan_v & 1);
    if (prof.on) {
        double sum[K_NUM] = {0}, cnt[K_NUM] = {0};
        for (size_t k = 0; k < prof.kind.size(); ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, prof.ev[2 * k], prof.ev[2 * k + 1]) == hipSuccess) {
                sum[prof.kind[k]] += ms;
                cnt[prof.kind[k]] += 1;
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
    }
    c->stats = st;
    return MSF_OK;
}

} // namespace msf_impl
