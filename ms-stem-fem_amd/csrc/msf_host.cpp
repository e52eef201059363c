// msf_host.cpp -- host-side FP64 rules (see msf_host.hpp).
#include "msf_host.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <system_error>
#include <thread>

namespace msf {

double ref_pow2(double x)
{
    volatile double two = 2.0;
    return std::pow(x, (double)two);
}

// tem.cpp:424-438 (constants as written there)
void wavelength(double VV, double& lambda, double& gamma)
{
    const double m_e0 = 9.109e-31;
    const double cc = 299792458;
    const double ee = 1.60217662e-19;
    const double hh = 6.62607004e-34;
    volatile double ten = 10.0;
    const double e10 = std::pow((double)ten, (double)ten);
    lambda = hh * cc * e10 / std::sqrt(ee * VV * (2.0 * m_e0 * cc * cc + ee * VV));
    gamma = (hh * e10 / lambda) * (hh * e10 / lambda) * (m_e0 * cc * cc + ee * VV)
            / ((2 * m_e0 * cc * cc + ee * VV) * ee * VV * m_e0);
}

Domain make_domain(int nx, int ny, double xperiod, double yperiod, double xmin, double ymin,
                   double voltage)
{
    Domain d;
    d.nx = nx;
    d.ny = ny;
    wavelength(voltage, d.lambda, d.gamma);
    // tem.cpp:683-695
    if (nx / (2 * xperiod) < ny / (2 * yperiod))
        d.bw_pap = nx / (2 * xperiod);
    else
        d.bw_pap = ny / (2 * yperiod);
    d.bw_t = (2.0 / 3.0) * d.bw_pap;
    d.nxny = (double)((int64_t)nx * ny);
    d.sqrt_nxny = (nx == ny) ? (double)nx : std::sqrt(d.nxny);
    // tem.cpp:763-768 spacings; func.cpp:92-106 and :147-167 grids
    const double dx = xperiod / nx, dy = yperiod / ny;
    const double dkx = 1.0 / xperiod, dky = 1.0 / yperiod;
    const double kxp = nx / xperiod, kyp = ny / yperiod;
    d.kx.resize(nx);
    d.ky.resize(ny);
    d.xx.resize(nx);
    d.yy.resize(ny);
    for (int i = 0; i < nx / 2; i++) d.kx[i] = i * dkx;
    for (int i = nx / 2; i < nx; i++) d.kx[i] = i * dkx - kxp;
    for (int i = 0; i < ny / 2; i++) d.ky[i] = i * dky;
    for (int i = ny / 2; i < ny; i++) d.ky[i] = i * dky - kyp;
    for (int i = 0; i < nx; i++) d.xx[i] = xmin + (i * dx);
    for (int i = 0; i < ny; i++) d.yy[i] = ymin + (i * dy);
    return d;
}

// ref_pow2 of every kx and ky, once per index rather than once per pixel
// (the same values: the same function of the same arguments)
std::vector<double> pow2_all(const std::vector<double>& k)
{
    std::vector<double> out(k.size());
    for (size_t i = 0; i < k.size(); ++i) out[i] = ref_pow2(k[i]);
    return out;
}

RowRanges slice_keep_ranges(const Domain& d)
{
    const double c_bw = ref_pow2(d.bw_t);   // bw_limit.cpp:36
    const double c_pr = d.bw_t * d.bw_t;     // slice.cpp:1375
    const std::vector<double> kx2 = pow2_all(d.kx), ky2 = pow2_all(d.ky);
    return row_ranges(d.nx, d.ny, [&](int i, int j) {
        const double k2 = kx2[i] + ky2[j];
        return !(k2 >= c_bw) && (k2 < c_pr);
    });
}

RowRanges column_keep_ranges(const RowRanges& rows, int nx, int ny, int ncp)
{
    const int ka = rows.amax, kb = rows.bmax, nc = ka + kb;
    auto kept = [&](int i, int j) { return j < rows.ab[2 * i] || j >= ny - rows.ab[2 * i + 1]; };
    RowRanges r;
    r.ab.assign(2 * (size_t)ncp, 0);
    for (int jc = 0; jc < ncp && jc < nc; ++jc) {
        const int j = jc < ka ? jc : jc - nc + ny;
        int c = 0;
        while (c < nx && kept(c, j)) ++c;
        int d = 0;
        if (c < nx)
            while (d < nx - c && kept(nx - 1 - d, j)) ++d;
        for (int i = c; i < nx - d; ++i)
            if (kept(i, j)) r.ok = false;
        r.ab[2 * jc] = c;
        r.ab[2 * jc + 1] = d;
        r.amax = std::max(r.amax, c);
        r.bmax = std::max(r.bmax, d);
    }
    return r;
}

RowRanges bwlimit_keep_ranges(const Domain& d)
{
    const double c_bw = ref_pow2(d.bw_t);
    const std::vector<double> kx2 = pow2_all(d.kx), ky2 = pow2_all(d.ky);
    return row_ranges(d.nx, d.ny, [&](int i, int j) { return !(kx2[i] + ky2[j] >= c_bw); });
}

RowRanges template_keep_ranges(const Domain& d)
{
    const double c2 = d.bw_pap * d.bw_pap; // scatterer_pap.hpp:82
    const std::vector<double> kx2 = pow2_all(d.kx), ky2 = pow2_all(d.ky);
    return row_ranges(d.nx, d.ny, [&](int i, int j) { return kx2[i] + ky2[j] < c2; });
}

RowRanges aperture_keep_ranges(const Domain& d, double alpha_max)
{
    const double l2 = d.lambda * d.lambda;      // adfstem.cpp:95
    const double a2 = alpha_max * alpha_max;    // adfstem.cpp:96
    return row_ranges(d.nx, d.ny, [&](int i, int j) {
        const double ksqr = d.kx[i] * d.kx[i] + d.ky[j] * d.ky[j];
        return l2 * ksqr < a2;
    });
}

int64_t shift_index(const std::vector<double>& grid, double q, int64_t previous)
{
    // The reference scans for the first i with grid[i] - q >= half.  The grid
    // is non-decreasing and rounded subtraction of a fixed q is monotone, so
    // the test holds from that index on and a bisection finds the same i.
    const double half = 0.5 * (grid[1] - grid[0]);
    size_t lo = 0, hi = grid.size();
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (grid[mid] - q >= half) hi = mid;
        else lo = mid + 1;
    }
    return lo < grid.size() ? (int64_t)lo - 1 : previous;
}

Slices autoslice(const double* q, uint32_t n, double zmin, double zperiod, double min_thickness)
{
    Slices s;
    if (n == 0) { // slice.cpp:632-639
        s.lower.push_back(zmin);
        s.thick.push_back(zperiod);
        s.atoms.emplace_back();
        return s;
    }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return q[3 * a + 2] < q[3 * b + 2]; });
    // slice.cpp:657-682: an atom joins the current slice while
    // z - z_first <= min_thickness; lower_bound accumulates member z
    s.lower.push_back(0.0);
    s.atoms.emplace_back();
    double tmp_lower = q[3 * order[0] + 2];
    for (uint32_t k = 0; k < n; ++k) {
        const int32_t a = order[k];
        const double z = q[3 * a + 2];
        if (z - tmp_lower <= min_thickness) {
            s.atoms.back().push_back(a);
            s.lower.back() += z;
        } else {
            s.atoms.emplace_back(1, a);
            s.lower.push_back(z);
            tmp_lower = z;
        }
    }
    const size_t ns = s.lower.size();
    for (size_t i = 0; i < ns; ++i) s.lower[i] = s.lower[i] / s.atoms[i].size(); // :690-699
    s.thick.resize(ns);
    for (size_t i = 0; i < ns; ++i) { // :709-730
        if (i + 1 == ns)
            s.thick[i] = q[3 * s.atoms[i].back() + 2] + min_thickness - s.lower[i];
        else
            s.thick[i] = s.lower[i + 1] - s.lower[i];
    }
    return s;
}

void propagator(double lambda, double thickness, const std::vector<double>& k, double bw_t,
                std::vector<double>& p)
{
    p.assign(2 * k.size(), 0.0);
    for (size_t i = 0; i < k.size(); ++i) {
        if (k[i] < bw_t && k[i] > -1.0 * bw_t) {
            const double trig = kRefPi * lambda * thickness * ref_pow2(k[i]);
            p[2 * i] = std::cos(trig);
            p[2 * i + 1] = -std::sin(trig);
        }
    }
}

int fem_bin_count(const Domain& d, double dr, double& delta_k)
{
    delta_k = dr * std::sqrt((d.kx[1] - d.kx[0]) * (d.kx[1] - d.kx[0])
                             + (d.ky[1] - d.ky[0]) * (d.ky[1] - d.ky[0]));
    const unsigned int nb = (unsigned int)(d.bw_t / delta_k);
    return (int)nb;
}

int correlograph_nphi(double delta_k, int nbins)
{
    unsigned int count = 0;
    for (unsigned int ii = 0; ii < (unsigned int)nbins + 1; ++ii) {
        const double k_bdy = ii * delta_k;
        if ((k_bdy >= 0.2) && (k_bdy <= 1.0)) ++count;
    }
    return count > 0 ? (int)count - 1 : 0;
}

std::vector<double> mtf_table_2d(const Domain& d, double xperiod, double yperiod,
                                 const double* mtf, const double* domain, int n, double rate)
{
    std::vector<double> dsq((size_t)n);
    for (int k = 0; k < n; ++k) dsq[k] = ref_pow2((0.5 * rate * domain[k])); // tem.cpp:283-284
    std::vector<double> m((size_t)d.nx * d.ny, 0.0);
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j) {
            const double xxx = ref_pow2(xperiod) * d.kx[i] / d.nx;
            const double yyy = ref_pow2(yperiod) * d.ky[j] / d.ny;
            const double rsqr = ref_pow2(xxx) + ref_pow2(yyy);
            double& v = m[(size_t)i * d.ny + j];
            // every segment holding rsqr overwrites (no early exit, as written)
            for (size_t idx = 0; idx + 1 < (size_t)n; ++idx) {
                if ((rsqr >= dsq[idx]) && (rsqr < dsq[idx + 1])) {
                    const double d1 = std::sqrt(dsq[idx]), d2 = std::sqrt(dsq[idx + 1]);
                    v = mtf[idx] + (mtf[idx + 1] - mtf[idx]) * (std::sqrt(rsqr) - d1) / (d2 - d1);
                    double arg = xxx / rate;
                    const double sx = (arg != 0) ? std::sin(arg) / arg : 1;
                    arg = yyy / rate;
                    const double sy = (arg != 0) ? std::sin(arg) / arg : 1;
                    v = v * sx * sy;
                }
            }
        }
    return m;
}

std::vector<uint8_t> outside_aperture(const Domain& d, double alpha_max)
{
    const double l2 = d.lambda * d.lambda, a2 = alpha_max * alpha_max; // adfstem.cpp:95-96
    std::vector<uint8_t> o((size_t)d.nx * d.ny);
    const std::vector<double> kx2 = pow2_all(d.kx), ky2 = pow2_all(d.ky);
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j) {
            const double ksqr = kx2[i] + ky2[j];
            o[(size_t)i * d.ny + j] = (l2 * ksqr > a2) ? 1 : 0;
        }
    return o;
}

// std::stable_sort on up to 8 threads: stable-sorted chunks, then pairwise
// std::inplace_merge, which keeps the left chunk's elements first among
// equals, so the order is exactly std::stable_sort's.
template <class It, class Cmp> static void stable_sort_threads(It first, It last, Cmp cmp)
{
    const size_t n = (size_t)(last - first);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    unsigned parts = 1;
    while (parts * 2 <= std::min(8u, hw)) parts *= 2;
    if (n < (1u << 16) || parts < 2) {
        std::stable_sort(first, last, cmp);
        return;
    }
    std::vector<size_t> b(parts + 1);
    for (unsigned k = 0; k <= parts; ++k) b[k] = n * k / parts;
    std::vector<std::thread> th;
    th.reserve(parts);
    try {
        for (unsigned k = 0; k < parts; ++k)
            th.emplace_back([&, k] { std::stable_sort(first + b[k], first + b[k + 1], cmp); });
        for (auto& t : th) t.join();
        for (unsigned w = 1; w < parts; w *= 2) {
            th.clear();
            for (unsigned k = 0; k + w < parts; k += 2 * w)
                th.emplace_back([&, k, w] {
                    std::inplace_merge(first + b[k], first + b[k + w],
                                       first + b[std::min(k + 2 * w, parts)], cmp);
                });
            for (auto& t : th) t.join();
        }
    } catch (const std::system_error&) {
        // no thread could be started: finish what started, then sort here
        for (auto& t : th)
            if (t.joinable()) t.join();
        std::stable_sort(first, last, cmp);
    }
}

Groups build_groups(const Domain& d, const RowRanges& keep, int ncp, bool adf, double k1,
                    double k2, bool fem, double dr, double alpha_max, bool fem_sqr,
                    bool correlograph)
{
    Groups g;
    const int nx = d.nx, ny = d.ny;
    const int A = keep.amax, B = keep.bmax;
    auto kept = [&](int i, int j) {
        return j < keep.ab[2 * i] || j >= ny - keep.ab[2 * i + 1];
    };
    auto compact = [&](int i, int j) -> int32_t {
        const int jc = (j < A) ? j : j - ny + A + B;
        return (int32_t)((int64_t)i * ncp + jc);
    };
    g.offsets.push_back(0);
    // group 0: ADF ring, k_sqr = kx*kx + ky*ky in (k1, k2] (func.cpp:2630-2637)
    if (adf) {
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                const double k_sqr = d.kx[i] * d.kx[i] + d.ky[j] * d.ky[j];
                if (k_sqr > k1 && k_sqr <= k2 && kept(i, j)) g.pixels.push_back(compact(i, j));
            }
    }
    g.offsets.push_back((int32_t)g.pixels.size());
    g.kind.push_back(0);
    if (fem) {
        double dk = 0.0;
        g.nbins = fem_bin_count(d, dr, dk);
        // sort key of adfstem.cpp:1043-1062; ties keep row-major order
        std::vector<std::pair<double, int64_t>> m;
        m.reserve((size_t)nx * ny);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                m.emplace_back(d.kx[i] * d.kx[i] + d.ky[j] * d.ky[j], (int64_t)i * ny + j);
        stable_sort_threads(m.begin(), m.end(),
                            [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
                                return a.first < b.first;
                            });
        size_t it = 0;
        g.fem_abs = (int)g.kind.size();
        g.bin_off.push_back(0);
        for (int ii = 0; ii < g.nbins; ++ii) {
            const double lower = (unsigned int)ii * dk;
            const double upper = (unsigned int)(ii + 1) * dk;
            g.edges.push_back(lower);
            int32_t count = 0;
            for (; it < m.size() && m[it].first <= (upper * upper); ++it) {
                const int i = (int)(m[it].second / ny), j = (int)(m[it].second % ny);
                ++count;
                const bool k = kept(i, j);
                if (k) g.pixels.push_back(compact(i, j));
                g.bin_full.push_back((int32_t)m[it].second);
                g.bin_compact.push_back(k ? compact(i, j) : -1);
            }
            g.bin_off.push_back((int32_t)g.bin_full.size());
            g.counts.push_back(count);
            g.zeroed.push_back(d.lambda * upper <= alpha_max ? 1 : 0);
            g.offsets.push_back((int32_t)g.pixels.size());
            g.kind.push_back(1);
        }
        if (correlograph) {
            // ring ii, sector jj: phi = PI + atan2(ky, kx) in (jj dphi, (jj+1) dphi]
            // (adfstem.cpp:1013-1027, variance.cpp:226-250)
            g.nphi = correlograph_nphi(dk, g.nbins);
            const double dphi = g.nphi > 0 ? 2 * kRefPi / g.nphi : 0.0;
            std::vector<std::vector<int32_t>> sec((size_t)g.nbins * std::max(g.nphi, 0));
            for (int ii = 0; ii < g.nbins && g.nphi > 0; ++ii)
                for (int32_t k = g.bin_off[ii]; k < g.bin_off[ii + 1]; ++k) {
                    if (g.bin_compact[k] < 0) continue;
                    const int i = g.bin_full[k] / ny, j = g.bin_full[k] % ny;
                    const double phi = kRefPi + std::atan2(d.ky[j], d.kx[i]);
                    double lower = 0 * dphi;
                    for (int jj = 0; jj < g.nphi; ++jj) {
                        if ((phi > lower) && (phi <= (unsigned int)(jj + 1) * dphi)) {
                            sec[(size_t)ii * g.nphi + jj].push_back(g.bin_compact[k]);
                            break;
                        }
                        lower = (unsigned int)jj * dphi;
                    }
                }
            g.corr = (int)g.kind.size();
            for (const std::vector<int32_t>& v : sec) {
                g.pixels.insert(g.pixels.end(), v.begin(), v.end());
                g.offsets.push_back((int32_t)g.pixels.size());
                g.kind.push_back(1);
            }
        }
        if (fem_sqr) { // the same rings summing |psi|^2 (integrate_out_phi_double, d2)
            g.fem_sqr = (int)g.kind.size();
            for (int ii = 0; ii < g.nbins; ++ii) {
                for (int32_t k = g.offsets[g.fem_abs + ii]; k < g.offsets[g.fem_abs + ii + 1]; ++k) {
                    const int32_t px = g.pixels[k];
                    g.pixels.push_back(px);
                }
                g.offsets.push_back((int32_t)g.pixels.size());
                g.kind.push_back(0);
            }
        }
    }
    return g;
}

SegmentTiles build_segments(const Groups& g, int seg)
{
    // A group of n pixels gets ns = ceil(n / seg) consecutive segments.  Its
    // pixels, in memory order, are dealt round-robin: pixel i goes to segment
    // i % ns at position i / ns.  Lanes of one tile handle consecutive
    // segments, so at every step they read neighbouring pixels of the group
    // (coalesced).  Each segment is summed in a fixed order, and the segments
    // of a group are added in segment order, independent of the batch.
    SegmentTiles t;
    std::vector<std::vector<int32_t>> segs;
    const int G = (int)g.kind.size();
    std::vector<int32_t> sorted;
    for (int k = 0; k < G; ++k) {
        t.seg_of_group.push_back((int32_t)segs.size());
        const int32_t n = g.offsets[k + 1] - g.offsets[k];
        if (n == 0) continue;
        const int32_t ns = (n + seg - 1) / seg;
        sorted.assign(g.pixels.begin() + g.offsets[k], g.pixels.begin() + g.offsets[k + 1]);
        std::sort(sorted.begin(), sorted.end());
        const size_t s0 = segs.size();
        segs.resize(s0 + (size_t)ns);
        for (int32_t i = 0; i < n; ++i) segs[s0 + (size_t)(i % ns)].push_back(sorted[(size_t)i]);
        for (int32_t j = 0; j < ns; ++j) t.seg_kind.push_back(g.kind[k]);
    }
    t.seg_of_group.push_back((int32_t)segs.size());
    t.n_segments = (int)segs.size();
    t.n_tiles = (t.n_segments + 63) / 64;
    t.tile_off.push_back(0);
    for (int w = 0; w < t.n_tiles; ++w) {
        size_t len = 0;
        for (int l = 0; l < 64; ++l) {
            const int s = w * 64 + l;
            if (s < t.n_segments) len = std::max(len, segs[(size_t)s].size());
        }
        const size_t base = t.pix.size();
        t.pix.resize(base + len * 64, -1);
        for (int l = 0; l < 64; ++l) {
            const int s = w * 64 + l;
            if (s >= t.n_segments) continue;
            for (size_t k = 0; k < segs[(size_t)s].size(); ++k) t.pix[base + k * 64 + l] = segs[(size_t)s][k];
        }
        t.tile_off.push_back(t.tile_off.back() + (int32_t)len);
    }
    t.seg_kind.resize((size_t)t.n_tiles * 64, 0);
    return t;
}

} // namespace msf
