// Host-only check of the ring-table build (msf_host.cpp build_groups /
// build_segments) at the 1024^2 FEM workload's settings: 32000-atom Si cell
// of 80 x 80 A, 300 kV, 20 mrad aperture, ADF 60-200 mrad, V_omega with
// dr = 0.1414213562, plus the |psi|^2 rings and the correlograph sectors (the
// largest tables: 1.18 M pixels sorted on up to 8 threads).  Built under
// AddressSanitizer + UBSan and under ThreadSanitizer by test_host_sanitizers.py;
// two builds also run concurrently, as two serving slots do.  Exits non-zero
// on any violated invariant.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../../ms-stem-fem_amd/csrc/msf_host.hpp"

static int check(const msf::Domain& d, int ncp, const msf::Groups& g, const msf::SegmentTiles& st)
{
    const int64_t plane = (int64_t)d.nx * ncp;
    const size_t G = g.kind.size();
    if (g.offsets.size() != G + 1) return 1;
    for (size_t k = 0; k < G; ++k)
        if (g.offsets[k] > g.offsets[k + 1]) return 2;
    if ((size_t)g.offsets[G] != g.pixels.size()) return 3;
    for (int32_t p : g.pixels)
        if (p < 0 || p >= plane) return 4;
    if (st.seg_of_group.size() != G + 1 || (int)st.tile_off.size() != st.n_tiles + 1) return 5;
    size_t listed = 0;
    for (int32_t p : st.pix) {
        if (p >= plane || p < -1) return 6;
        listed += p >= 0;
    }
    if (listed != g.pixels.size()) return 7;
    // the sums the reduction forms (segments in lane order, then a group's
    // segments in order) equal each group's direct sum of |psi|^2 or |psi|
    std::vector<float> re((size_t)plane), im((size_t)plane);
    uint64_t x = 88172645463325252ull;
    for (int64_t k = 0; k < plane; ++k) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        re[(size_t)k] = (float)((x >> 11) * (1.0 / 9007199254740992.0)) - 0.5f;
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        im[(size_t)k] = (float)((x >> 11) * (1.0 / 9007199254740992.0)) - 0.5f;
    }
    auto val = [&](int64_t k, bool amp) {
        const double a = re[(size_t)k], b = im[(size_t)k], i2 = a * a + b * b;
        return amp ? std::sqrt(i2) : i2;
    };
    // (segment w * 64 + l sums its lane's pixels in order; a group adds its
    // consecutive segments)
    std::vector<double> seg((size_t)st.n_tiles * 64, 0.0);
    for (int w = 0; w < st.n_tiles; ++w)
        for (int l = 0; l < 64; ++l) {
            double acc = 0.0;
            for (int k = st.tile_off[w]; k < st.tile_off[w + 1]; ++k) {
                const int32_t p = st.pix[(size_t)k * 64 + l];
                if (p >= 0) acc += val(p, st.seg_kind[(size_t)w * 64 + l] != 0);
            }
            seg[(size_t)w * 64 + l] = acc;
        }
    for (size_t k = 0; k < G; ++k) {
        double a = 0.0, b = 0.0;
        for (int32_t s = st.seg_of_group[k]; s < st.seg_of_group[k + 1]; ++s) a += seg[(size_t)s];
        for (int32_t q = g.offsets[k]; q < g.offsets[k + 1]; ++q) b += val(g.pixels[(size_t)q], g.kind[k] != 0);
        if (std::fabs(a - b) > 1e-9 * std::max(1.0, std::fabs(b))) return 15;
    }
    if ((int)g.bin_off.size() != g.nbins + 1) return 8;
    for (size_t k = 0; k < g.bin_full.size(); ++k)
        if (g.bin_full[k] < 0 || g.bin_full[k] >= d.nx * d.ny || g.bin_compact[k] >= plane) return 9;
    return 0;
}

static int run_once(int& pixels)
{
    const int n = 1024;
    const msf::Domain d = msf::make_domain(n, n, 80.0, 80.0, 0.0, 0.0, 300e3);
    const msf::RowRanges keep = msf::slice_keep_ranges(d);
    if (!keep.ok) return 20;
    const int W = 16;
    const int ncp = std::max(W, ((keep.amax + keep.bmax + W - 1) / W) * W);
    const double k1 = 0.06 / d.lambda, k2 = 0.2 / d.lambda;
    const msf::Groups g = msf::build_groups(d, keep, ncp, true, k1, k2, true, 0.1414213562, 0.02,
                                            true, true);
    const msf::SegmentTiles st = msf::build_segments(g, 32);
    pixels = (int)g.pixels.size();
    return check(d, ncp, g, st);
}

int main()
{
    int pa = 0, pb = 0, ra = 0, rb = 0;
    std::thread t1([&] { ra = run_once(pa); });
    std::thread t2([&] { rb = run_once(pb); });
    t1.join();
    t2.join();
    int pc = 0;
    const int rc = run_once(pc);
    std::printf("groups at 1024^2: %d pixels listed; checks %d %d %d\n", pc, ra, rb, rc);
    if (ra || rb || rc || pa != pc || pb != pc || pc < 1000000) return 1;
    return 0;
}
