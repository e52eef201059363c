// msf_tune.cpp -- the launch-shape autotuner of the slice loop and its
// tuning file.  The reference plans its FFTs with FFTW_MEASURE
// (adfstem.cpp:1499-1553) and keeps the plans as wisdom in a file named per
// grid (tem.cpp:588-616, exported at :1749-1765); here the measured choice is
// the launch shape of the whole batch step (positions per batch, column-pass
// tile, batch chains and, on request, the radix plans), kept process-wide
// per device type and job shape, and saved to / loaded from a text file.
#include "msf_internal.hpp"

#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>

namespace msf_impl {

namespace {

std::mutex g_tune_mu;
std::map<std::string, TunedShape> g_tuned; // "<arch> <shape key>" -> choice

} // namespace

std::string device_arch(int device)
{
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::string a(p.gcnArchName);
    const size_t colon = a.find(':');
    if (colon != std::string::npos) a.resize(colon); // feature suffixes do not change the choice
    return a + "-" + std::to_string(p.multiProcessorCount) + "cu";
}

// With the candidates' trial costs (seconds per position-slice of a whole
// batch step: both slice passes of every slice, the reductions and the FEM
// statistics, with the candidate's chains) keep the fastest, unless it
// beats the rule's shape by less than 3 % (trial noise).
int autotune_shape(msf_ctx* c, const std::string& key, const TunedShape& rule,
                   const std::vector<TunedShape>& cands, const TrialFn& trial, bool measure,
                   TunedShape& pick, int& state)
{
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            pick = it->second;
            state = 2;
            return MSF_OK;
        }
    }
    if (!measure) {
        pick = rule;
        state = 0;
        return MSF_OK;
    }
    double warm = 0.0; // clocks and caches to steady state first
    if (int r = trial(rule, warm)) return r;
    const double inf = std::numeric_limits<double>::infinity();
    double best = inf, rule_cost = inf;
    pick = rule;
    for (const TunedShape& cand : cands) {
        double cost = inf;
        if (int r = trial(cand, cost)) return r;
        if (cand == rule) rule_cost = cost;
        if (cost < best) {
            best = cost;
            pick = cand;
        }
    }
    if (rule_cost < inf && best > rule_cost / 1.03) pick = rule;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = pick;
    }
    state = 1;
    (void)c;
    return MSF_OK;
}

} // namespace msf_impl

using namespace msf_impl;

extern "C" {

int msf_tuning_save(const char* path)
{
    try {
        if (!path) return MSF_ERR_INVALID;
        std::ofstream f(path);
        if (!f.is_open()) return MSF_ERR_INVALID;
        f << "# msf tuning v1: <device> <nx> <ny> <kept columns> <positions up to 256> <reductions>"
             " <plan rule> = <batch> <column tile> <chains> <radix plans>\n";
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto& e : g_tuned)
            f << e.first << " = " << e.second[0] << " " << e.second[1] << " " << e.second[2] << " "
              << e.second[3] << "\n";
        f.flush();
        return f.good() ? MSF_OK : MSF_ERR_INVALID;
    } catch (...) {
        return MSF_ERR_INVALID;
    }
}

int msf_tuning_load(const char* path, uint32_t* entries)
{
    try {
        if (!path) return MSF_ERR_INVALID;
        std::ifstream f(path);
        if (!f.is_open()) return MSF_ERR_INVALID; // as fftw_import_wisdom_from_filename: no file
        std::map<std::string, TunedShape> got;
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            const size_t eq = line.find(" = ");
            if (eq == std::string::npos) return MSF_ERR_INVALID;
            std::istringstream v(line.substr(eq + 3));
            TunedShape s{};
            if (!(v >> s[0] >> s[1] >> s[2] >> s[3])) return MSF_ERR_INVALID;
            if (s[0] < 1 || (s[1] != 1 && s[1] != 4 && s[1] != 8 && s[1] != 16) ||
                (s[2] < 1 || s[2] > 3) || s[3] < 0 || s[3] > 3)
                return MSF_ERR_INVALID;
            got[line.substr(0, eq)] = s;
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto& e : got) g_tuned[e.first] = e.second;
        if (entries) *entries = (uint32_t)got.size();
        return MSF_OK;
    } catch (...) {
        return MSF_ERR_INVALID;
    }
}

int msf_tuning_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
    return MSF_OK;
}

} // extern "C"
