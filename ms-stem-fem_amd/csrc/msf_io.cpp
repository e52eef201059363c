// msf_io.cpp -- model readers and result writers of the command line
// (include/msf_io.h).  Host code only; parsing goes through iostreams so
// that number and token extraction behaves as in the reference readers.
#include "../../include/msf_io.h"
#include "../../include/msf.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int io_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// Every entry point turns a C++ exception into an error code: none may cross
// the C ABI.
#define MSF_IO_CATCH_ALL                                                                      \
    catch (const std::bad_alloc&)                                                             \
    {                                                                                         \
        return io_fail(MSF_ERR_MEMORY, "out of host memory");                                \
    }                                                                                         \
    catch (const std::exception& e_)                                                          \
    {                                                                                         \
        return io_fail(MSF_ERR_INVALID, std::string("unexpected error: ") + e_.what());      \
    }                                                                                         \
    catch (...)                                                                               \
    {                                                                                         \
        return io_fail(MSF_ERR_INVALID, "unexpected error");                                 \
    }

// Room reserved up front for a declared atom count: never more than this, so
// a malformed count (a negative number read as size_t wraps to 2^64 - 1)
// fails by running out of atom lines, as the reference's reader does, instead
// of in the allocator.
constexpr size_t kReserveAtoms = size_t(1) << 20;

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    return s;
}

// Element symbols by atomic number (standard periodic table).
const char* const kSymbols[103] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",
    "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
    "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh",
    "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd",
    "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re",
    "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
    "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr"};

// Mass windows of guessZ (tem_io.cpp:1030-1257), tested in this order: a
// mass matches an entry if it equals `exact` (when set) or lies in
// [lo, hi].  Values are the reference's; its one strict bound (Ar, mass <
// 40.0) is expressed as hi = the largest double below 40.0.  Z = 97 and
// 118 have no window there.
struct MassWindow {
    double exact, lo, hi;
    uint32_t Z;
};
constexpr double kNone = -1.0;
const double kBelow40 = std::nextafter(40.0, 0.0);
const MassWindow kMasses[] = {
    {kNone, 0.9, 1.1, 1},         {kNone, 3.9, 4.1, 2},         {kNone, 6.9, 7.0, 3},
    {kNone, 9.0, 9.1, 4},         {kNone, 10.8, 11.0, 5},       {kNone, 12.0, 12.1, 6},
    {kNone, 14.0, 14.1, 7},       {kNone, 15.9, 16.0, 8},       {kNone, 18.9, 19.0, 9},
    {kNone, 20.0, 20.2, 10},      {kNone, 22.9, 23.0, 11},      {kNone, 24.0, 24.4, 12},
    {kNone, 26.9, 27.0, 13},      {kNone, 28.0, 28.1, 14},      {kNone, 30.9, 31.0, 15},
    {kNone, 32.0, 32.1, 16},      {kNone, 35.4, 35.5, 17},      {36.0, 39.9, kBelow40, 18},
    {38.0, 39.0, 39.1, 19},       {kNone, 40.0, 40.1, 20},      {kNone, 44.9, 45.0, 21},
    {44.0, 47.8, 48.0, 22},       {46.0, 50.0, 51.0, 23},       {kNone, 51.8, 51.99999, 24},
    {kNone, 54.0, 55.0, 25},      {52.0, 55.0, 55.9, 26},       {kNone, 58.0, 59.0, 27},
    {56.0, 58.0, 58.7, 28},       {kNone, 63.0, 63.6, 29},      {60.0, 65.0, 65.4, 30},
    {62.0, 69.0, 69.8, 31},       {64.0, 72.0, 72.7, 32},       {66.0, 75.8, 75.9999, 33},
    {68.0, 78.9, 78.999, 34},     {70.0, 79.0, 80.0, 35},       {kNone, 83.0, 84.0, 36},
    {74.0, 85.0, 85.5, 37},       {76.0, 87.6, 88.0, 38},       {78.0, 88.9, 89.0, 39},
    {kNone, 91.0, 91.3, 40},      {82.0, 92.9, 93.0, 41},       {kNone, 95.9, 96.0, 42},
    {86.0, 96.9, 97.1, 43},       {kNone, 101.0, 101.1, 44},    {90.0, 102.9, 103.0, 45},
    {92.0, 106.0, 106.5, 46},     {94.0, 107.8, 108.0, 47},     {kNone, 112.4, 112.5, 48},
    {98.0, 114.8, 114.9, 49},     {100.0, 118.7, 118.8, 50},    {102.0, 121.7, 121.8, 51},
    {104.0, 127.5, 127.7, 52},    {kNone, 126.8, 127.0, 53},    {kNone, 131.2, 131.3, 54},
    {110.0, 132.9, 133.0, 55},    {112.0, 137.0, 137.4, 56},    {114.0, 138.9, 139.0, 57},
    {116.0, 140.0, 140.2, 58},    {118.0, 140.9, 141.0, 59},    {120.0, 144.2, 144.3, 60},
    {122.0, 144.9, 145.1, 61},    {124.0, 150.3, 150.4, 62},    {126.0, 151.9, 152.0, 63},
    {128.0, 157.0, 157.3, 64},    {130.0, 158.9, 159.0, 65},    {132.0, 162.0, 163.0, 66},
    {134.0, 164.9, 165.0, 67},    {136.0, 167.0, 167.3, 68},    {138.0, 168.9, 168.0, 69},
    {kNone, 173.0, 173.1, 70},    {142.0, 174.9, 175.0, 71},    {kNone, 178.0, 178.5, 72},
    {146.0, 180.9, 181.0, 73},    {148.0, 183.8, 184.0, 74},    {150.0, 186.0, 186.3, 75},
    {kNone, 190.0, 190.3, 76},    {154.0, 192.0, 192.3, 77},    {156.0, 195.0, 195.1, 78},
    {kNone, 196.9, 197.0, 79},    {160.0, 200.5, 201.0, 80},    {kNone, 204.0, 204.4, 81},
    {164.0, 207.0, 207.3, 82},    {166.0, 208.9, 208.9999, 83}, {168.0, 208.99991, 209.1, 84},
    {170.0, 209.9, 210.1, 85},    {172.0, 221.9, 222.1, 86},    {174.0, 222.9, 223.1, 87},
    {176.0, 225.9, 226.1, 88},    {kNone, 226.9, 227.1, 89},    {180.0, 232.0, 232.1, 90},
    {182.0, 231.0, 231.1, 91},    {kNone, 238.0, 238.1, 92},    {186.0, 236.9, 237.1, 93},
    {188.0, 243.9, 244.1, 94},    {kNone, 242.9, 243.1, 95},    {192.0, 246.9, 247.1, 96},
    {kNone, 250.9, 251.1, 98},    {198.0, 251.9, 252.1, 99},    {kNone, 256.9, 257.1, 100},
    {202.0, 257.9, 258.1, 101},   {kNone, 258.9, 259.1, 102},   {206.0, 265.9, 266.1, 103},
    {208.0, 266.9, 267.1, 104},   {kNone, 267.9, 268.1, 105},   {212.0, 268.9, 269.1, 106},
    {214.0, 269.9, 270.1, 107},   {216.0, 268.9, 269.1, 108},   {218.0, 277.9, 278.1, 109},
    {220.0, 280.9, 281.1, 110},   {kNone, 281.9, 282.1, 111},   {224.0, 284.9, 285.1, 112},
    {kNone, 285.9, 286.1, 113},   {228.0, 288.9, 289.1, 114},   {230.0, 289.9, 290.1, 115},
    {232.0, 292.9, 293.1, 116},   {234.0, 293.9, 294.1, 117},
};

int fill_model(msf_model* out, const std::vector<double>& q, const std::vector<uint32_t>& Z,
               double xp, double yp, double zp)
{
    out->n_atoms = (uint32_t)Z.size();
    out->q = (double*)std::malloc(std::max<size_t>(1, q.size()) * sizeof(double));
    out->Z = (uint32_t*)std::malloc(std::max<size_t>(1, Z.size()) * sizeof(uint32_t));
    if (!out->q || !out->Z) {
        std::free(out->q);
        std::free(out->Z);
        out->q = nullptr;
        out->Z = nullptr;
        return io_fail(MSF_ERR_MEMORY, "out of host memory");
    }
    if (!q.empty()) std::memcpy(out->q, q.data(), q.size() * sizeof(double));
    if (!Z.empty()) std::memcpy(out->Z, Z.data(), Z.size() * sizeof(uint32_t));
    out->xperiod = xp;
    out->yperiod = yp;
    out->zperiod = zp;
    out->xmin = out->ymin = out->zmin = 0.0;
    return MSF_OK;
}

// getline followed by the stream's good() test, as every reader loop does
bool next_line(std::ifstream& f, std::string& line) { return (bool)std::getline(f, line) && f.good(); }

// the LAMMPS reader's skip of whitespace-only lines (tem_io.cpp:87-92 and
// its repeats); the reference loops forever at end of file, here it fails
bool skip_blank(std::ifstream& f, std::string& line)
{
    while (line.find_first_not_of(" \t") == std::string::npos) {
        if (!std::getline(f, line)) return false;
    }
    return true;
}

// double -> uint16_t as x86-64 code converts it (truncation through a
// 32-bit integer; NaN and out-of-range give 0)
uint16_t to_u16(double d)
{
    if (!(d > -2147483649.0 && d < 2147483648.0)) return 0;
    return (uint16_t)(uint32_t)(int32_t)d;
}

void put16(std::vector<uint8_t>& b, size_t at, uint16_t v)
{
    b[at] = (uint8_t)(v & 0xff);
    b[at + 1] = (uint8_t)(v >> 8);
}
void put32(std::vector<uint8_t>& b, size_t at, uint32_t v)
{
    for (int k = 0; k < 4; ++k) b[at + k] = (uint8_t)(v >> (8 * k));
}

// TIFF RATIONAL from a non-negative double: exact for integers, else the
// largest power-of-ten denominator that keeps the numerator in 32 bits
void rational(double v, uint32_t& num, uint32_t& den)
{
    if (!(v >= 0.0) || v > 4294967295.0) {
        num = 0;
        den = 1;
        return;
    }
    den = 1;
    while (den < 1000000000u && std::floor(v * den) != v * den && v * den * 10.0 < 4294967295.0)
        den *= 10;
    num = (uint32_t)std::llround(v * den);
}

} // namespace

extern "C" {

const char* msf_io_last_error(void) { return g_err.c_str(); }

uint32_t msf_element_z(const char* name)
{
    if (!name) return 0;
    const std::string s = lower(name);
    for (uint32_t z = 1; z <= 103; ++z) {
        if (s == lower(kSymbols[z - 1])) return z;
        if (s == std::to_string(z)) return z;
    }
    return 0;
}

int msf_guess_z(double mass, uint32_t* Z)
{
    try {
        if (!Z) return io_fail(MSF_ERR_INVALID, "Z is NULL");
        for (const MassWindow& w : kMasses) {
            if ((w.exact != kNone && mass == w.exact) || (mass >= w.lo && mass <= w.hi)) {
                *Z = w.Z;
                return MSF_OK;
            }
        }
        *Z = 0;
        std::ostringstream m;
        m << "could not guess element having mass " << mass;
        return io_fail(MSF_ERR_INVALID, m.str());
    }
    MSF_IO_CATCH_ALL
}

void msf_model_free(msf_model* m)
{
    if (!m) return;
    std::free(m->q);
    std::free(m->Z);
    m->q = nullptr;
    m->Z = nullptr;
    m->n_atoms = 0;
}

void msf_mtf_table_free(msf_mtf_table* t)
{
    if (!t) return;
    std::free(t->domain);
    std::free(t->mtf);
    t->domain = nullptr;
    t->mtf = nullptr;
    t->n = 0;
}

int msf_read_mtf(const char* path, msf_mtf_table* out)
{
    try {
        if (!path || !out) return io_fail(MSF_ERR_INVALID, "NULL argument");
        *out = msf_mtf_table{};
        std::ifstream in(path);
        if (!in.is_open()) return io_fail(MSF_ERR_INVALID, std::string("Error opening MTF file: ") + path);
        std::vector<double> rows[2]; // domain, mtf
        for (std::vector<double>& v : rows) {
            std::string line;
            if (!(std::getline(in, line) && in.good())) break;
            std::replace(line.begin(), line.end(), ',', ' ');
            std::istringstream ls(line);
            // one value per line, as declared by the reference before its loop: an
            // extraction that finds only trailing blanks leaves it unchanged, so the
            // last value repeats (for a blank line the reference reads an
            // uninitialised double; 0 here)
            double value = 0.0;
            while (!ls.eof()) {
                ls >> value;
                if (ls.fail() && !ls.eof()) // the reference's loop never ends here
                    return io_fail(MSF_ERR_INVALID, std::string("non-numeric entry in MTF file: ") +
                                                        path);
                v.push_back(value);
            }
        }
        if (rows[0].size() != rows[1].size())
            return io_fail(MSF_ERR_INVALID, std::string("Error reading MTF file: ") + path +
                                                "\n MTF (1st line) and MTF domain (2nd line) do not "
                                                "have the same number of elements.");
        if (rows[0].empty()) return io_fail(MSF_ERR_INVALID, std::string("empty MTF table: ") + path);
        const size_t n = rows[0].size();
        out->domain = (double*)std::malloc(n * sizeof(double));
        out->mtf = (double*)std::malloc(n * sizeof(double));
        if (!out->domain || !out->mtf) {
            msf_mtf_table_free(out);
            return io_fail(MSF_ERR_INVALID, "out of memory");
        }
        std::copy(rows[0].begin(), rows[0].end(), out->domain);
        std::copy(rows[1].begin(), rows[1].end(), out->mtf);
        out->n = (uint32_t)n;
        return MSF_OK;
    }
    MSF_IO_CATCH_ALL
}

int msf_read_xyz(const char* path, msf_model* out)
{
    try {
        if (!path || !out) return io_fail(MSF_ERR_INVALID, "NULL argument");
        *out = msf_model{};
        std::ifstream f(path);
        if (!f.is_open()) return io_fail(MSF_ERR_INVALID, std::string("Error opening position file: ") + path);
        std::string line;
        // line 1: declared population (tem_io.cpp:4659-4688)
        size_t declared = 0;
        if (!next_line(f, line)) return io_fail(MSF_ERR_INVALID, "Error reading position data file; first line");
        {
            std::istringstream s(line);
            s >> declared;
            if (!declared)
                return io_fail(MSF_ERR_INVALID, "Error reading position data file; declared_population : 0");
        }
        // line 2: periods, all > 0 (tem_io.cpp:4690-4731)
        double xp = 0, yp = 0, zp = 0;
        if (!next_line(f, line)) return io_fail(MSF_ERR_INVALID, "Error reading position data file; second line");
        {
            std::istringstream s(line);
            s >> xp >> yp >> zp;
            if (!(xp > 0.0 && yp > 0.0 && zp > 0.0))
                return io_fail(MSF_ERR_INVALID,
                               "Error reading xyz position data file; second line should begin with "
                               "<xperiod> <yperiod> <zperiod>");
        }
        // atoms (tem_io.cpp:4747-4815); extrema start at 0
        std::vector<double> q;
        std::vector<uint32_t> Z;
        q.reserve(3 * std::min(declared, kReserveAtoms));
        Z.reserve(std::min(declared, kReserveAtoms));
        double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
        while (Z.size() < declared) {
            if (!next_line(f, line)) {
                std::ostringstream m;
                m << "Error reading position data file; atoms read : " << Z.size()
                  << ", declared population : " << declared;
                return io_fail(MSF_ERR_INVALID, m.str());
            }
            std::istringstream s(line);
            std::string element;
            s >> element;
            const uint32_t z = msf_element_z(lower(element).c_str());
            if (z < 1 || z > 103) {
                std::ostringstream m;
                m << "Error reading position data file; atom " << Z.size() << ", element " << element
                  << " invalid";
                return io_fail(MSF_ERR_INVALID, m.str());
            }
            double r[3] = {0, 0, 0};
            s >> r[0] >> r[1] >> r[2];
            for (int k = 0; k < 3; ++k) {
                if (r[k] < mn[k]) mn[k] = r[k];
                if (r[k] > mx[k]) mx[k] = r[k];
                q.push_back(r[k]);
            }
            Z.push_back(z);
        }
        if (mx[0] - mn[0] > xp || mx[1] - mn[1] > yp || mx[2] - mn[2] > zp)
            return io_fail(MSF_ERR_INVALID, "Error reading position data file; positions exceed period");
        // shift by the (non-positive) minima (tem_io.cpp:4843-4858)
        for (size_t a = 0; a < Z.size(); ++a)
            for (int k = 0; k < 3; ++k) q[3 * a + k] = q[3 * a + k] - mn[k];
        return fill_model(out, q, Z, xp, yp, zp);
    }
    MSF_IO_CATCH_ALL
}

int msf_read_lammps_data(const char* path, msf_model* out)
{
    try {
        if (!path || !out) return io_fail(MSF_ERR_INVALID, "NULL argument");
        *out = msf_model{};
        std::ifstream f(path);
        if (!f.is_open()) return io_fail(MSF_ERR_INVALID, std::string("Error opening position file: ") + path);
        const std::string where = "Error reading position data file; ";
        std::string line;
        std::getline(f, line); // title line (tem_io.cpp:62)
        // "<N> atoms" (tem_io.cpp:66-118)
        size_t declared = 0;
        if (!next_line(f, line) || !skip_blank(f, line)) return io_fail(MSF_ERR_INVALID, where + "atoms line");
        {
            std::istringstream s(line);
            std::string d;
            s >> declared >> d;
            if (!declared || lower(d) != "atoms")
                return io_fail(MSF_ERR_INVALID, where + "line should be : <number of atoms> atoms");
        }
        // "<n> atom types" (tem_io.cpp:120-176)
        size_t n_species = 0;
        if (!next_line(f, line) || !skip_blank(f, line)) return io_fail(MSF_ERR_INVALID, where + "atom types line");
        {
            std::istringstream s(line);
            std::string d1, d2;
            s >> n_species >> d1 >> d2;
            if (!n_species || lower(d1) != "atom" || lower(d2) != "types")
                return io_fail(MSF_ERR_INVALID, where + "line should be : <number> atom types");
        }
        // box bounds (tem_io.cpp:178-364)
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        const char* names[3][2] = {{"xlo", "xhi"}, {"ylo", "yhi"}, {"zlo", "zhi"}};
        for (int k = 0; k < 3; ++k) {
            if (!next_line(f, line) || !skip_blank(f, line))
                return io_fail(MSF_ERR_INVALID, where + "box bounds line");
            std::istringstream s(line);
            std::string d1, d2;
            s >> lo[k] >> hi[k] >> d1 >> d2;
            if (lower(d1) != names[k][0] || lower(d2) != names[k][1])
                return io_fail(MSF_ERR_INVALID, where + "line should be : <lo> <hi> " + names[k][0] + " " +
                                                    names[k][1]);
        }
        for (int k = 0; k < 3; ++k)
            if (hi[k] <= lo[k]) return io_fail(MSF_ERR_INVALID, where + "upper bound <= lower bound");
        // "Masses" section and one "<id> <mass>" line per species (tem_io.cpp:398-690)
        if (!next_line(f, line) || !skip_blank(f, line)) return io_fail(MSF_ERR_INVALID, where + "Masses");
        {
            std::istringstream s(line);
            std::string sec;
            s >> sec;
            if (lower(sec) != "masses") return io_fail(MSF_ERR_INVALID, where + "expected section Masses");
        }
        // grown line by line: a malformed type count fails when the lines run out
        std::vector<size_t> species_id;
        std::vector<double> mass;
        size_t n_read = 0;
        while (f.good() && n_read < n_species) {
            species_id.push_back(0);
            mass.push_back(0.0);
            if (next_line(f, line)) {
                if (!skip_blank(f, line)) break;
                std::istringstream s(line);
                s >> species_id[n_read] >> mass[n_read];
            }
            ++n_read;
        }
        if (n_read != n_species) return io_fail(MSF_ERR_INVALID, where + "fewer masses than atom types");
        std::vector<uint32_t> species_z(n_species, 0);
        for (size_t k = 0; k < n_species; ++k)
            if (int r = msf_guess_z(mass[k], &species_z[k])) return r;
        // "Atoms" section (tem_io.cpp:700-760)
        if (!next_line(f, line) || !skip_blank(f, line)) return io_fail(MSF_ERR_INVALID, where + "Atoms");
        {
            std::istringstream s(line);
            std::string sec;
            s >> sec;
            if (sec.empty() || lower(sec) != "atoms")
                return io_fail(MSF_ERR_INVALID, where + "expected section Atoms");
        }
        // atom lines "<number> <type> <x> <y> <z>" (tem_io.cpp:762-900)
        std::vector<double> q;
        std::vector<uint32_t> Z;
        q.reserve(3 * std::min(declared, kReserveAtoms));
        Z.reserve(std::min(declared, kReserveAtoms));
        while (Z.size() < declared) {
            if (!next_line(f, line) || !skip_blank(f, line))
                return io_fail(MSF_ERR_INVALID, where + "fewer atoms than declared");
            std::istringstream s(line);
            size_t number = 0, type = 0;
            double r[3] = {0, 0, 0};
            s >> number >> type >> r[0] >> r[1] >> r[2];
            size_t k = 0;
            for (; k < n_species; ++k)
                if (type == species_id[k]) break;
            if (k >= n_species) return io_fail(MSF_ERR_INVALID, where + "atom type without a mass");
            const uint32_t z = species_z[k];
            if (number > declared || z == 0 || r[0] > hi[0] || r[0] < lo[0] || r[1] > hi[1] ||
                r[1] < lo[1] || r[2] > hi[2] || r[2] < lo[2])
                return io_fail(MSF_ERR_INVALID, where + "atom outside the box or numbered beyond the count");
            for (int c = 0; c < 3; ++c) q.push_back(r[c]);
            Z.push_back(z);
        }
        // distinct atomic numbers must equal the number of atom types (tem_io.cpp:908-960)
        std::vector<uint32_t> uz(Z);
        std::sort(uz.begin(), uz.end());
        uz.erase(std::unique(uz.begin(), uz.end()), uz.end());
        if (uz.size() != n_species)
            return io_fail(MSF_ERR_INVALID, where + "distinct elements differ from the number of atom types");
        // shift by the box's lower bounds (tem_io.cpp:918-923)
        for (size_t a = 0; a < Z.size(); ++a)
            for (int c = 0; c < 3; ++c) q[3 * a + c] = q[3 * a + c] - lo[c];
        return fill_model(out, q, Z, hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
    }
    MSF_IO_CATCH_ALL
}

int msf_write_tiff16(const char* path, const double* data, uint32_t height, uint32_t width,
                     double min, double max, int logscale, uint32_t resolution_unit,
                     double x_resolution, double y_resolution)
{
    try {
        if (!path || !data || height == 0 || width == 0) return io_fail(MSF_ERR_INVALID, "bad image");
        // scaling (write_tif.hpp:230-264)
        const double digits = 65535.0; // 2^16 - 1
        std::vector<uint16_t> img((size_t)height * width);
        if (logscale) {
            const double len_log = std::log(max - min + 1);
            for (size_t i = 0; i < img.size(); ++i) {
                const double t = std::log(data[i] - min + 1) / len_log;
                img[i] = (t > 0.0) ? to_u16(t * digits + 0.5) : (uint16_t)0;
            }
        } else {
            const double scale = digits / (max - min);
            for (size_t i = 0; i < img.size(); ++i) img[i] = to_u16((data[i] - min) * scale + 0.5);
        }
        // baseline little-endian TIFF: header, pixel strips, IFD, rationals.
        // RowsPerStrip is the image width, the value the reference requests
        // from TIFFDefaultStripSize (write_tif.hpp:298-299).
        const uint32_t rps = width;
        const uint32_t n_strips = (height + rps - 1) / rps;
        const size_t row_bytes = (size_t)width * 2;
        const size_t pix_off = 8;
        const size_t pix_bytes = row_bytes * height;
        size_t ifd_off = pix_off + pix_bytes;
        ifd_off += ifd_off & 1;
        const int n_tags = 14;
        const size_t ifd_bytes = 2 + 12 * (size_t)n_tags + 4;
        const size_t extra_off = ifd_off + ifd_bytes;  // rationals, then strip arrays
        const size_t xres_off = extra_off, yres_off = extra_off + 8;
        const size_t offs_off = extra_off + 16;
        const size_t cnts_off = offs_off + 4 * (size_t)n_strips;
        std::vector<uint8_t> b(cnts_off + 4 * (size_t)n_strips, 0);
        b[0] = 'I';
        b[1] = 'I';
        put16(b, 2, 42);
        put32(b, 4, (uint32_t)ifd_off);
        for (uint32_t r = 0; r < height; ++r) { // bottom-up rows (write_tif.hpp:317-321)
            const uint16_t* src = &img[(size_t)(height - r - 1) * width];
            for (uint32_t j = 0; j < width; ++j) put16(b, pix_off + r * row_bytes + 2 * (size_t)j, src[j]);
        }
        for (uint32_t s = 0; s < n_strips; ++s) {
            const uint32_t rows = std::min(rps, height - s * rps);
            put32(b, offs_off + 4 * (size_t)s, (uint32_t)(pix_off + (size_t)s * rps * row_bytes));
            put32(b, cnts_off + 4 * (size_t)s, (uint32_t)(rows * row_bytes));
        }
        uint32_t xn, xd, yn, yd;
        rational(x_resolution, xn, xd);
        rational(y_resolution, yn, yd);
        put32(b, xres_off, xn);
        put32(b, xres_off + 4, xd);
        put32(b, yres_off, yn);
        put32(b, yres_off + 4, yd);
        size_t p = ifd_off;
        put16(b, p, (uint16_t)n_tags);
        p += 2;
        auto tag = [&](uint16_t id, uint16_t type, uint32_t count, uint32_t value) {
            put16(b, p, id);
            put16(b, p + 2, type);
            put32(b, p + 4, count);
            if (type == 3 && count == 1)
                put16(b, p + 8, (uint16_t)value);
            else
                put32(b, p + 8, value);
            p += 12;
        };
        const uint16_t SHORT = 3, LONG = 4, RATIONAL = 5;
        tag(256, LONG, 1, width);
        tag(257, LONG, 1, height);
        tag(258, SHORT, 1, 16);
        tag(259, SHORT, 1, 1);        // no compression
        tag(262, SHORT, 1, 1);        // black is zero
        tag(273, LONG, n_strips, n_strips == 1 ? (uint32_t)pix_off : (uint32_t)offs_off);
        tag(274, SHORT, 1, 1);        // top-left origin
        tag(277, SHORT, 1, 1);
        tag(278, LONG, 1, rps);
        tag(279, LONG, n_strips, n_strips == 1 ? (uint32_t)pix_bytes : (uint32_t)cnts_off);
        tag(282, RATIONAL, 1, (uint32_t)xres_off);
        tag(283, RATIONAL, 1, (uint32_t)yres_off);
        tag(284, SHORT, 1, 1);        // contiguous planes
        tag(296, SHORT, 1, resolution_unit);
        put32(b, p, 0); // no further IFD
        FILE* fo = std::fopen(path, "wb");
        if (!fo) return io_fail(MSF_ERR_INVALID, std::string("Output file could not be opened: ") + path);
        const size_t w = std::fwrite(b.data(), 1, b.size(), fo);
        const int c = std::fclose(fo);
        if (w != b.size() || c != 0) return io_fail(MSF_ERR_INVALID, std::string("short write: ") + path);
        return MSF_OK;
    }
    MSF_IO_CATCH_ALL
}

int msf_write_variance_txt(const char* path, const double* values, const double* lower_edges,
                           uint32_t n_bins)
{
    try {
        if (!path || (n_bins && (!values || !lower_edges))) return io_fail(MSF_ERR_INVALID, "NULL argument");
        FILE* fo = std::fopen(path, "w");
        if (!fo) return io_fail(MSF_ERR_INVALID, std::string("Unable to write variance to file: ") + path);
        for (uint32_t i = 0; i < n_bins; ++i) std::fprintf(fo, "%.10e %.10e\n", lower_edges[i], values[i]);
        if (std::fclose(fo) != 0) return io_fail(MSF_ERR_INVALID, std::string("short write: ") + path);
        return MSF_OK;
    }
    MSF_IO_CATCH_ALL
}

int msf_write_correlograph_txt(const char* path, const double* data, const double* k_edges,
                               uint32_t n_k, const double* phi_edges, uint32_t n_phi)
{
    try {
        if (!path || (n_k && !k_edges) || (n_phi && !phi_edges) || (n_k && n_phi && !data))
            return io_fail(MSF_ERR_INVALID, "NULL argument");
        FILE* fo = std::fopen(path, "w");
        if (!fo) return io_fail(MSF_ERR_INVALID, std::string("Unable to write variance to file: ") + path);
        std::fputs("0 ", fo); // the (0,0) corner, not data
        for (uint32_t jj = 0; jj < n_phi; ++jj) std::fprintf(fo, "%.10e ", phi_edges[jj]);
        std::fputc('\n', fo);
        for (uint32_t ii = 0; ii < n_k; ++ii) {
            std::fprintf(fo, "%.10e ", k_edges[ii]);
            for (uint32_t jj = 0; jj < n_phi; ++jj)
                std::fprintf(fo, "%.10e ", data[(size_t)ii + (size_t)jj * n_k]);
            std::fputc('\n', fo);
        }
        if (std::fclose(fo) != 0) return io_fail(MSF_ERR_INVALID, std::string("short write: ") + path);
        return MSF_OK;
    }
    MSF_IO_CATCH_ALL
}

int msf_write_f64(const char* path, const double* data, uint32_t rows, uint32_t cols)
{
    try {
        if (!path || (rows && cols && !data)) return io_fail(MSF_ERR_INVALID, "NULL argument");
        FILE* fo = std::fopen(path, "wb");
        if (!fo) return io_fail(MSF_ERR_INVALID, std::string("cannot open ") + path);
        std::fprintf(fo, "msf-f64 %u %u\n", rows, cols);
        const size_t n = (size_t)rows * cols;
        const size_t w = n ? std::fwrite(data, sizeof(double), n, fo) : 0;
        if (std::fclose(fo) != 0 || w != n) return io_fail(MSF_ERR_INVALID, std::string("short write: ") + path);
        return MSF_OK;
    }
    MSF_IO_CATCH_ALL
}

} // extern "C"
