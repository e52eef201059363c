// ms-stem-fem -- command line front end over the C ABI (include/msf.h,
// include/msf_io.h).  Replaces the reference executable's main
// (tem.cpp:81-1700) for the ADF-STEM and FEM paths: the same command
// line flags and parameter-file keys (tem_io.cpp:5277-5700, :4063-4500),
// the same required-parameter rules (check_runtime_flags,
// tem_io.cpp:5858-6060), defaults, raster, Scherzer conditions and output
// file names, running on one GPU instead of MPI ranks.
//
// Extensions: --plan_only prints the resolved job and exits before any
// device work; a float64 sidecar (<prefix>..._stem_image_output.f64) holds
// the unquantised ADF image next to the 16-bit TIFFs.
#include "../../include/msf.h"
#include "../../include/msf_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Flags {
    bool o = false, m = false, a = false, pf = false, nx = false, ny = false, voltage = false;
    bool scherzer_defocus = false, scherzer_alphamax = false, scherzer_cs3 = false;
    bool defocus = false, alpha_max = false, spread = false, cs3 = false, cs5 = false;
    bool detector_angles = false, raster_spacing = false;
    bool adfstem_corrected = false, adfstem_uncorrected = false;
    bool bfctem_corrected = false, bfctem_uncorrected = false;
    bool fem = false, d1 = false, d2 = false, d3 = false, d4 = false, gt17 = false, rva = false;
    bool correlograph = false, image_output = false, diffraction_output = false;
    bool correlograph_variance = false, correlograph_everyimage = false, correlograph_everytxt = false;
    bool netcdf_images = false, netcdf_variance = false, pap_tif = false, dupe = false;
    bool mtf_file = false, mtf_resolution = false;
    bool lammps_pre = false, lammps_steps = false, lammps_samples = false;
    bool debug = false, fail = false;
    bool plan_only = false; // extension
};

struct Params {
    std::string prefix, model;
    long Nx = 0, Ny = 0;
    double VV = 0, defocus = 0, alpha_max = 0, defocus_spread = 1.0, condenser_angle = 0;
    double Cs3 = 0, Cs5 = 0, det_inner = 0, det_outer = 0, raster_spacing = 0;
    double dr = 0.70710678118654746; // 1/sqrt(2) (tem.cpp:178)
    double minslice = 1.0;           // (tem.cpp:180)
    unsigned dupe_x = 1, dupe_y = 1, dupe_z = 1;
    std::vector<double> mtf, mtf_domain; // as read, appended per read (tem_io.cpp:5758,5776)
    double mtf_res = 0;
};

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    return s;
}

// read_mtf_file (tem_io.cpp:5722-5800), called where the key or flag is
// parsed; the values are appended as the reference appends them
bool read_mtf(const std::string& name, Params& p)
{
    msf_mtf_table t{};
    if (msf_read_mtf(name.c_str(), &t) != MSF_OK) {
        std::cout << msf_io_last_error() << std::endl;
        std::cout << "Error, could not read MTF file : " << name << std::endl;
        return false;
    }
    p.mtf_domain.insert(p.mtf_domain.end(), t.domain, t.domain + t.n);
    p.mtf.insert(p.mtf.end(), t.mtf, t.mtf + t.n);
    msf_mtf_table_free(&t);
    return true;
}

// read_parameter_file (tem_io.cpp:4063-4500): one key per line, key case
// folded, values read with stream extraction.  As in the reference, the
// loop stops at a line that is not followed by a newline.
bool read_parameter_file(const std::string& name, Flags& f, Params& p)
{
    std::ifstream in(name.c_str());
    if (!in.is_open()) {
        std::cerr << "Error opening parameter file: " << name << std::endl;
        return false;
    }
    std::string line;
    while (std::getline(in, line) && in.good()) {
        if (line.find_first_not_of(" \t") == std::string::npos) continue;
        std::istringstream s(line);
        std::string key;
        s >> key;
        key = lower(key);
        if (key == "debug") f.debug = true;
        else if (key == "output_prefix") { s >> p.prefix; f.o = true; }
        else if (key == "samples_x") { int v = 0; s >> v; p.Nx = v; f.nx = true; }
        else if (key == "samples_y") { int v = 0; s >> v; p.Ny = v; f.ny = true; }
        else if (key == "model_file") { s >> p.model; f.a = true; }
        else if (key == "microscope_voltage") { s >> p.VV; f.voltage = true; }
        else if (key == "scherzer_defocus") f.scherzer_defocus = true;
        else if (key == "scherzer_alphamax") f.scherzer_alphamax = true;
        else if (key == "scherzer_cs3") f.scherzer_cs3 = true;
        else if (key == "defocus") { s >> p.defocus; f.defocus = true; }
        else if (key == "alphamax") { s >> p.alpha_max; f.alpha_max = true; }
        else if (key == "spread") { s >> p.defocus_spread >> p.condenser_angle; f.spread = true; }
        else if (key == "cs3") { s >> p.Cs3; f.cs3 = true; }
        else if (key == "cs5") { s >> p.Cs5; f.cs5 = true; }
        else if (key == "detectorangles") {
            s >> p.det_inner >> p.det_outer;
            if (p.det_inner > p.det_outer) std::swap(p.det_inner, p.det_outer);
            f.detector_angles = true;
        }
        else if (key == "mtf_file") {
            std::string name;
            s >> name;
            if (read_mtf(name, p)) f.mtf_file = true;
            else f.fail = true;
        }
        else if (key == "mtf_sampling_rate") {
            p.mtf_res = 0;
            if (!s.eof()) s >> p.mtf_res;
            f.mtf_resolution = p.mtf_res != 0;
        }
        else if (key == "raster_spacing") { s >> p.raster_spacing; f.raster_spacing = true; }
        else if (key == "adfstemuncorrfem") { f.adfstem_uncorrected = true; f.fem = true; }
        else if (key == "adfstemcorrfem") { f.adfstem_corrected = true; f.fem = true; }
        else if (key == "v_gt17") f.gt17 = true;
        else if (key == "v_omega") f.d1 = true;
        else if (key == "vbar_r") f.d2 = true;
        else if (key == "v_re") f.d3 = true;
        else if (key == "omega_vimage") f.d4 = true;
        else if (key == "v_rbar") f.rva = true;
        else if (key == "adfstemcorr") f.adfstem_corrected = true;
        else if (key == "adfstemuncorr") f.adfstem_uncorrected = true;
        else if (key == "bfctemcorr") f.bfctem_corrected = true;
        else if (key == "bfctemuncorr") f.bfctem_uncorrected = true;
        else if (key == "paptif") f.pap_tif = true;
        else if (key == "dupe") f.dupe = true;
        else if (key == "dr") s >> p.dr;
        else if (key == "minslice") s >> p.minslice;
        else if (key == "images") f.image_output = true;
        else if (key == "diffractionimages") f.diffraction_output = true;
        else if (key == "netcdfimages") f.netcdf_images = true;
        else if (key == "netcdfvariance") f.netcdf_variance = true;
        else if (key == "correlograph") f.correlograph = true;
        else if (key == "correlograph_variance") f.correlograph = f.correlograph_variance = true;
        else if (key == "correlograph_everyimage") f.correlograph = f.correlograph_everyimage = true;
        else if (key == "correlograph_everytxt") f.correlograph = f.correlograph_everytxt = true;
        else if (key == "lammps_pretem_file_name") f.lammps_pre = true;
        else if (key == "lammps_tem_steps") f.lammps_steps = true;
        else if (key == "lammps_tem_samples") f.lammps_samples = true;
        else {
            std::cerr << "Error, unexpected argument in parameter file: " << line << std::endl;
            return false;
        }
    }
    if (f.fem && !(f.d1 || f.d2 || f.d3 || f.d4 || f.gt17 || f.rva || f.correlograph)) f.d1 = true;
    return true;
}

// read_cmdline_options (tem_io.cpp:5277-5700).  Two deliberate differences:
// --detectorangles takes two values (the reference reads the first value
// twice and then rejects the second as an unexpected argument), and
// --mtf_file consumes its file name (the reference reads the file but does
// not advance past the name, which it then rejects as an unexpected argument).
void read_cmdline(const std::vector<std::string>& args, Flags& f, Params& p)
{
    auto num = [&](size_t i, auto& v) {
        if (i < args.size()) std::istringstream(args[i]) >> v;
    };
    for (size_t i = 1; i < args.size(); ++i) {
        const std::string& a = args[i];
        if (a == "--parameter_file") {
            const std::string name = i + 1 < args.size() ? args[i + 1] : std::string();
            if (!read_parameter_file(name, f, p)) {
                std::cout << "Error, could not read parameter file : " << name << std::endl;
                f.fail = true;
            }
            f.m = f.voltage && f.nx && f.ny;
            f.pf = true;
            i += 1;
        }
        else if (a == "--mtf_file") {
            const std::string name = i + 1 < args.size() ? args[i + 1] : std::string();
            if (read_mtf(name, p)) f.mtf_file = true;
            else f.fail = true;
            i += 1;
        }
        else if (a == "--mtf_sampling_rate") {
            p.mtf_res = 0;
            if (i + 1 < args.size()) {
                num(i + 1, p.mtf_res);
                i += 1;
            }
            f.mtf_resolution = p.mtf_res > 0;
            if (!f.mtf_resolution) std::cout << " Warning: attempted to use MTF sampling rate <= 0" << std::endl;
        }
        else if (a == "--debug") f.debug = true;
        else if (a == "--output_prefix") { if (i + 1 < args.size()) p.prefix = args[i + 1]; f.o = true; i += 1; }
        else if (a == "-m") {
            int nx = 0, ny = 0;
            num(i + 1, nx);
            num(i + 2, ny);
            num(i + 3, p.VV);
            p.Nx = nx;
            p.Ny = ny;
            f.m = true;
            i += 3;
        }
        else if (a == "-a") { if (i + 1 < args.size()) p.model = args[i + 1]; f.a = true; i += 1; }
        else if (a == "--scherzer_defocus") f.scherzer_defocus = true;
        else if (a == "--scherzer_alphamax") f.scherzer_alphamax = true;
        else if (a == "--scherzer_cs3") f.scherzer_cs3 = true;
        else if (a == "--defocus") { num(i + 1, p.defocus); f.defocus = true; i += 1; }
        else if (a == "--alphamax") { num(i + 1, p.alpha_max); f.alpha_max = true; i += 1; }
        else if (a == "--spread") { num(i + 1, p.defocus_spread); num(i + 2, p.condenser_angle); f.spread = true; i += 2; }
        else if (a == "--cs3") { num(i + 1, p.Cs3); f.cs3 = true; i += 1; }
        else if (a == "--cs5") { num(i + 1, p.Cs5); f.cs5 = true; i += 1; }
        else if (a == "--detectorangles") {
            num(i + 1, p.det_inner);
            num(i + 2, p.det_outer);
            if (p.det_inner > p.det_outer) std::swap(p.det_inner, p.det_outer);
            f.detector_angles = true;
            i += 2;
        }
        else if (a == "--rasterspacing") { num(i + 1, p.raster_spacing); f.raster_spacing = true; i += 1; }
        else if (a == "--adfstemuncorrfem") { f.adfstem_uncorrected = true; f.fem = true; }
        else if (a == "--adfstemcorrfem") { f.adfstem_corrected = true; f.fem = true; }
        else if (a == "--V_gt17") f.gt17 = true;
        else if (a == "--V_omega") f.d1 = true;
        else if (a == "--Vbar_r") f.d2 = true;
        else if (a == "--V_re") f.d3 = true;
        else if (a == "--Omega_vimage") f.d4 = true;
        else if (a == "--V_rbar") f.rva = true;
        else if (a == "--adfstemcorr") f.adfstem_corrected = true;
        else if (a == "--adfstemuncorr") f.adfstem_uncorrected = true;
        else if (a == "--bfctemcorr") f.bfctem_corrected = true;
        else if (a == "--bfctemuncorr") f.bfctem_uncorrected = true;
        else if (a == "--paptif") f.pap_tif = true;
        else if (a == "--dupe") { num(i + 1, p.dupe_x); num(i + 2, p.dupe_y); num(i + 3, p.dupe_z); f.dupe = true; i += 3; }
        else if (a == "--dr") { num(i + 1, p.dr); i += 1; }
        else if (a == "--minslice") { num(i + 1, p.minslice); i += 1; }
        else if (a == "--images") f.image_output = true;
        else if (a == "--netcdfimages") f.netcdf_images = true;
        else if (a == "--netcdfvariance") f.netcdf_variance = true;
        else if (a == "--correlograph") f.correlograph = true;
        else if (a == "--correlograph_variance") f.correlograph = f.correlograph_variance = true;
        else if (a == "--correlograph_everyimage") f.correlograph = f.correlograph_everyimage = true;
        else if (a == "--correlograph_everytxt") f.correlograph = f.correlograph_everytxt = true;
        else if (a == "--lammps_preTEM_file_name") { f.lammps_pre = true; i += 1; }
        else if (a == "--lammps_TEM_steps") { f.lammps_steps = true; i += 1; }
        else if (a == "--lammps_TEM_samples") { f.lammps_samples = true; i += 1; }
        else if (a == "--plan_only") f.plan_only = true;
        else {
            std::cerr << "Error, unexpected argument : " << a << std::endl;
            f.fail = true;
        }
    }
    if (!f.gt17 && !f.d1 && !f.d2 && !f.d3 && !f.d4 && !f.rva && !f.correlograph) f.d1 = true;
}

// check_runtime_flags (tem_io.cpp:5858-6060)
bool check_runtime_flags(const Flags& f, const std::string& argv0)
{
    const bool corr = f.adfstem_corrected || f.bfctem_corrected;
    const bool uncorr = f.adfstem_uncorrected || f.bfctem_uncorrected;
    const bool ok = f.o && f.m && !(corr && !(f.cs5 && (f.scherzer_cs3 || f.cs3))) &&
                    !(uncorr && !f.cs3) && f.mtf_file == f.mtf_resolution &&
                    (f.scherzer_defocus || f.defocus) && (f.scherzer_alphamax || f.alpha_max) &&
                    f.lammps_pre == f.lammps_steps && f.lammps_pre == f.lammps_samples;
    if (!ok) {
        std::cout << argv0 << " : lacking required parameters" << std::endl;
        if (corr && !((f.cs3 || f.scherzer_cs3) && f.cs5))
            std::cout << "Use of aberraction correction requires specifying --defocus_spread and "
                         "--cs5. If not using --scherzer_cs3, then --cs3 is also required."
                      << std::endl;
    }
    return ok;
}

// Features of the reference outside this build: rejected with a message
// rather than silently ignored.
bool check_supported(const Flags& f)
{
    struct Item { bool on; const char* what; };
    const Item items[] = {
        {f.bfctem_corrected || f.bfctem_uncorrected, "--bfctem*"},
        {f.pap_tif, "--paptif"},
        {f.diffraction_output, "diffractionimages"},
        {f.netcdf_images || f.netcdf_variance, "--netcdf*"},
        {f.lammps_pre, "--lammps_* (molecular dynamics coupling)"}};
    bool ok = true;
    for (const Item& it : items)
        if (it.on) {
            std::cerr << "Error, " << it.what << " is not supported by this build" << std::endl;
            ok = false;
        }
    return ok;
}

std::string to_string_g(double v)
{
    std::ostringstream s; // default stream format, as TEM_NS::to_string (to_string.hpp:32-37)
    s << v;
    return s.str();
}

int die(const char* what, msf_ctx* c)
{
    std::cerr << "Error, " << what << ": " << (c ? msf_last_error(c) : "") << std::endl;
    return EXIT_FAILURE;
}

// log(1 + s |psi|^2) of a complex wave (write_tif.cpp:817-1030) or
// log(1 + s v^2) of a real array (write_tif.cpp:1413-1548), quadrants
// swapped, then a linear 16-bit min-max TIFF.  NaN pixels (0/0 in a 2-D
// variance) are left out of the min and max, as by the reference's
// comparisons, and written as 0.
int write_log_diffraction(const std::string& prefix, std::vector<double>& mag, long nx, long ny,
                          double s, double xres, double yres)
{
    for (double& v : mag) v = std::log(1.0 + (s * v));
    double mn = mag[0], mx = mag[0];
    for (double v : mag) {
        if (v > mx) mx = v;
        if (v < mn) mn = v;
    }
    for (long i = 0; i < nx / 2; ++i) {
        for (long j = 0; j < ny / 2; ++j) std::swap(mag[j + i * ny], mag[(j + ny / 2) + (i + nx / 2) * ny]);
        for (long j = ny / 2; j < ny; ++j) std::swap(mag[j + i * ny], mag[(j - ny / 2) + (i + nx / 2) * ny]);
    }
    const std::string name = prefix + "_diffraction_scalefactor_" + to_string_g(s) + ".tif";
    return msf_write_tiff16(name.c_str(), mag.data(), (uint32_t)nx, (uint32_t)ny, mn, mx, 0, 1, xres, yres);
}

int write_diffraction(const std::string& prefix, const float* psi, long nx, long ny, double s,
                      double xres, double yres)
{
    std::vector<double> i2((size_t)nx * ny);
    for (size_t k = 0; k < i2.size(); ++k) {
        const double re = psi[2 * k], im = psi[2 * k + 1];
        i2[k] = re * re + im * im;
    }
    return write_log_diffraction(prefix, i2, nx, ny, s, xres, yres);
}

int write_diffraction_real(const std::string& prefix, const std::vector<double>& v, long nx,
                           long ny, double s, double xres, double yres)
{
    std::vector<double> i2(v.size());
    for (size_t k = 0; k < v.size(); ++k) i2[k] = v[k] * v[k];
    return write_log_diffraction(prefix, i2, nx, ny, s, xres, yres);
}

// Correlograph outputs (adfstem.cpp:3617-3745): the average as text and a
// linear 16-bit image (phi along x, k along y); with the variance flag, the
// text file of the sums of squares (the reference writes those under the
// variance name) and log and linear images of the variance over the k bins
// from the one holding 0.2 up to, not including, the one holding 1.0 1/A.
// Empty images are skipped (the reference reads element 0 of an empty array).
// output_correlograph_image (write_tif.cpp:1607-1654): min-max 16-bit TIFF,
// phi along x (w columns), k along y (h rows); the resolution tags come from
// the full bin counts nk and nphi (adfstem.cpp:2651-2654)
int write_correlograph_image(const double* v, uint32_t w, uint32_t h, const std::string& name,
                             int lg, uint32_t nk, uint32_t nphi)
{
    if ((size_t)w * h == 0) return MSF_OK;
    const double kres = (nk / (1.0 - 0.2)) * 1.0e8;
    const double phires = (double)(nphi / 360);  // integer division, as in the reference
    double mn = v[0], mx = v[0];
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        if (v[i] > mx) mx = v[i];
        if (v[i] < mn) mn = v[i];
    }
    return msf_write_tiff16(name.c_str(), v, w, h, mn, mx, lg, 3, phires, kres);
}

int write_correlograph(const std::string& out, const std::vector<double>& avg,
                       const std::vector<double>& sqr, const std::vector<double>& k_edges,
                       const std::vector<double>& phi_edges, uint32_t nk, uint32_t nphi,
                       size_t P, bool variance)
{
    auto image = [&](const double* v, uint32_t w, uint32_t h, const std::string& name, int lg) {
        return write_correlograph_image(v, w, h, name, lg, nk, nphi);
    };
    int w = msf_write_correlograph_txt((out + "_correlograph_avg.txt").c_str(), avg.data(),
                                       k_edges.data(), nk, phi_edges.data(), nphi);
    w |= image(avg.data(), nphi, nk, out + "_correlograph_avg.tif", 0);
    if (!variance) return w;
    size_t start = 0, end = nk ? nk - 1 : 0;
    for (size_t ii = 0; ii < nk; ++ii) {
        if (k_edges[ii] <= 0.2 && k_edges[ii + 1] > 0.2) start = ii;
        if (k_edges[ii] <= 1.0 && k_edges[ii + 1] > 1.0) end = ii;
    }
    const size_t nred = (nk && end >= start) ? end - start : 0;
    std::vector<double> var(nred * nphi);
    for (size_t ii = start; ii < start + nred; ++ii)
        for (size_t jj = 0; jj < nphi; ++jj) {
            const size_t ci = ii + jj * nk;
            const double t = sqr[ci] / (P * avg[ci] * avg[ci]);
            var[(ii - start) + jj * nred] =
                (sqr[ci] != 0 && P != 0 && avg[ci] != 0 && t >= 1) ? t - 1 : 0;
        }
    w |= msf_write_correlograph_txt((out + "_correlograph_variance.txt").c_str(), sqr.data(),
                                    k_edges.data(), nk, phi_edges.data(), nphi);
    w |= image(var.data(), nphi, (uint32_t)nred, out + "_correlograph_variance_logscale.tif", 1);
    w |= image(var.data(), nphi, (uint32_t)nred, out + "_correlograph_variance.tif", 0);
    return w;
}

} // namespace

int main(int argc, char** argv)
{
    std::vector<std::string> args(argv, argv + argc);
    Flags f;
    Params p;
    read_cmdline(args, f, p);

    // model file: ".xyz" after the last '.' selects the xyz reader (tem.cpp:270-330)
    msf_model model{};
    bool have_model = false;
    if (!f.fail) {
        const size_t dot = p.model.find_last_of('.');
        const std::string ext = p.model.substr(dot == std::string::npos ? 0 : dot + 1);
        const int r = (ext == "xyz") ? msf_read_xyz(p.model.c_str(), &model)
                                     : msf_read_lammps_data(p.model.c_str(), &model);
        if (r != MSF_OK) {
            std::cerr << msf_io_last_error() << std::endl;
            std::cerr << "Error, file could not be read : " << p.model << std::endl;
            f.fail = true;
        } else {
            have_model = true;
        }
    }
    if (!check_runtime_flags(f, args[0])) f.fail = true;
    if (!f.fail && !check_supported(f)) f.fail = true;
    if (f.fail) {
        if (have_model) msf_model_free(&model);
        std::cout << "Exiting ms-stem-fem due to failure." << std::endl;
        return EXIT_FAILURE;
    }

    // --dupe replicates the cell for FEM only (tem.cpp:1053-1117)
    const bool dupe = f.dupe && f.fem && (p.dupe_x > 1 || p.dupe_y > 1 || p.dupe_z > 1) &&
                      (p.dupe_x > 0 && p.dupe_y > 0 && p.dupe_z > 0);
    const unsigned dx = dupe ? p.dupe_x : 1, dy = dupe ? p.dupe_y : 1, dz = dupe ? p.dupe_z : 1;
    const uint32_t n0 = model.n_atoms;
    const size_t nd = (size_t)n0 * dx * dy * dz;
    std::vector<double> q(3 * nd);
    std::vector<uint32_t> Z(nd);
    for (unsigned ii = 0; ii < dx; ++ii)
        for (unsigned jj = 0; jj < dy; ++jj)
            for (unsigned kk = 0; kk < dz; ++kk)
                for (uint32_t a = 0; a < n0; ++a) {
                    const size_t o = a + (size_t)n0 * (kk + dz * (jj + (size_t)dy * ii));
                    q[3 * o] = model.q[3 * a] + ii * model.xperiod;
                    q[3 * o + 1] = model.q[3 * a + 1] + jj * model.yperiod;
                    q[3 * o + 2] = model.q[3 * a + 2] + kk * model.zperiod;
                    Z[o] = model.Z[a];
                }
    msf_grid grid{};
    grid.nx = (uint32_t)p.Nx;
    grid.ny = (uint32_t)p.Ny;
    grid.xperiod = model.xperiod * dx;
    grid.yperiod = model.yperiod * dy;
    grid.zperiod = model.zperiod * dz;
    grid.voltage = p.VV;
    const double xperiod = model.xperiod, yperiod = model.yperiod;
    msf_model_free(&model);

    double lambda = 0, gamma = 0;
    msf_wavelength(p.VV, &lambda, &gamma);
    // band limits (tem.cpp:683-695) for the detector check (:710-732)
    const double bw_pap = (p.Nx / (2 * grid.xperiod) < p.Ny / (2 * grid.yperiod))
                              ? p.Nx / (2 * grid.xperiod) : p.Ny / (2 * grid.yperiod);
    const double bw_t = (2.0 / 3.0) * bw_pap;
    const bool adf_mode = f.adfstem_corrected || f.adfstem_uncorrected;
    if (adf_mode && f.detector_angles) {
        bool bad = false;
        if (p.det_inner < 0.0) { std::cout << "Error: detector_inner_angle < 0" << std::endl; bad = true; }
        if (p.det_outer < 0.0) { std::cout << "Error: detector_outer_angle < 0" << std::endl; bad = true; }
        if (p.det_inner > bw_t * lambda) {
            std::cout << "Error: detector inner angle falls beyond bandwidth limits" << std::endl
                      << " detector angles: " << p.det_inner << ", " << p.det_outer << std::endl
                      << " bandwidth limit: " << bw_t * lambda << std::endl;
            bad = true;
        }
        if (bad) return EXIT_FAILURE;
    }
    if (!adf_mode) return EXIT_SUCCESS; // nothing else of the reference runs on this path
    if (!f.detector_angles) { // defaults (tem.cpp:1324-1337)
        p.det_inner = p.VV >= 200000 ? 0.040 : 0.045;
        p.det_outer = 0.200;
    }
    if (!f.raster_spacing) p.raster_spacing = f.fem ? 10 : 1.5; // (tem.cpp:1339-1345)
    // raster of the undupled cell (tem.cpp:1357-1378, func.cpp:108-127)
    const int nrx = (int)(size_t)std::floor(xperiod / p.raster_spacing);
    const int nry = (int)(size_t)std::floor(yperiod / p.raster_spacing);
    std::vector<double> xp, yp;
    for (int i = 0; i < nrx; ++i) xp.push_back(0.0 + (i * (xperiod / nrx)));
    for (int j = 0; j < nry; ++j) yp.push_back(0.0 + (j * (yperiod / nry)));

    struct Mode { bool on; int corrected; const char* suffix; };
    const Mode modes[2] = {{f.adfstem_uncorrected, 0, "_adfstem_uncorr"},
                           {f.adfstem_corrected, 1, "_adfstem_corr"}};
    std::vector<msf_optics> optics;
    for (const Mode& m : modes) {
        if (!m.on) continue;
        msf_optics o{};
        o.corrected = m.corrected;
        o.defocus = p.defocus;
        o.alpha_max = p.alpha_max;
        o.cs5 = p.Cs5;
        if (!m.corrected) { // tem.cpp:1395-1410
            o.cs3 = p.Cs3;
            if (f.scherzer_defocus) o.defocus = msf_scherzer_defocus(0, lambda, p.Cs3);
            if (f.scherzer_alphamax) o.alpha_max = msf_scherzer_alphamax(0, lambda, p.Cs3);
        } else {            // tem.cpp:1455-1480
            if (f.scherzer_defocus) o.defocus = msf_scherzer_defocus(1, lambda, p.Cs5);
            if (f.scherzer_alphamax) o.alpha_max = msf_scherzer_alphamax(1, lambda, p.Cs5);
            o.cs3 = f.scherzer_cs3 ? msf_scherzer_cs3(lambda, p.Cs5) : p.Cs3;
        }
        optics.push_back(o);
    }

    if (f.plan_only) {
        std::printf("grid %ld %ld %.17g %.17g %.17g\n", p.Nx, p.Ny, grid.xperiod, grid.yperiod, grid.zperiod);
        std::printf("atoms %zu\nvoltage %.17g\nlambda %.17g\nminslice %.17g\ndr %.17g\n", nd, p.VV,
                    lambda, p.minslice, p.dr);
        std::printf("raster %d %d %.17g\ndetector %.17g %.17g\n", nrx, nry, p.raster_spacing,
                    p.det_inner, p.det_outer);
        std::printf("images %d\nfem %d\n", f.image_output ? 1 : 0, (f.fem && f.d1) ? 1 : 0);
        std::printf("fem_modes%s%s%s%s%s%s%s%s\n", (f.fem && f.d1) ? " V_omega" : "",
                    (f.fem && f.d2) ? " Vbar_r" : "", (f.fem && f.d3) ? " V_re" : "",
                    (f.fem && f.d4) ? " Omega_vimage" : "", (f.fem && f.gt17) ? " V_gt17" : "",
                    (f.fem && f.rva) ? " V_rbar" : "",
                    (f.fem && f.correlograph) ? " correlograph" : "",
                    (f.fem && f.correlograph_variance) ? " correlograph_variance" : "");
        size_t k = 0;
        for (const Mode& m : modes)
            if (m.on) {
                const msf_optics& o = optics[k++];
                std::printf("mode %s cs3 %.17g cs5 %.17g defocus %.17g alphamax %.17g\n",
                            m.suffix + 1, o.cs3, o.cs5, o.defocus, o.alpha_max);
            }
        const bool mtf_on = f.mtf_file && f.mtf_resolution;
        std::printf("mtf %zu %.17g\n", mtf_on ? p.mtf.size() : (size_t)0, mtf_on ? p.mtf_res : 0.0);
        std::printf("prefix %s\n", p.prefix.c_str());
        return EXIT_SUCCESS;
    }
    if (nrx < 1 || nry < 1) {
        std::cerr << "Error, raster spacing leaves no probe positions" << std::endl;
        return EXIT_FAILURE;
    }

    msf_ctx* c = nullptr;
    if (msf_create(&c, 0) != MSF_OK) {
        std::cerr << "Error, no usable GPU" << std::endl;
        return EXIT_FAILURE;
    }
    // the tuning file of this grid, read before the run and written after it,
    // as the reference reads and writes its FFTW wisdom file per grid
    // (tem.cpp:588-616, :1749-1765); a missing file is not an error
    const std::string tuning_file =
        "msf_tuning_" + std::to_string(p.Nx) + "x" + std::to_string(p.Ny) + ".txt";
    (void)msf_tuning_load(tuning_file.c_str(), nullptr);
    msf_specimen* spec = nullptr;
    if (msf_specimen_build(c, &grid, q.data(), Z.data(), (uint32_t)nd, p.minslice, &spec) != MSF_OK)
        return die("specimen", c);
    msf_detector det{};
    det.enabled = f.image_output ? 1 : 0; // the ADF sum runs only with --images (adfstem.cpp:2245)
    det.inner = p.det_inner;
    det.outer = p.det_outer;
    det.threshold_mode = MSF_ADF_REFERENCE;
    msf_fem fem{};
    const bool corr = f.fem && f.correlograph;
    fem.enabled = (f.fem && (f.d1 || f.d2 || f.d3 || f.d4 || f.gt17 || f.rva || corr)) ? 1 : 0;
    fem.dr = p.dr;
    // detector MTF (adfstem.cpp:238, :2272): both the table and a sampling rate
    const bool mtf = f.mtf_file && f.mtf_resolution;
    if (mtf) {
        fem.mtf = p.mtf.data();
        fem.mtf_domain = p.mtf_domain.data();
        fem.mtf_n = (uint32_t)p.mtf.size();
        fem.mtf_sampling_rate = p.mtf_res;
    }
    uint32_t nbins = 0, nphi = 0;
    std::vector<double> edges, k_edges, phi_edges;
    if (fem.enabled) {
        if (msf_fem_bins(&grid, p.dr, &nbins, nullptr) != MSF_OK) return die("fem bins", c);
        edges.resize(std::max<uint32_t>(nbins, 1));
        msf_fem_bins(&grid, p.dr, &nbins, edges.data());
    }
    if (corr) {
        k_edges.resize(nbins + 1);
        if (msf_correlograph_bins(&grid, p.dr, &nphi, k_edges.data(), nullptr) != MSF_OK)
            return die("correlograph bins", c);
        phi_edges.resize(nphi + 1);
        msf_correlograph_bins(&grid, p.dr, &nphi, k_edges.data(), phi_edges.data());
    }
    std::vector<int32_t> shifts(2 * xp.size() * yp.size());
    if (msf_probe_shifts(&grid, xp.data(), (uint32_t)xp.size(), yp.data(), (uint32_t)yp.size(),
                         shifts.data()) != MSF_OK)
        return die("probe shifts", c);

    size_t k = 0;
    int rc = EXIT_SUCCESS;
    for (const Mode& m : modes) {
        if (!m.on) continue;
        const msf_optics& o = optics[k++];
        const std::string out = p.prefix + m.suffix;
        std::vector<double> img(xp.size() * yp.size());
        // one output array per requested FEM mode, with its file suffix
        // (adfstem.cpp:2886, :2932, :3311, :3418, :3454, :3601)
        struct FemFile { bool on; const char* suffix; std::vector<double> v; };
        FemFile files[6] = {{f.fem && f.d1, "_fem_V_omega", {}}, {f.fem && f.d2, "_fem_Vbar_r", {}},
                            {f.fem && f.d3, "_fem_V_re", {}}, {f.fem && f.d4, "_fem_Omega_vimage", {}},
                            {f.fem && f.gt17, "_fem_V_gt17", {}}, {f.fem && f.rva, "_fem_V_rbar", {}}};
        double* ptr[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        for (int m = 0; m < 6; ++m)
            if (files[m].on) {
                files[m].v.assign(std::max<uint32_t>(nbins, 1), 0.0);
                ptr[m] = files[m].v.data();
            }
        const size_t ncorr = std::max<size_t>((size_t)nbins * nphi, 1);
        std::vector<double> c_avg(corr ? ncorr : 0), c_sqr(corr ? ncorr : 0);
        const bool every = corr && (f.correlograph_everyimage || f.correlograph_everytxt);
        std::vector<double> c_each(every ? xp.size() * yp.size() * ncorr : 0);
        // per-pixel averages over the raster and the 2-D variance, written as
        // images (adfstem.cpp:2966-3099, :3468-3520)
        const size_t npix = (size_t)p.Nx * p.Ny;
        const bool img_avg = f.fem && f.image_output && (f.gt17 || f.d3 || f.d4 || f.rva);
        const bool img_sqr = img_avg && (f.gt17 || f.d3 || f.d4);
        const bool img_var = f.fem && f.image_output && f.d4;
        std::vector<double> p_abs(img_avg ? npix : 0), p_sqr(img_sqr ? npix : 0), p_var(img_var ? npix : 0);
        const msf_fem_outputs fo{ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5],
                                 corr ? c_avg.data() : nullptr, corr ? c_sqr.data() : nullptr,
                                 every ? c_each.data() : nullptr, img_avg ? p_abs.data() : nullptr,
                                 img_sqr ? p_sqr.data() : nullptr, img_var ? p_var.data() : nullptr};
        if (msf_run_stem_fem(c, spec, &o, &det, &fem, xp.data(), (uint32_t)xp.size(), yp.data(),
                             (uint32_t)yp.size(), img.data(), &fo) != MSF_OK)
            return die("adfstem", c);
        if (files[0].on) // adfstem.cpp:2851-2856
            for (uint32_t b = 0; b < nbins; ++b)
                if (files[0].v[b] < 0.0)
                    std::cout << " WARNING, fem --V_omega, ftem_d1[" << b << "] < 0.0 : " << files[0].v[b]
                              << std::endl;
        for (int m = 0; m < 6; ++m)
            if (files[m].on &&
                msf_write_variance_txt((out + files[m].suffix + ".txt").c_str(), files[m].v.data(),
                                       edges.data(), nbins))
                std::cout << " failed to write variance data to a txt file : " << out << std::endl;
        // every position's correlograph, <prefix>_<x index>_<y index>.tif / .txt
        // (adfstem.cpp:2655-2685)
        for (size_t ix = 0; every && ix < xp.size(); ++ix)
            for (size_t iy = 0; iy < yp.size(); ++iy) {
                const double* cp = c_each.data() + (ix * yp.size() + iy) * ncorr;
                const std::string name = out + "_" + std::to_string(ix) + "_" + std::to_string(iy);
                int w = MSF_OK;
                if (f.correlograph_everyimage && nbins * nphi > 0)
                    w |= write_correlograph_image(cp, nphi, nbins, name + ".tif", 0, nbins, nphi);
                if (f.correlograph_everytxt)
                    w |= msf_write_correlograph_txt((name + ".txt").c_str(), cp, k_edges.data(),
                                                    nbins, phi_edges.data(), nphi);
                if (w) {
                    std::cerr << "Error writing the correlograph: " << msf_io_last_error() << std::endl;
                    rc = EXIT_FAILURE;
                }
            }
        {
            const double rx = grid.xperiod * 1.0e-8, ry = grid.yperiod * 1.0e-8;
            int w = MSF_OK;
            if (img_avg) w |= write_diffraction_real(out + "_diffracted_wave_avg", p_abs, p.Nx, p.Ny, 1.0e13, rx, ry);
            if (img_sqr)
                w |= write_diffraction_real(out + "_diffracted_wave_sqr_avg", p_sqr, p.Nx, p.Ny, 1.0e13, rx, ry);
            for (double s : {1.0e0, 1.0e-10})
                if (img_var) w |= write_diffraction_real(out + "_variance_2-D", p_var, p.Nx, p.Ny, s, rx, ry);
            if (w) {
                std::cerr << "Error writing images: " << msf_io_last_error() << std::endl;
                rc = EXIT_FAILURE;
            }
        }
        if (corr && write_correlograph(out, c_avg, c_sqr, k_edges, phi_edges, nbins, nphi,
                                       xp.size() * yp.size(), f.correlograph_variance)) {
            std::cerr << "Error writing the correlograph: " << msf_io_last_error() << std::endl;
            rc = EXIT_FAILURE;
        }
        if (f.image_output) { // adfstem.cpp:3764-3797, write_tif.cpp:1540-1600
            double mn = img[0], mx = img[0];
            for (double a : img) {
                if (a > mx) mx = a;
                if (a < mn) mn = a;
            }
            const double xres = (xp.size() / xperiod) * 1.0e8, yres = (yp.size() / yperiod) * 1.0e8;
            const std::string base = out + "_stem_image_output";
            int w = msf_write_tiff16((base + ".tif").c_str(), img.data(), (uint32_t)xp.size(),
                                     (uint32_t)yp.size(), mn, mx, 0, 3, xres, yres);
            w |= msf_write_tiff16((base + "_logscale.tif").c_str(), img.data(), (uint32_t)xp.size(),
                                  (uint32_t)yp.size(), mn, mx, 1, 3, xres, yres);
            w |= msf_write_f64((base + ".f64").c_str(), img.data(), (uint32_t)xp.size(), (uint32_t)yp.size());
            // diffraction of the first position at five scale factors and of
            // the last position (adfstem.cpp:2140-2230, :3780-3797)
            const size_t P = xp.size() * yp.size();
            std::vector<int32_t> sel = {shifts[0], shifts[1], shifts[2 * (P - 1)], shifts[2 * (P - 1) + 1]};
            std::vector<float> ex(2 * (size_t)p.Nx * p.Ny * 2);
            if (msf_exit_waves_shifted(c, spec, &o, sel.data(), 2, ex.data()) != MSF_OK)
                return die("exit waves", c);
            const double rx = grid.xperiod * 1.0e-8, ry = grid.yperiod * 1.0e-8;
            for (double s : {1.0e-5, 1.0e-1, 1.0e0, 1.0e10, 1.0e20})
                w |= write_diffraction(out + "_first", ex.data(), p.Nx, p.Ny, s, rx, ry);
            if (mtf) { // the final wave has been through the MTF (adfstem.cpp:2272-2385)
                const size_t M = (size_t)p.Nx * p.Ny;
                std::vector<double> w64(2 * M);
                for (size_t k = 0; k < 2 * M; ++k) w64[k] = ex[2 * M + k];
                if (msf_apply_mtf(c, &grid, &fem, o.alpha_max, w64.data(), 1) != MSF_OK)
                    return die("mtf", c);
                for (size_t k = 0; k < 2 * M; ++k) ex[2 * M + k] = (float)w64[k];
            }
            w |= write_diffraction(out + "_final", ex.data() + 2 * (size_t)p.Nx * p.Ny, p.Nx, p.Ny, 1.0, rx, ry);
            if (w) {
                std::cerr << "Error writing images: " << msf_io_last_error() << std::endl;
                rc = EXIT_FAILURE;
            }
        }
    }
    msf_specimen_free(spec);
    msf_destroy(c);
    if (msf_tuning_save(tuning_file.c_str()) != MSF_OK)
        std::cerr << "Could not write tuning file: " << tuning_file << std::endl;
    return rc;
}
