// msf_noalloc.cpp -- msf_run_stem_noalloc: the raster run of msf_run_stem
// with every device buffer carved out of one static device array, for
// callers that cannot allocate device memory.  No reference counterpart.
#include "msf_internal.hpp"

#include <map>
#include <mutex>

namespace {

// The region of the allocation-free run.  Each device the library runs on
// holds its own copy until the process ends.  On the host, `scratch` names a
// host-side stand-in, not a device address: only the address returned by
// hipGetSymbolAddress for the current device may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[MSF_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;     // device -> its copy of the array
std::map<int, std::mutex> g_device_mu;  // device -> lock held by a run in its array

} // namespace

namespace msf_impl {

// Called once per context at msf_create with `device` current: looks up the
// address of this device's copy of the array (once per device).
int scratch_attach(int device, char** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return MSF_ERR_DEVICE;
        }
        it = g_scratch_ptr.emplace(device, static_cast<char*>(p)).first;
        g_device_mu[device]; // the device's run lock
    }
    *out = it->second;
    return MSF_OK;
}

} // namespace msf_impl

using namespace msf_impl;

extern "C" {

int msf_run_stem_noalloc(msf_ctx* c, const msf_specimen* sp, const msf_optics* o,
                         const msf_detector* det, const msf_fem* fem, const double* x_p,
                         uint32_t n_xp, const double* y_p, uint32_t n_yp, double* adf_image,
                         double* fem_v)
{
    try {
        if (!c) return MSF_ERR_INVALID;
        MSF_LOCK(c);
        HIPCHK(c, hipSetDevice(c->device));
        if (det && det->enabled && !adf_image) return fail(c, MSF_ERR_INVALID, "adf_image is NULL");
        if (fem && fem->enabled && !fem_v) return fail(c, MSF_ERR_INVALID, "fem_v is NULL");
        if (!sp) return fail(c, MSF_ERR_INVALID, "specimen is NULL");
        if (!x_p || !y_p || n_xp == 0 || n_yp == 0) return fail(c, MSF_ERR_INVALID, "empty raster");
        if (!c->scratch) return fail(c, MSF_ERR_DEVICE, "no scratch region on this device");
        std::vector<int2> shifts;
        if (!raster_shifts(sp->dom, x_p, n_xp, y_p, n_yp, shifts))
            return fail(c, MSF_ERR_DOMAIN, "probe position below the domain minimum");
        RunOut out;
        out.adf = adf_image;
        out.fem = fem_v;
        // Every call on the device shares its one array: runs in it are
        // serialised, and each run synchronises its stream before it returns.
        std::mutex* dm = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            dm = &g_device_mu[c->device];
        }
        std::lock_guard<std::mutex> run_lock(*dm);
        Mem m;
        m.base = c->scratch;
        m.cap = MSF_NOALLOC_SCRATCH_BYTES;
        return run_internal(c, m, sp, o, det, fem, shifts, out);
    }
    MSF_CATCH_ALL(c)
}

} // extern "C"
