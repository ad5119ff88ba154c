// tdw_noalloc.hip -- tdw_normalize_noalloc: tdw_normalize (normalize_wf,
// pca_kernels.cu:955-976) for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It runs the same kernels as tdw_normalize with their
// partial-sum buffer (and, without a caller's norms array, the norms) in the
// file-scope device array `scratch` instead of context memory that
// tdw_normalize grows on demand.
//
// The array is TDW_NOALLOC_SCRATCH_DOUBLES doubles (8 MiB) and lives on
// every device the library runs on until the process ends.  Its device
// address is taken once per device, with the device current, when the
// first context on that device is created (tdw_create calls
// noalloc_setup); host code uses only that address, never the array's name.
// Every call on one device shares the array: calls that overlap on a device
// (two host threads, or work queued on different streams) race.  Nothing in
// it needs to start at zero: each call writes every slot it reads.
#include <mutex>

#include "tdw_internal.hpp"
#include "tdw_physics.hpp"

using namespace tdw;

namespace {

__device__ __align__(256) double scratch[TDW_NOALLOC_SCRATCH_DOUBLES];

constexpr int MAXDEV = 64;
double *g_scratch[MAXDEV];   // device address of `scratch` per device
std::mutex g_mu;

}  // namespace

namespace tdw {

int noalloc_setup(int device)
{
    if (device < 0 || device >= MAXDEV) return TDW_EINVAL;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch[device] = (double *)p;
    return 0;
}

}  // namespace tdw

extern "C" int tdw_normalize_noalloc(tdw_ctx *c, int n, tdw_complex *wf, double *norms)
{
    if (!c || !wf || n <= 0 || n > c->nwf) return TDW_EINVAL;
    double *s = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (c->device >= 0 && c->device < MAXDEV) s = g_scratch[c->device];
    }
    if (!s) return TDW_ENOCTX;
    const size_t part = (size_t)n * (size_t)norm_chunks(c);
    if (part + (norms ? 0 : (size_t)n) > (size_t)TDW_NOALLOC_SCRATCH_DOUBLES) return TDW_ESCRATCH;
    // per-launch timing records create events: off for this call
    TimingOff off(c);
    return normalize_with(c, n, (double2 *)wf, norms ? norms : s + part, s);
}
