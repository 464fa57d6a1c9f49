// bh_noalloc.hip -- bh_set_potential_noalloc: the import of a host potential
// into a plan's warm-start state through a fixed device staging array, for
// callers that may not allocate device memory.  bh_set_potential from a host
// array stages the whole (numr, numz, numphi) field in a buffer the library
// allocates on first use (137 MB at 258 x 130 x 512); this entry point
// streams the field through `scratch` in column chunks instead and calls no
// allocating routine.
#include "bh_plan_impl.h"
#include "bh_mixed.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

namespace {

// The staging array: 32 MiB of doubles on every device the library runs on,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, never the device memory: kernels and copies get the address
// that hipGetSymbolAddress returns for the current device (scratch_address).
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;
std::vector<double *> g_addr;  // per device ordinal; null until looked up

// scratch's device address on `device` (which must be current), looked up
// once per device.
int scratch_address(int device, double **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_addr.size() <= device) g_addr.resize(device + 1, nullptr);
    if (!g_addr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return bh::fail(BH_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_addr[device] = (double *)ptr;
    }
    *out = g_addr[device];
    return BH_OK;
}

}  // namespace

extern "C" int bh_set_potential_noalloc(bh_plan *p, const double *pot)
{
    using namespace bh;
    if (!p || !pot) return fail(BH_E_INVALID, "null argument");
    int device = 0;
    HIPCHK(hipGetDevice(&device));
    if (device != p->device)
        return fail(BH_E_INVALID, "the plan lives on device " + std::to_string(p->device) + ", device " +
                                      std::to_string(device) + " is current");
    const int N = p->g.numphi;
    const long M = (long)p->g.numr * p->g.numz;
    // whole 64-column panels of all N rows per chunk
    const long cols = (long)(kScratchBytes / sizeof(double)) / N / 64 * 64;
    if (cols < 64)
        return fail(BH_E_INVALID, "numphi = " + std::to_string(N) + ": a 64-column chunk needs " +
                                      std::to_string(64L * N * 8) + " bytes, over the " +
                                      std::to_string(kScratchBytes) + "-byte staging array");
    double *st = nullptr;
    int rc = scratch_address(device, &st);
    if (rc != BH_OK) return rc;
    if ((rc = wait_plan_idle(p)) != BH_OK) return rc;  // a solve still running reads and writes XJ
    p->lazy_state = false;  // XJ is rewritten whole
    p->faces_clean = false;
    if (p->lazy_flag) launch_fillf(p->lazy_flag, 1, 0.f, 0);
    std::vector<Op> ops;
    for (long c0 = 0; c0 < M; c0 += cols) {
        const long w = std::min(cols, M - c0);
        // columns c0 .. c0 + w - 1 of all N planes; the copy on the null
        // stream waits for the previous chunk's transform to have read them
        HIPCHK(hipMemcpy2D(st, w * sizeof(double), pot + c0, M * sizeof(double), w * sizeof(double), N,
                           hipMemcpyHostToDevice));
        ops.clear();
        dft_fwd64_ops(p, st, w, (int)w, p->XJ + c0, M, false, &ops);
        for (const Op &op : ops) launch_op(op, 0);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(0));
    p->has_solution = true;
    return BH_OK;
}
