// Entry point that allocates nothing: the volume constraint (LinearElasticity.cc:1283-1296) for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops).
//
// tt_volume_constraint sums xPhys with per-block partial sums held in its context's reduction buffer,
// which tt_create allocates.  This variant needs no context: the partials live in `scratch`, a
// file-scope device array of kScratchDoubles doubles (2 KB) that every GPU holds until the process
// ends.  Host code never uses the array's name as an address: at the first call on a device (with that
// device current) hipGetSymbolAddress gives that device's copy, kept per device in g_scratch.
#include <climits>
#include <mutex>

#include "../../include/topopt_gpu.h"
#include "tt_internal.h"

namespace {

// one partial per block of the sum kernel: launch_sum_partials uses at most 256 blocks, whatever n is
constexpr int kScratchDoubles = 256;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// gx = (sum of the nb partials, in block order) / nel - volfrac: the order and the operations of the
// host sum in tt_volume_constraint, so both entry points return the same bits
__global__ void k_volume_finish(const double* __restrict__ partials, int nb, long nel, double volfrac,
                                double* __restrict__ gx) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += partials[b];
        gx[0] = s / (double)nel - volfrac;
    }
}

double* scratch_of_current_device() {
    int dev = -1;
    TT_HIP(hipGetDevice(&dev));
    TT_CHECK(dev >= 0 && dev < kMaxDevices, TT_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        TT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TT_CHECK(p != nullptr, TT_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" int tt_volume_constraint_noalloc(const double* xphys, long nel, double volfrac, double* gx, double* dgdx,
                                            void* stream) {
    try {
        TT_CHECK(xphys && gx, TT_ERR_ARG, "tt_volume_constraint_noalloc: null xphys or gx");
        TT_CHECK(nel >= 1 && nel <= (long)INT_MAX, TT_ERR_ARG,
                 "tt_volume_constraint_noalloc: nel must be in 1 .. 2^31-1 (int32 element indices)");
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* part = scratch_of_current_device();
        const int nb = tt::launch_sum_partials(xphys, nel, part, kScratchDoubles, s);
        hipLaunchKernelGGL(k_volume_finish, dim3(1), dim3(64), 0, s, part, nb, nel, volfrac, gx);
        TT_HIP(hipGetLastError());
        if (dgdx) tt::launch_fill(dgdx, 1.0 / (double)nel, nel, true, s);
    } catch (const tt::Error& e) {
        tt::set_contextless_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        tt::set_contextless_error(e.what());
        return TT_ERR_INTERNAL;
    }
    return TT_OK;
}
