// Allocation-free MSD entry point.  Callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc) get the MSD of anl_msd_calculate_async with its
// scratch -- the block table, the per-type term counts and the partial sums
// of the atom chunks -- in a file-scope device array of fixed size instead of
// freshly allocated buffers.  Each GPU the library runs on holds its own copy
// of the array until the process ends; every call on a GPU reuses it, so
// calls that overlap on one GPU race.  Nothing needs clearing between calls:
// every entry a call reads is written by that call first.
#include <mutex>

#include "anl_internal.h"

namespace anl {
constexpr size_t NOALLOC_SCRATCH_BYTES = 16u << 20;  // 16 MiB per GPU
}

static __device__ __align__(256) unsigned char scratch[anl::NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[MAX_DEVICES] = {};

// Device address of the current GPU's copy of scratch, looked up once per
// device with that device current.  In host code the array's name is only a
// host-side stand-in; kernels get this address.
unsigned char *scratch_of_current_device() {
    int dev = -1;
    ANL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) anl::fail(ANL_ERR_ARG, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        ANL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) anl::fail(ANL_ERR_HIP, "no device address for the MSD scratch");
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" int anl_msd_calculate_noalloc(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self,
                                         uint32_t n, uint64_t primo, double *out_dev, void *stream) {
    using namespace anl;
    return guard([&] {
        if (!tr || !out_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_msd_calculate_noalloc");
        ensure_device();
        unsigned char *p = scratch_of_current_device();
        msd_calculate_in_scratch(tr, skip, tmax, cm_msd, cm_self, n, primo, out_dev, p, NOALLOC_SCRATCH_BYTES,
                                 (hipStream_t)stream);
    });
}
