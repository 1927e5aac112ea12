// Block engine: the per-block loop of BlockAverageG::calcola_custom
// (blockaverage.h:107-190) with MediaVar (calcoliblocchi.h:21-65) for Gofrt
// and MSD, as device work.  Blocks may have different lengths and starts;
// every block's g(r,t) runs in one launch of the pair kernel, every block's
// MSD in one launch of the MSD kernels, and the mean / variance of the mean
// over blocks in one element-parallel kernel that applies the reference's
// Welford steps in block order.  The sequence is captured once into a HIP
// graph (fills are kernels, no memsets or host copies inside) and replayed.
#include <algorithm>
#include <cstring>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

// MediaVar over blocks for element k: x_b = counts_b[k] * incr_b (g(r,t),
// the reference's vdata) or values_b[k] (MSD); calcoliblocchi.h:35-61.
__global__ void mediavar_blocks_kernel(const unsigned long long *__restrict__ counts, const double *__restrict__ values,
                                       const double *__restrict__ incr, uint32_t nblocks, uint64_t len,
                                       double *__restrict__ mean, double *__restrict__ var) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= len) return;
    double m = 0.0, v = 0.0;
    for (uint32_t b = 0; b < nblocks; b++) {
        const double x = counts ? (double)counts[(uint64_t)b * len + k] * incr[b] : values[(uint64_t)b * len + k];
        const double delta = x - m;
        double tmp = delta / (double)(b + 1);
        m = m + tmp;
        tmp = x - m;
        tmp = tmp * delta;
        v = v + tmp;
    }
    mean[k] = m;
    var[k] = v / (double)((nblocks - 1) * nblocks);
}

void mediavar_blocks(const unsigned long long *counts, const double *values, const double *incr, uint32_t nblocks,
                     uint64_t len, double *mean, double *var, hipStream_t s) {
    if (len == 0) return;
    hipLaunchKernelGGL(mediavar_blocks_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, counts, values,
                       incr, nblocks, len, mean, var);
    ANL_HIP(hipGetLastError());
}

}  // namespace anl

using namespace anl;

struct anl_blocks {
    anl_traj *tr = nullptr;
    anl_gofrt *g = nullptr;
    uint32_t nblocks = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // g(r,t)
    void *gprep = nullptr;
    uint32_t g_leff = 0;
    uint64_t g_len = 0;  // per block
    unsigned long long *d_counts = nullptr;
    double *d_incr = nullptr, *d_gmean = nullptr, *d_gvar = nullptr;
    // MSD
    void *mprep = nullptr;
    uint32_t m_leff = 0;
    uint64_t m_len = 0;
    double *d_msd = nullptr, *d_mmean = nullptr, *d_mvar = nullptr;

    void enqueue(hipStream_t s) {
        if (gprep) {
            gofrt_enqueue_prepared(g, gprep, d_counts, s);
            mediavar_blocks(d_counts, nullptr, d_incr, nblocks, g_len, d_gmean, d_gvar, s);
        }
        if (mprep) {
            msd_enqueue_prepared(mprep, d_msd, s);
            mediavar_blocks(nullptr, d_msd, nullptr, nblocks, m_len, d_mmean, d_mvar, s);
        }
    }
    ~anl_blocks() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (gprep) gofrt_free_prepared(gprep);
        if (mprep) msd_free_prepared(mprep);
        for (void *p : {(void *)d_counts, (void *)d_incr, (void *)d_gmean, (void *)d_gvar, (void *)d_msd,
                        (void *)d_mmean, (void *)d_mvar})
            (void)hipFree(p);
    }
};

extern "C" {

int anl_blocks_create(anl_traj *tr, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, anl_gofrt *gofrt,
                      int msd, uint32_t msd_skip, uint32_t msd_tmax, int cm_msd, int cm_self, int use_graph,
                      anl_blocks **out) {
    return guard([&] {
        if (!tr || !primo || !n || !out) fail(ANL_ERR_ARG, "null pointer passed to anl_blocks_create");
        if (nblocks < 2) fail(ANL_ERR_ARG, "the variance over blocks needs at least two blocks");
        if (!gofrt && !msd) fail(ANL_ERR_ARG, "nothing to compute");
        ensure_device();
        auto *B = new anl_blocks();
        try {
            B->tr = tr;
            B->g = gofrt;
            B->nblocks = nblocks;
            ANL_HIP(hipStreamCreate(&B->cap));
            if (gofrt) {
                B->gprep = gofrt_prepare_blocks(gofrt, primo, n, nblocks, &B->g_leff, B->cap);
                uint32_t leff0 = gofrt_block_leff(gofrt, n[0]);
                for (uint32_t b = 0; b < nblocks; b++)
                    if (gofrt_block_leff(gofrt, n[b]) != leff0)
                        fail(ANL_ERR_ARG, "blocks with different numbers of lags cannot be averaged together");
                B->g_len = (uint64_t)B->g_leff * tr->nt * (tr->nt + 1) * gofrt_nbin(gofrt);
                std::vector<double> incr(nblocks);
                const uint32_t skip = gofrt_skip(gofrt);
                for (uint32_t b = 0; b < nblocks; b++)  // gofrt.cpp:91-92
                    incr[b] = (n[b] / skip > 0) ? 1.0 / (double)(int)(n[b] / skip) : 1.0;
                ANL_HIP(hipMalloc(&B->d_counts, nblocks * B->g_len * sizeof(unsigned long long)));
                ANL_HIP(hipMalloc(&B->d_incr, nblocks * sizeof(double)));
                ANL_HIP(hipMemcpy(B->d_incr, incr.data(), nblocks * sizeof(double), hipMemcpyHostToDevice));
                ANL_HIP(hipMalloc(&B->d_gmean, B->g_len * sizeof(double)));
                ANL_HIP(hipMalloc(&B->d_gvar, B->g_len * sizeof(double)));
            }
            if (msd) {
                B->mprep = msd_prepare_blocks(tr, primo, n, nblocks, msd_skip, msd_tmax, cm_msd, cm_self, &B->m_leff);
                B->m_len = (uint64_t)B->m_leff * tr->nt * (cm_msd ? 2 : 1);
                ANL_HIP(hipMalloc(&B->d_msd, std::max<uint64_t>(1, nblocks * B->m_len) * sizeof(double)));
                ANL_HIP(hipMalloc(&B->d_mmean, std::max<uint64_t>(1, B->m_len) * sizeof(double)));
                ANL_HIP(hipMalloc(&B->d_mvar, std::max<uint64_t>(1, B->m_len) * sizeof(double)));
            }
            if (use_graph) {
                ANL_HIP(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
                try {
                    B->enqueue(B->cap);
                } catch (...) {
                    hipGraph_t gtmp = nullptr;
                    (void)hipStreamEndCapture(B->cap, &gtmp);
                    if (gtmp) (void)hipGraphDestroy(gtmp);
                    throw;
                }
                ANL_HIP(hipStreamEndCapture(B->cap, &B->graph));
                ANL_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
            }
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int anl_blocks_run(anl_blocks *b, void *stream) {
    return guard([&] {
        if (!b) fail(ANL_ERR_ARG, "null handle");
        ensure_device();
        if (b->exec) ANL_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
        else b->enqueue((hipStream_t)stream);
    });
}

int anl_blocks_gofrt_results(anl_blocks *b, double *mean, double *var, uint64_t *counts, uint32_t *leff) {
    return guard([&] {
        if (!b || !b->gprep) fail(ANL_ERR_ARG, "no g(r,t) in this block set");
        ANL_HIP(hipDeviceSynchronize());
        if (mean) ANL_HIP(hipMemcpy(mean, b->d_gmean, b->g_len * sizeof(double), hipMemcpyDeviceToHost));
        if (var) ANL_HIP(hipMemcpy(var, b->d_gvar, b->g_len * sizeof(double), hipMemcpyDeviceToHost));
        if (counts)
            ANL_HIP(hipMemcpy(counts, b->d_counts, b->nblocks * b->g_len * sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (leff) *leff = b->g_leff;
    });
}

int anl_blocks_msd_results(anl_blocks *b, double *mean, double *var, double *per_block, uint32_t *leff) {
    return guard([&] {
        if (!b || !b->mprep) fail(ANL_ERR_ARG, "no MSD in this block set");
        ANL_HIP(hipDeviceSynchronize());
        if (mean) ANL_HIP(hipMemcpy(mean, b->d_mmean, b->m_len * sizeof(double), hipMemcpyDeviceToHost));
        if (var) ANL_HIP(hipMemcpy(var, b->d_mvar, b->m_len * sizeof(double), hipMemcpyDeviceToHost));
        if (per_block)
            ANL_HIP(hipMemcpy(per_block, b->d_msd, b->nblocks * b->m_len * sizeof(double), hipMemcpyDeviceToHost));
        if (leff) *leff = b->m_leff;
    });
}

int anl_blocks_destroy(anl_blocks *b) {
    return guard([&] { delete b; });
}

}  // extern "C"
