// Mean square displacement (msd.cpp:40-146) on the device.
//
// vdata[t][type] = mean over (t0 = 0, skip, .. < n ; atoms of the type) of
// |x(t0) - x(t0+t) [- (cm(t0) - cm(t0+t))]|^2 and, with cm_msd, the second
// half vdata[t][nt + type] = mean over t0 of |cm(t0) - cm(t0+t)|^2.
//
// The atom part is a sum of non-negative terms: each thread accumulates one
// atom over all t0 for a block of lags (FP64), a workgroup reduces its 256
// atoms per type in slot order, and a second kernel adds the workgroup
// partials in a fixed order and divides by the term count -- deterministic,
// and within a few ulps of the reference's running mean (no cancellation).
// The centre-of-mass channel replays the reference's running mean exactly.
// Several blocks (primo_b, n_b) run in one launch (grid z = block).
#include <algorithm>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

constexpr int MSD_WG = 256;
constexpr int MSD_TL = 8;  // lags per thread

struct MsdBlock {
    uint64_t primo;
    uint32_t n;
    uint32_t pad;
};

template <bool CMSELF>
__global__ __launch_bounds__(MSD_WG) void msd_atoms_kernel(const double4 *__restrict__ pos, const double *__restrict__ cm,
                                                          const int32_t *__restrict__ type_slot, uint64_t N, uint32_t nt,
                                                          const MsdBlock *__restrict__ blocks, uint32_t skip,
                                                          uint32_t leff, uint32_t nchunks, double *__restrict__ partial) {
    __shared__ double acc_lds[MSD_TL][MSD_WG];
    __shared__ int type_lds[MSD_WG];
    const MsdBlock blk = blocks[blockIdx.z];
    const uint64_t primo = blk.primo;
    const uint32_t n = blk.n;
    const uint64_t slot = (uint64_t)blockIdx.x * MSD_WG + threadIdx.x;
    const uint32_t tb = blockIdx.y * MSD_TL;
    double acc[MSD_TL];
#pragma unroll
    for (int l = 0; l < MSD_TL; l++) acc[l] = 0.0;
    int ti = -1;
    if (slot < N) {
        ti = type_slot[slot];
        for (uint32_t t0 = 0; t0 < n; t0 += skip) {
            const uint64_t fa = primo + t0;
            const double4 xa = pos[fa * N + slot];
            double ca0 = 0, ca1 = 0, ca2 = 0;
            if (CMSELF) {
                const double *c = cm + (fa * nt + ti) * 3;
                ca0 = c[0];
                ca1 = c[1];
                ca2 = c[2];
            }
#pragma unroll
            for (int l = 0; l < MSD_TL; l++) {
                const uint32_t t = tb + l;
                if (t < leff) {
                    const uint64_t fb = fa + t;
                    const double4 xb = pos[fb * N + slot];
                    double e0 = xa.x - xb.x, e1 = xa.y - xb.y, e2 = xa.z - xb.z;
                    if (CMSELF) {
                        const double *c = cm + (fb * nt + ti) * 3;
                        e0 = e0 - (ca0 - c[0]);
                        e1 = e1 - (ca1 - c[1]);
                        e2 = e2 - (ca2 - c[2]);
                    }
                    acc[l] += (e0 * e0 + e1 * e1) + e2 * e2;
                }
            }
        }
    }
#pragma unroll
    for (int l = 0; l < MSD_TL; l++) acc_lds[l][threadIdx.x] = acc[l];
    type_lds[threadIdx.x] = ti;
    __syncthreads();
    // thread (l, type) sums its lane set in slot order
    for (uint32_t w = threadIdx.x; w < MSD_TL * nt; w += MSD_WG) {
        const uint32_t l = w / nt, tt = w % nt;
        const uint32_t t = tb + l;
        if (t >= leff) continue;
        double s = 0.0;
        for (int k = 0; k < MSD_WG; k++)
            if (type_lds[k] == (int)tt) s += acc_lds[l][k];
        partial[(((uint64_t)blockIdx.z * leff + t) * nchunks + blockIdx.x) * nt + tt] = s;
    }
}

__global__ void msd_finish_kernel(const double *__restrict__ partial, uint32_t nchunks, uint32_t nt, uint32_t leff,
                                  uint32_t nblocks, const uint64_t *__restrict__ count, uint32_t fcm,
                                  double *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nblocks * leff * nt) return;
    const uint32_t b = w / (leff * nt), r = w % (leff * nt);
    const uint32_t t = r / nt, tt = r % nt;
    double s = 0.0;
    for (uint32_t c = 0; c < nchunks; c++) s += partial[(((uint64_t)b * leff + t) * nchunks + c) * nt + tt];
    const uint64_t cn = count[(uint64_t)b * nt + tt];
    out[(uint64_t)b * leff * nt * fcm + (uint64_t)t * nt * fcm + tt] = cn ? s / (double)cn : 0.0;
}

// msd.cpp:113-123: |cm(t0) - cm(t0+t)|^2 running mean over t0, per type.
__global__ void msd_cm_kernel(const double *__restrict__ cm, uint32_t nt, const MsdBlock *__restrict__ blocks,
                              uint32_t skip, uint32_t leff, uint32_t nblocks, double *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nblocks * leff * nt) return;
    const uint32_t b = w / (leff * nt), r = w % (leff * nt);
    const uint32_t t = r / nt, tt = r % nt;
    const MsdBlock blk = blocks[b];
    double m = 0.0;
    uint64_t cont = 0;
    for (uint32_t t0 = 0; t0 < blk.n; t0 += skip) {
        const double *a = cm + ((blk.primo + t0) * nt + tt) * 3;
        const double *c = cm + ((blk.primo + t0 + t) * nt + tt) * 3;
        double e0 = a[0] - c[0], e1 = a[1] - c[1], e2 = a[2] - c[2];
        double delta = ((e0 * e0 + e1 * e1) + e2 * e2) - m;
        m = m + delta / (double)(++cont);
    }
    out[(uint64_t)b * leff * nt * 2 + (uint64_t)t * nt * 2 + nt + tt] = m;
}

struct MsdPrepared {
    anl_traj *tr = nullptr;
    uint32_t skip = 1, leff = 0, nblocks = 0, nt = 0, fcm = 1, nchunks = 0;
    bool cm_msd = false, cm_self = false;
    MsdBlock *d_blocks = nullptr;
    uint64_t *d_count = nullptr;
    double *d_partial = nullptr;
    ~MsdPrepared() {
        (void)hipFree(d_blocks);
        (void)hipFree(d_count);
        (void)hipFree(d_partial);
    }
};

// msd.cpp:52-61 checks (leff is common to all blocks: min(n_b, tmax) must agree)
void *msd_prepare_blocks(anl_traj *tr, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, uint32_t skip,
                         uint32_t tmax, int cm_msd, int cm_self, uint32_t *leff_out) {
    if (nblocks == 0) fail(ANL_ERR_ARG, "no blocks");
    if (skip == 0) skip = 1;
    const uint32_t leff = (n[0] < tmax || tmax == 0) ? n[0] : tmax;  // msd.cpp:42
    for (uint32_t b = 0; b < nblocks; b++) {
        const uint32_t lb = (n[b] < tmax || tmax == 0) ? n[b] : tmax;
        if (lb != leff) fail(ANL_ERR_ARG, "blocks with different numbers of lags cannot be averaged together");
        if ((uint64_t)leff + n[b] + primo[b] > tr->F)
            fail(ANL_ERR_TOO_SHORT,
                 "trajectory is too short for this kind of calculation. Select a different starting timestep or "
                 "lower the size of the average or the lenght of the time lag");
        if ((uint64_t)leff + n[b] > tr->F)
            fail(ANL_ERR_NOT_LOADED, "there are not enough loaded timesteps inside the trajectory object. I need at "
                                     "least " + std::to_string((uint64_t)leff + n[b]) +
                                         " timesteps to do the requested calculation");
    }
    if ((cm_msd || cm_self) && !tr->cm) fail(ANL_ERR_ARG, "centre-of-mass series not provided");
    ensure_device();
    auto *P = new MsdPrepared();
    try {
        P->tr = tr;
        P->skip = skip;
        P->leff = leff;
        P->nblocks = nblocks;
        P->nt = tr->nt;
        P->fcm = cm_msd ? 2 : 1;
        P->cm_msd = cm_msd;
        P->cm_self = cm_self;
        P->nchunks = (uint32_t)((tr->N + MSD_WG - 1) / MSD_WG);
        std::vector<MsdBlock> hb(nblocks);
        std::vector<uint64_t> hc((size_t)nblocks * tr->nt);
        for (uint32_t b = 0; b < nblocks; b++) {
            hb[b] = MsdBlock{primo[b], n[b], 0};
            const uint64_t n0 = (n[b] + skip - 1) / skip;
            for (uint32_t k = 0; k < tr->nt; k++) hc[(size_t)b * tr->nt + k] = n0 * tr->type_count[k];
        }
        ANL_HIP(hipMalloc(&P->d_blocks, nblocks * sizeof(MsdBlock)));
        ANL_HIP(hipMemcpy(P->d_blocks, hb.data(), nblocks * sizeof(MsdBlock), hipMemcpyHostToDevice));
        ANL_HIP(hipMalloc(&P->d_count, hc.size() * sizeof(uint64_t)));
        ANL_HIP(hipMemcpy(P->d_count, hc.data(), hc.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        ANL_HIP(hipMalloc(&P->d_partial, std::max<size_t>(1, (size_t)nblocks * leff * P->nchunks * P->nt) * sizeof(double)));
    } catch (...) {
        delete P;
        throw;
    }
    if (leff_out) *leff_out = leff;
    return P;
}

// out: [nblocks][leff][fcm][nt]
void msd_enqueue_prepared(void *prep, double *out, hipStream_t s) {
    const MsdPrepared &P = *(const MsdPrepared *)prep;
    if (P.leff == 0) return;
    const anl_traj *tr = P.tr;
    const uint32_t nlb = (P.leff + MSD_TL - 1) / MSD_TL;
    dim3 grid(P.nchunks, nlb, P.nblocks);
    if (P.cm_self)
        hipLaunchKernelGGL(msd_atoms_kernel<true>, grid, dim3(MSD_WG), 0, s, tr->pos, tr->cm, tr->type_slot, tr->N,
                           P.nt, P.d_blocks, P.skip, P.leff, P.nchunks, P.d_partial);
    else
        hipLaunchKernelGGL(msd_atoms_kernel<false>, grid, dim3(MSD_WG), 0, s, tr->pos, tr->cm, tr->type_slot, tr->N,
                           P.nt, P.d_blocks, P.skip, P.leff, P.nchunks, P.d_partial);
    ANL_HIP(hipGetLastError());
    const uint32_t nw = P.nblocks * P.leff * P.nt;
    hipLaunchKernelGGL(msd_finish_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, P.d_partial, P.nchunks, P.nt,
                       P.leff, P.nblocks, P.d_count, P.fcm, out);
    ANL_HIP(hipGetLastError());
    if (P.cm_msd) {
        hipLaunchKernelGGL(msd_cm_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, tr->cm, P.nt, P.d_blocks, P.skip,
                           P.leff, P.nblocks, out);
        ANL_HIP(hipGetLastError());
    }
}

void msd_free_prepared(void *prep) { delete (MsdPrepared *)prep; }

}  // namespace anl

using namespace anl;

extern "C" {

int anl_msd_calculate_async(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n,
                            uint64_t primo, double *out_dev, void *stream) {
    return guard([&] {
        if (!tr || !out_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_msd_calculate_async");
        void *P = msd_prepare_blocks(tr, &primo, &n, 1, skip, tmax, cm_msd, cm_self, nullptr);
        try {
            msd_enqueue_prepared(P, out_dev, (hipStream_t)stream);
            ANL_HIP(hipStreamSynchronize((hipStream_t)stream));  // the prepared buffers are freed below
        } catch (...) {
            msd_free_prepared(P);
            throw;
        }
        msd_free_prepared(P);
    });
}

int anl_msd_calculate(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n, uint64_t primo,
                      double *out_host) {
    return guard([&] {
        if (!tr || !out_host) fail(ANL_ERR_ARG, "null pointer passed to anl_msd_calculate");
        const uint32_t leff = (n < tmax || tmax == 0) ? n : tmax;
        const size_t len = (size_t)leff * tr->nt * (cm_msd ? 2 : 1);
        ensure_device();
        double *d = nullptr;
        ANL_HIP(hipMalloc(&d, std::max<size_t>(len, 1) * sizeof(double)));
        int rc = anl_msd_calculate_async(tr, skip, tmax, cm_msd, cm_self, n, primo, d, nullptr);
        if (rc != ANL_OK) {
            std::string m = anl_last_error();
            (void)hipFree(d);
            fail(rc, m);
        }
        hipError_t e = hipMemcpy(out_host, d, len * sizeof(double), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        ANL_HIP(e);
    });
}

}  // extern "C"
