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
#include <algorithm>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

constexpr int MSD_WG = 256;
constexpr int MSD_TL = 8;  // lags per thread

template <bool CMSELF>
__global__ __launch_bounds__(MSD_WG) void msd_atoms_kernel(const double4 *__restrict__ pos, const double *__restrict__ cm,
                                                          const int32_t *__restrict__ type_slot, uint64_t N, uint32_t nt,
                                                          uint64_t primo, uint32_t n, uint32_t skip, uint32_t leff,
                                                          uint32_t nchunks, double *__restrict__ partial) {
    __shared__ double acc_lds[MSD_TL][MSD_WG];
    __shared__ int type_lds[MSD_WG];
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
        partial[((uint64_t)t * nchunks + blockIdx.x) * nt + tt] = s;
    }
}

__global__ void msd_finish_kernel(const double *__restrict__ partial, uint32_t nchunks, uint32_t nt, uint32_t leff,
                                  const uint64_t *__restrict__ count, uint32_t fcm, double *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= leff * nt) return;
    const uint32_t t = w / nt, tt = w % nt;
    double s = 0.0;
    for (uint32_t c = 0; c < nchunks; c++) s += partial[((uint64_t)t * nchunks + c) * nt + tt];
    out[(uint64_t)t * nt * fcm + tt] = count[tt] ? s / (double)count[tt] : 0.0;
}

// msd.cpp:113-123: |cm(t0) - cm(t0+t)|^2 running mean over t0, per type.
__global__ void msd_cm_kernel(const double *__restrict__ cm, uint32_t nt, uint64_t primo, uint32_t n, uint32_t skip,
                              uint32_t leff, double *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= leff * nt) return;
    const uint32_t t = w / nt, tt = w % nt;
    double m = 0.0;
    uint64_t cont = 0;
    for (uint32_t t0 = 0; t0 < n; t0 += skip) {
        const double *a = cm + ((primo + t0) * nt + tt) * 3;
        const double *b = cm + ((primo + t0 + t) * nt + tt) * 3;
        double e0 = a[0] - b[0], e1 = a[1] - b[1], e2 = a[2] - b[2];
        double delta = ((e0 * e0 + e1 * e1) + e2 * e2) - m;
        m = m + delta / (double)(++cont);
    }
    out[(uint64_t)t * nt * 2 + nt + tt] = m;
}

}  // namespace anl

using namespace anl;

extern "C" {

int anl_msd_calculate_async(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n,
                            uint64_t primo, double *out_dev, void *stream) {
    return guard([&] {
        if (!tr || !out_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_msd_calculate_async");
        if (skip == 0) skip = 1;
        const uint32_t leff = (n < tmax || tmax == 0) ? n : tmax;  // msd.cpp:42
        // msd.cpp:54-60
        if ((uint64_t)leff + n + primo > tr->F)
            fail(ANL_ERR_TOO_SHORT,
                 "trajectory is too short for this kind of calculation. Select a different starting timestep or "
                 "lower the size of the average or the lenght of the time lag");
        if ((uint64_t)leff + n > tr->F)
            fail(ANL_ERR_NOT_LOADED, "there are not enough loaded timesteps inside the trajectory object. I need at "
                                     "least " + std::to_string((uint64_t)leff + n) +
                                         " timesteps to do the requested calculation");
        if ((cm_msd || cm_self) && !tr->cm) fail(ANL_ERR_ARG, "centre-of-mass series not provided");
        if (leff == 0) return;
        ensure_device();
        hipStream_t s = (hipStream_t)stream;
        const uint32_t nt = tr->nt, fcm = cm_msd ? 2 : 1;
        const uint32_t nchunks = (uint32_t)((tr->N + MSD_WG - 1) / MSD_WG);
        const uint32_t nlb = (leff + MSD_TL - 1) / MSD_TL;
        double *partial = nullptr;
        uint64_t *count = nullptr;
        ANL_HIP(hipMallocAsync((void **)&partial, (size_t)leff * nchunks * nt * sizeof(double), s));
        ANL_HIP(hipMallocAsync((void **)&count, nt * sizeof(uint64_t), s));
        const uint64_t n0 = (n + skip - 1) / skip;
        std::vector<uint64_t> hc(nt);
        for (uint32_t k = 0; k < nt; k++) hc[k] = n0 * tr->type_count[k];
        ANL_HIP(hipMemcpyAsync(count, hc.data(), nt * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        dim3 grid(nchunks, nlb);
        if (cm_self)
            hipLaunchKernelGGL(msd_atoms_kernel<true>, grid, dim3(MSD_WG), 0, s, tr->pos, tr->cm, tr->type_slot, tr->N,
                               nt, primo, n, skip, leff, nchunks, partial);
        else
            hipLaunchKernelGGL(msd_atoms_kernel<false>, grid, dim3(MSD_WG), 0, s, tr->pos, tr->cm, tr->type_slot,
                               tr->N, nt, primo, n, skip, leff, nchunks, partial);
        ANL_HIP(hipGetLastError());
        const uint32_t nw = leff * nt;
        hipLaunchKernelGGL(msd_finish_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, partial, nchunks, nt, leff,
                           count, fcm, out_dev);
        ANL_HIP(hipGetLastError());
        if (cm_msd) {
            hipLaunchKernelGGL(msd_cm_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, tr->cm, nt, primo, n, skip, leff,
                               out_dev);
            ANL_HIP(hipGetLastError());
        }
        ANL_HIP(hipFreeAsync(partial, s));
        ANL_HIP(hipFreeAsync(count, s));
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
