// Mean square displacement (msd.cpp:40-146) on the device.
//
// vdata[t][type] = mean over (t0 = 0, skip, .. < n ; atoms of the type) of
// |x(t0) - x(t0+t) [- (cm(t0) - cm(t0+t))]|^2 and, with cm_msd, the second
// half vdata[t][nt + type] = mean over t0 of |cm(t0) - cm(t0+t)|^2.
//
// The atom part is a sum of non-negative terms: each thread accumulates one
// atom over all t0 for a tile of lags (FP64), the workgroup sums its atoms per
// type in slot order, and a second kernel adds the chunk partials in a fixed
// order and divides by the term count -- deterministic, and within a few ulps
// of the reference's running mean (no cancellation).  The centre-of-mass
// channel replays the reference's running mean exactly.  Several blocks
// (primo_b, n_b) run in one launch.
//
// t0 stride 1 (the usual case) runs msd_window_kernel: a thread holds the R
// frames x(t0 + tb .. t0 + tb + R - 1) of its atom in registers; moving to t0+1
// drops the first and loads one more, so each t0 costs two loads (x(t0) and
// the new window frame) for R updates.  The t0 loop is unrolled R times so the
// window rotates through fixed registers.  R (the lag tile) is tunable.
// Other strides run msd_atoms_kernel (eight lags per thread, every frame
// loaded).
#include <algorithm>
#include <type_traits>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

constexpr int MSD_WG = 256;

// f(std::integral_constant<int, 0>) .. f(std::integral_constant<int, N-1>),
// unrolled at compile time.
template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}
constexpr int MSD_TL = 8;  // lags per thread of msd_atoms_kernel

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

// t0 stride 1.  Workgroup = 64 atom slots (the lanes) x 4 wavefronts, each
// wavefront a tile of R consecutive lags: lags (4 quad + wave) R .. + R - 1.
// The 1-D grid is decoded so that the lag tiles of one atom chunk are
// consecutive logical ids on one XCD (hardware id h runs on XCD h % 8), where
// they share the x(t0) stream in L2.  y(f) = x(f) [- cm(f, type)]: with
// drift removal the centre of mass is subtracted as each frame is loaded.
template <int R, int WV, bool CMSELF>
__global__ __launch_bounds__(64 * WV) void msd_window_kernel(const double4 *__restrict__ pos,
                                                           const double *__restrict__ cm,
                                                           const int32_t *__restrict__ type_slot, uint64_t N,
                                                           uint32_t nt, const MsdBlock *__restrict__ blocks,
                                                           uint32_t leff, uint32_t nchunks, uint32_t nquads,
                                                           uint32_t per_xcd, uint32_t total,
                                                           double *__restrict__ partial) {
    __shared__ double red[WV][R][64];
    __shared__ int type_lds[64];
    const uint32_t h = blockIdx.x;
    const uint32_t L = (h & 7u) * per_xcd + (h >> 3);
    if (L >= total) return;  // whole workgroup: uniform
    const uint32_t quad = L % nquads;
    const uint32_t rest = L / nquads;
    const uint32_t chunk = rest % nchunks;
    const uint32_t blk = rest / nchunks;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tb = (quad * (uint32_t)WV + wave) * (uint32_t)R;
    const uint64_t slot = (uint64_t)chunk * 64u + lane;
    const uint64_t sl = slot < N ? slot : N - 1;  // lanes past N load a real atom and are not summed
    const MsdBlock b = blocks[blk];
    const int ti = type_slot[sl];
    // the last frame a counted (t0, lag) reads; window frames of lags past
    // leff are clamped to it (their sums are never used)
    const uint64_t flast = b.primo + (uint64_t)b.n - 1u + (uint64_t)leff - 1u;
    auto load = [&](uint64_t f) -> double3 {
        f = f < flast ? f : flast;
        const double4 x = pos[f * N + sl];
        double3 y = make_double3(x.x, x.y, x.z);
        if (CMSELF) {
            const double *c = cm + (f * nt + (uint32_t)ti) * 3u;
            y.x = y.x - c[0];
            y.y = y.y - c[1];
            y.z = y.z - c[2];
        }
        return y;
    };
    double acc[R];
#pragma unroll
    for (int k = 0; k < R; k++) acc[k] = 0.0;
    if (tb < leff) {  // uniform per wavefront
        double3 W[R];
#pragma unroll
        for (int k = 0; k < R; k++) W[k] = load(b.primo + tb + (uint32_t)k);
        // step s of a pass starting at t0: W[(s + k) % R] holds frame
        // t0 + s + tb + k; afterwards W[s] is refilled with frame
        // t0 + s + tb + R.  Every index is a compile-time constant once the
        // pass is unrolled, so the window stays in registers.
        // x(t0) is loaded one step ahead (yn), so that its latency hides
        // behind the previous step's R updates
        double3 yn = load(b.primo);
        auto step = [&](uint32_t t0, auto sc) {
            constexpr int s = decltype(sc)::value;
            const double3 ya = yn;
            yn = load(b.primo + t0 + (uint32_t)s + 1u);
#pragma unroll
            for (int k = 0; k < R; k++) {
                const double3 wv = W[(s + k) % R];
                const double e0 = ya.x - wv.x, e1 = ya.y - wv.y, e2 = ya.z - wv.z;
                acc[k] = fma(e0, e0, acc[k]);
                acc[k] = fma(e1, e1, acc[k]);
                acc[k] = fma(e2, e2, acc[k]);
            }
            W[s] = load(b.primo + t0 + (uint32_t)s + tb + (uint32_t)R);
        };
        uint32_t t0 = 0;
        for (; t0 + R <= b.n; t0 += R)
            static_for<R>([&](auto sc) { step(t0, sc); });
        // last partial pass (uniform condition per step)
        static_for<R>([&](auto sc) {
            if (t0 + (uint32_t)decltype(sc)::value < b.n) step(t0, sc);
        });
    }
#pragma unroll
    for (int k = 0; k < R; k++) red[wave][k][lane] = acc[k];
    if (threadIdx.x < 64) type_lds[lane] = slot < N ? ti : -1;
    __syncthreads();
    // thread (wave, lag, type) sums the 64 slots in order
    for (uint32_t w = threadIdx.x; w < (uint32_t)WV * R * nt; w += 64u * WV) {
        const uint32_t wv = w / ((uint32_t)R * nt), r = w % ((uint32_t)R * nt);
        const uint32_t k = r / nt, tt = r % nt;
        const uint32_t t = (quad * (uint32_t)WV + wv) * (uint32_t)R + k;
        if (t >= leff) continue;
        double sum = 0.0;
        for (int j = 0; j < 64; j++)
            if (type_lds[j] == (int)tt) sum += red[wv][k][j];
        partial[(((uint64_t)blk * leff + t) * nchunks + chunk) * nt + tt] = sum;
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

// Launch shapes of msd_window_kernel: (lag tile R, wavefronts per workgroup).
// A workgroup covers R * waves consecutive lags of 64 atoms and streams the
// atoms' frames once for all of them, so wider workgroups read each frame
// fewer times; the register window (R frames) and the workgroup size must
// fit the VGPRs (R 24 needs 256: at most 8 wavefronts; R 16 at most 12).
struct MsdShape {
    uint32_t tile, waves;
};
constexpr MsdShape MSD_SHAPES[] = {{8, 4}, {12, 4}, {16, 4}, {24, 4}, {16, 8}, {24, 8}, {12, 12}, {16, 12}, {8, 16}};
bool msd_shape_ok(uint32_t tile, uint32_t waves) {
    if (tile == 0) return true;  // the eight-lags-per-thread kernel (waves ignored)
    for (const MsdShape &m : MSD_SHAPES)
        if (m.tile == tile && m.waves == waves) return true;
    return false;
}

struct MsdPrepared {
    anl_traj *tr = nullptr;
    uint32_t skip = 1, leff = 0, nblocks = 0, nt = 0, fcm = 1, nchunks = 0;
    uint32_t tile = 0;  // lag tile R of msd_window_kernel; 0: msd_atoms_kernel
    uint32_t waves = 4; // wavefronts per workgroup of msd_window_kernel
    bool cm_msd = false, cm_self = false;
    bool owns = true;   // the three buffers were allocated for this launch
    MsdBlock *d_blocks = nullptr;
    uint64_t *d_count = nullptr;
    double *d_partial = nullptr;
    ~MsdPrepared() {
        if (!owns) return;
        (void)hipFree(d_blocks);
        (void)hipFree(d_count);
        (void)hipFree(d_partial);
    }
};

// msd.cpp:52-61 checks (leff is common to all blocks: min(n_b, tmax) must
// agree); fills P's sizes and the host tables of the block list and the
// term counts.  Touches no device memory.
static void msd_plan(anl_traj *tr, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, uint32_t skip,
                     uint32_t tmax, int cm_msd, int cm_self, MsdPrepared *P, std::vector<MsdBlock> *hb,
                     std::vector<uint64_t> *hc) {
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
    P->tr = tr;
    P->skip = skip;
    P->leff = leff;
    P->nblocks = nblocks;
    P->nt = tr->nt;
    P->fcm = cm_msd ? 2 : 1;
    P->cm_msd = cm_msd;
    P->cm_self = cm_self;
    // the window kernel needs t0 stride 1; its chunks are 64 atoms
    P->tile = skip == 1 ? tr->msd_tile : 0;
    P->waves = tr->msd_waves;
    P->nchunks = (uint32_t)(P->tile ? (tr->N + 63) / 64 : (tr->N + MSD_WG - 1) / MSD_WG);
    hb->assign(nblocks, MsdBlock{});
    hc->assign((size_t)nblocks * tr->nt, 0);
    for (uint32_t b = 0; b < nblocks; b++) {
        (*hb)[b] = MsdBlock{primo[b], n[b], 0};
        const uint64_t n0 = (n[b] + skip - 1) / skip;
        for (uint32_t k = 0; k < tr->nt; k++) (*hc)[(size_t)b * tr->nt + k] = n0 * tr->type_count[k];
    }
}

// partial sums of a planned launch: [nblocks][leff][nchunks][nt] doubles
static size_t msd_partial_len(const MsdPrepared &P) {
    return std::max<size_t>(1, (size_t)P.nblocks * P.leff * P.nchunks * P.nt);
}

void *msd_prepare_blocks(anl_traj *tr, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, uint32_t skip,
                         uint32_t tmax, int cm_msd, int cm_self, uint32_t *leff_out) {
    auto *P = new MsdPrepared();
    try {
        std::vector<MsdBlock> hb;
        std::vector<uint64_t> hc;
        msd_plan(tr, primo, n, nblocks, skip, tmax, cm_msd, cm_self, P, &hb, &hc);
        ensure_device();
        ANL_HIP(hipMalloc(&P->d_blocks, nblocks * sizeof(MsdBlock)));
        ANL_HIP(hipMemcpy(P->d_blocks, hb.data(), nblocks * sizeof(MsdBlock), hipMemcpyHostToDevice));
        ANL_HIP(hipMalloc(&P->d_count, hc.size() * sizeof(uint64_t)));
        ANL_HIP(hipMemcpy(P->d_count, hc.data(), hc.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        ANL_HIP(hipMalloc(&P->d_partial, msd_partial_len(*P) * sizeof(double)));
    } catch (...) {
        delete P;
        throw;
    }
    if (leff_out) *leff_out = P->leff;
    return P;
}

// out: [nblocks][leff][fcm][nt]
void msd_enqueue_prepared(void *prep, double *out, hipStream_t s) {
    const MsdPrepared &P = *(const MsdPrepared *)prep;
    if (P.leff == 0) return;
    const anl_traj *tr = P.tr;
    if (P.tile) {
        const uint32_t span = P.waves * P.tile;  // lags per workgroup
        const uint32_t nquads = (P.leff + span - 1) / span;
        const uint64_t total = (uint64_t)nquads * P.nchunks * P.nblocks;
        const uint64_t per = (total + 7) / 8;
        if (8 * per >= (1ull << 32)) fail(ANL_ERR_ARG, "MSD launch too large (reduce the number of blocks)");
        const dim3 grid((uint32_t)(8 * per));
#define ANL_MSD_WIN(R, W, C)                                                                             \
    hipLaunchKernelGGL((msd_window_kernel<R, W, C>), grid, dim3(64 * W), 0, s, tr->pos, tr->cm, tr->type_slot, \
                       tr->N, P.nt, P.d_blocks, P.leff, P.nchunks, nquads, (uint32_t)per, (uint32_t)total,       \
                       P.d_partial)
#define ANL_MSD_CASE(R, W)                          \
    case (R) * 64 + (W) * 2: ANL_MSD_WIN(R, W, false); break; \
    case (R) * 64 + (W) * 2 + 1: ANL_MSD_WIN(R, W, true); break;
        switch (P.tile * 64 + P.waves * 2 + (P.cm_self ? 1 : 0)) {
            ANL_MSD_CASE(8, 4)
            ANL_MSD_CASE(12, 4)
            ANL_MSD_CASE(16, 4)
            ANL_MSD_CASE(24, 4)
            ANL_MSD_CASE(16, 8)
            ANL_MSD_CASE(24, 8)
            ANL_MSD_CASE(12, 12)
            ANL_MSD_CASE(16, 12)
            ANL_MSD_CASE(8, 16)
            default: fail(ANL_ERR_ARG, "unsupported MSD launch shape");
        }
#undef ANL_MSD_CASE
#undef ANL_MSD_WIN
    } else {
        const uint32_t nlb = (P.leff + MSD_TL - 1) / MSD_TL;
        dim3 grid(P.nchunks, nlb, P.nblocks);
        if (P.cm_self)
            hipLaunchKernelGGL(msd_atoms_kernel<true>, grid, dim3(MSD_WG), 0, s, tr->pos, tr->cm, tr->type_slot, tr->N,
                               P.nt, P.d_blocks, P.skip, P.leff, P.nchunks, P.d_partial);
        else
            hipLaunchKernelGGL(msd_atoms_kernel<false>, grid, dim3(MSD_WG), 0, s, tr->pos, tr->cm, tr->type_slot,
                               tr->N, P.nt, P.d_blocks, P.skip, P.leff, P.nchunks, P.d_partial);
    }
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

// One MSD calculation whose block table, term counts and partial sums live
// in caller-provided device memory (scratch, scratch_bytes, 256-byte
// aligned); allocates nothing.  The tables are copied in on stream s, which
// is synchronised before returning (the host tables are locals).
void msd_calculate_in_scratch(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n,
                              uint64_t primo, double *out, unsigned char *scratch, size_t scratch_bytes,
                              hipStream_t s) {
    MsdPrepared P;
    P.owns = false;
    std::vector<MsdBlock> hb;
    std::vector<uint64_t> hc;
    msd_plan(tr, &primo, &n, 1, skip, tmax, cm_msd, cm_self, &P, &hb, &hc);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_blocks = al(sizeof(MsdBlock)), b_count = al(hc.size() * sizeof(uint64_t));
    const size_t b_partial = al(msd_partial_len(P) * sizeof(double));
    if (b_blocks + b_count + b_partial > scratch_bytes)
        fail(ANL_ERR_ARG, "MSD of " + std::to_string(tr->N) + " atoms and " + std::to_string(P.leff) + " lags needs " +
                              std::to_string(b_blocks + b_count + b_partial) + " bytes of scratch, more than the " +
                              std::to_string(scratch_bytes) + " of the allocation-free entry point");
    P.d_blocks = reinterpret_cast<MsdBlock *>(scratch);
    P.d_count = reinterpret_cast<uint64_t *>(scratch + b_blocks);
    P.d_partial = reinterpret_cast<double *>(scratch + b_blocks + b_count);
    ANL_HIP(hipMemcpyAsync(P.d_blocks, hb.data(), sizeof(MsdBlock), hipMemcpyHostToDevice, s));
    ANL_HIP(hipMemcpyAsync(P.d_count, hc.data(), hc.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    msd_enqueue_prepared(&P, out, s);
    ANL_HIP(hipStreamSynchronize(s));
}

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

static const char *MSD_SHAPE_MSG =
    "MSD launch shape must be lag tile 0, or (lag tile, wavefronts) one of (8,4) (12,4) (16,4) (24,4) (16,8) "
    "(24,8) (12,12) (16,12) (8,16)";

int anl_msd_set_tuning(anl_traj *tr, uint32_t lag_tile, uint32_t waves) {
    return guard([&] {
        if (!tr) fail(ANL_ERR_ARG, "null trajectory");
        if (!msd_shape_ok(lag_tile, waves)) fail(ANL_ERR_ARG, MSD_SHAPE_MSG);
        tr->msd_tile = lag_tile;
        tr->msd_waves = lag_tile ? waves : 4;
    });
}

int anl_msd_get_tuning(anl_traj *tr, uint32_t *lag_tile, uint32_t *waves) {
    return guard([&] {
        if (!tr || !lag_tile) fail(ANL_ERR_ARG, "null pointer passed to anl_msd_get_tuning");
        *lag_tile = tr->msd_tile;
        if (waves) *waves = tr->msd_waves;
    });
}

// Times the MSD calculation (reset(n) at primo) with each candidate lag tile:
// one untimed and one timed call each (HIP events on the null stream); keeps
// the fastest in the trajectory.
int anl_msd_autotune(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n, uint64_t primo,
                     const uint32_t *cand, uint32_t ncand, float *ms_out, uint32_t *best_out) {
    return guard([&] {
        if (!tr) fail(ANL_ERR_ARG, "null trajectory");
        // (lag tile, wavefronts) pairs
        static const uint32_t defaults[] = {8, 4, 12, 4, 16, 4, 24, 4, 16, 8, 24, 8, 12, 12, 16, 12, 8, 16, 0, 4};
        if (!cand) {
            cand = defaults;
            ncand = sizeof(defaults) / sizeof(defaults[0]) / 2;
        }
        if (ncand == 0) fail(ANL_ERR_ARG, "no candidates");
        for (uint32_t c = 0; c < ncand; c++)
            if (!msd_shape_ok(cand[2 * c], cand[2 * c + 1]))
                fail(ANL_ERR_ARG, "candidate " + std::to_string(c) + ": " + MSD_SHAPE_MSG);
        ensure_device();
        const uint32_t saved = tr->msd_tile, saved_w = tr->msd_waves;
        const uint32_t leff = (n < tmax || tmax == 0) ? n : tmax;
        double *d = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        void *P = nullptr;
        float best = -1.0f;
        uint32_t besti = 0;
        try {
            ANL_HIP(hipMalloc(&d, std::max<size_t>(1, (size_t)leff * tr->nt * 2) * sizeof(double)));
            ANL_HIP(hipEventCreate(&e0));
            ANL_HIP(hipEventCreate(&e1));
            for (uint32_t c = 0; c < ncand; c++) {
                tr->msd_tile = cand[2 * c];
                tr->msd_waves = cand[2 * c] ? cand[2 * c + 1] : 4;
                P = msd_prepare_blocks(tr, &primo, &n, 1, skip, tmax, cm_msd, cm_self, nullptr);
                msd_enqueue_prepared(P, d, 0);
                ANL_HIP(hipEventRecord(e0, 0));
                msd_enqueue_prepared(P, d, 0);
                ANL_HIP(hipEventRecord(e1, 0));
                ANL_HIP(hipEventSynchronize(e1));
                msd_free_prepared(P);
                P = nullptr;
                float ms = 0.0f;
                ANL_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms_out) ms_out[c] = ms;
                if (best < 0.0f || ms < best) {
                    best = ms;
                    besti = c;
                }
            }
        } catch (...) {
            tr->msd_tile = saved;
            tr->msd_waves = saved_w;
            if (P) {
                (void)hipDeviceSynchronize();
                msd_free_prepared(P);
            }
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipFree(d);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(d);
        tr->msd_tile = cand[2 * besti];
        tr->msd_waves = cand[2 * besti] ? cand[2 * besti + 1] : 4;
        if (best_out) *best_out = besti;
    });
}

}  // extern "C"
