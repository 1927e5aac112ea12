// Green-Kubo lag correlations on the matrix cores (greenkuboNcomponentionicfluid.cpp:
// 156-249): split-BF16 products on MFMA, FP32 accumulation within a K chunk,
// FP64 across chunks; the epilogue (integrals, Schur complement) is the FP64
// one of gk.hip.
//
// For currents (a, b) and lag l the reference averages J_a(t0).J_b(t0 + l)/3
// over the selected t0.  With the series cut into rows of 32 steps,
//     S_q[r][s] = sum_{p, x} Xa_x(32p + r) * Yb_x(32(p + q) + s)
// is a plain 32 x 32 x K GEMM (K = p), and every (t0, l) product lands in
// exactly one tile: t0 = 32p + r, l = 32q + s - r.  Lag l is the sum of
// diagonal s - r = l - 32q of tiles q = floor(l/32) and floor(l/32) + 1.
// Xa carries the t0 selection of one segment (:207-209); Yb is unmasked.
// Each FP64 value v is split into hi = bf16(v), lo = bf16(v - hi); a product
// is hi*hi' + hi*lo' + lo*hi' (relative error ~2^-16 per product, with random
// sign across t0, so the t0 average reduces it further).  The dropped lo*lo'
// is the one term whose sign is fixed, at lag 0 of a == b (lo^2): the tile
// holding lag 0 adds it as a fourth product.
//
// Layouts in HBM (bf16 bit patterns, u16):
//   X[cx][r][p]          = X_cx(32p + r),          p < PP        (cx = 3c + x)
//   Y[c8][cx][s][p']     = Y_cx(32(p' + c8) + s),  p' < YP, c8 < 8
// so that the 8 k-values one lane feeds to an MFMA are one aligned 16-byte
// load for every tile offset q (the eight c8 copies absorb q mod 8).
// A wavefront owns tiles q0 + 16j (j < GKM_Q) of one current pair and one K
// chunk: tile q0 + 16j at step st reads the same Y vector as tile q0 at step
// st + j, so Y operands live in a rolling register window and each step loads
// one new Y vector and one X vector for GKM_Q tiles.
#include <algorithm>
#include <vector>

#include "anl_internal.h"

namespace anl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t GKM_Q = 8;    // tiles per wavefront, 16 apart
constexpr uint32_t GKM_CS = 16;  // MFMA steps (16 rows of 32 t0 each) per FP32 chunk

__device__ __forceinline__ unsigned short bf16_rn(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

// hi/lo split of one double: hi = RN_bf16((float)v), lo = RN_bf16((float)(v - hi)).
__device__ __forceinline__ void split_bf16(double v, unsigned short &h, unsigned short &l) {
    h = bf16_rn((float)v);
    l = bf16_rn((float)(v - (double)bf16_to_f(h)));
}

// X side, one segment: element (cx, r, p) = data[primo + t][idx[c] + x] at
// t = 32p + r when t is a selected t0 of the segment, else 0.
struct GkIdx {
    uint32_t v[4];
};

__global__ void gkm_split_x_kernel(const double *__restrict__ data, uint32_t ncols, const GkIdx idx, uint64_t primo, uint32_t PP, uint32_t seg_lo, uint32_t seg_hi, uint32_t first,
                                   uint32_t skip, unsigned short *__restrict__ xh, unsigned short *__restrict__ xl) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // e = r * PP + p within one cx
    if (e >= 32ull * PP) return;
    const uint32_t cx = blockIdx.y, c = cx / 3, x = cx % 3;
    const uint32_t r = (uint32_t)(e / PP), p = (uint32_t)(e % PP);
    const uint64_t t = 32ull * p + r;
    double v = 0.0;
    if (t >= seg_lo && t < seg_hi && t >= first && (t - first) % skip == 0) v = data[(primo + t) * ncols + idx.v[c] + x];
    unsigned short h, l;
    split_bf16(v, h, l);
    const uint64_t o = (uint64_t)cx * 32 * PP + e;
    xh[o] = h;
    xl[o] = l;
}

// Y side: element (c8, cx, s, p') = data[primo + t][idx[c] + x] at
// t = 32(p' + c8) + s when t < avail, else 0.
__global__ void gkm_split_y_kernel(const double *__restrict__ data, uint32_t ncols, const GkIdx idx, uint64_t primo, uint64_t avail, uint32_t YP, uint32_t ncx,
                                   unsigned short *__restrict__ yh, unsigned short *__restrict__ yl) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // s * YP + p'
    if (e >= 32ull * YP) return;
    const uint32_t cx = blockIdx.y % ncx, c8 = blockIdx.y / ncx, c = cx / 3, x = cx % 3;
    const uint32_t s = (uint32_t)(e / YP), pp = (uint32_t)(e % YP);
    const uint64_t t = 32ull * ((uint64_t)pp + c8) + s;
    double v = 0.0;
    if (t < avail) v = data[(primo + t) * ncols + idx.v[c] + x];
    unsigned short h, l;
    split_bf16(v, h, l);
    const uint64_t o = ((uint64_t)c8 * ncx + cx) * 32 * YP + e;
    yh[o] = h;
    yl[o] = l;
}

struct GkmArgs {
    const unsigned short *xh, *xl, *yh, *yl;
    uint32_t ncur, PP, YP, ntiles, nchunks;
    double *partial;  // [pair][ntiles][nchunks][63]
};

__device__ __forceinline__ bf16x8 ld8(const unsigned short *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

// Operand layouts of v_mfma_f32_32x32x16_bf16: lane l feeds row l % 32 of A
// and column l % 32 of B, each with the 8 k-values of slot l / 32 (the same k
// slots in A and B); result item e of lane l is row 8(e/4) + 4(l/32) + e%4,
// column l % 32.
__global__ __launch_bounds__(64) void gkm_tile_kernel(const GkmArgs a) {
    __shared__ double tile[32][33];
    const uint32_t g = blockIdx.x, pair = blockIdx.y, chunk = blockIdx.z;
    const uint32_t qbase = (g / 16) * 16 * GKM_Q, q0 = g % 16;
    const uint32_t ca = pair / a.ncur, cb = pair % a.ncur, ncx = 3 * a.ncur;
    const uint32_t lane = threadIdx.x, r = lane & 31u, h = lane >> 5;
    const uint32_t c8 = q0 & 7u, qoff = qbase + (q0 & ~7u);
    const uint32_t st0 = chunk * GKM_CS;
    f32x16 acc[GKM_Q];
#pragma unroll
    for (uint32_t j = 0; j < GKM_Q; j++) acc[j] = f32x16{};
    for (uint32_t x = 0; x < 3; x++) {
        const uint64_t xo = ((uint64_t)(3 * ca + x) * 32 + r) * a.PP + 8 * h;
        const uint64_t yo = (((uint64_t)c8 * ncx + 3 * cb + x) * 32 + r) * a.YP + qoff + 8 * h;
        const unsigned short *XH = a.xh + xo, *XL = a.xl + xo, *YH = a.yh + yo, *YL = a.yl + yo;
        bf16x8 wh[GKM_Q], wl[GKM_Q];
#pragma unroll
        for (uint32_t j = 0; j + 1 < GKM_Q; j++) {
            wh[j] = ld8(YH + 16 * (st0 + j));
            wl[j] = ld8(YL + 16 * (st0 + j));
        }
        // X operands of the next step are loaded while this step's MFMAs run
        bf16x8 ahc = ld8(XH + 16 * st0), alc = ld8(XL + 16 * st0);
        for (uint32_t st = st0; st < st0 + GKM_CS; st += GKM_Q) {
#pragma unroll
            for (uint32_t u = 0; u < GKM_Q; u++) {
                const uint32_t sn = st + u;
                wh[(u + GKM_Q - 1) % GKM_Q] = ld8(YH + 16 * (sn + GKM_Q - 1));
                wl[(u + GKM_Q - 1) % GKM_Q] = ld8(YL + 16 * (sn + GKM_Q - 1));
                const bf16x8 ah = ahc, al = alc;
                if (u + 1 < GKM_Q || st + GKM_Q < st0 + GKM_CS) {  // next step inside this chunk
                    ahc = ld8(XH + 16 * (sn + 1));
                    alc = ld8(XL + 16 * (sn + 1));
                }
#pragma unroll
                for (uint32_t j = 0; j < GKM_Q; j++) {
                    const uint32_t w = (u + j) % GKM_Q;
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, wh[w], acc[j], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, wl[w], acc[j], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, wh[w], acc[j], 0, 0, 0);
                    // tile 0 holds lag 0, where lo*lo' = lo^2 >= 0 for a == b
                    // does not average out over t0: add it there
                    if (j == 0 && g == 0)
                        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, wl[w], acc[0], 0, 0, 0);
                }
            }
        }
    }
    // diagonal sums of each tile, in FP64, into this chunk's partial
#pragma unroll
    for (uint32_t j = 0; j < GKM_Q; j++) {
        const uint32_t q = qbase + q0 + 16 * j;
        if (q >= a.ntiles) break;  // uniform across the wavefront
#pragma unroll
        for (int e = 0; e < 16; e++) tile[8 * (e / 4) + 4 * h + (e % 4)][r] = (double)acc[j][e];
        __syncthreads();
        if (lane < 63) {
            const int d = (int)lane - 31;  // s - r
            double sum = 0.0;
            const int r0 = d < 0 ? -d : 0, r1 = d < 0 ? 32 : 32 - d;
            for (int rr = r0; rr < r1; rr++) sum += tile[rr][rr + d];
            a.partial[(((uint64_t)pair * a.ntiles + q) * a.nchunks + chunk) * 63 + lane] = sum;
        }
        __syncthreads();
    }
}

// Segment mean of product/3 for lag l and pair (diagonal l - 32q of tiles
// q = floor(l/32), floor(l/32) + 1; chunks in order), folded into the running
// mean over segments (:230-237).
__global__ void gkm_reduce_kernel(const double *__restrict__ partial, uint32_t ntiles, uint32_t nchunks,
                                  uint32_t npairs, uint32_t leff, uint32_t narr, double cont, uint32_t iseg,
                                  double *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= leff * npairs) return;
    const uint32_t l = w / npairs, pair = w % npairs;
    double s = 0.0;
    for (uint32_t q = l / 32; q <= l / 32 + 1 && q < ntiles; q++) {
        const int d = (int)l - 32 * (int)q;
        if (d < -31 || d > 31) continue;
        const double *p = partial + ((uint64_t)pair * ntiles + q) * nchunks * 63 + (d + 31);
        for (uint32_t ch = 0; ch < nchunks; ch++) s += p[(uint64_t)ch * 63];
    }
    const double jj = cont > 0 ? (s / 3.0) / cont : 0.0;
    double *v = out + (uint64_t)l * narr + pair;
    *v = *v + (jj - *v) / (double)(iseg + 1);
}

__global__ void gkm_zero_kernel(double *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace anl

using namespace anl;

namespace {

struct GkmPlan {
    uint32_t leff, ncorr, narr, ncx;
    uint64_t nchunks, PP, ntiles, ngroups, YP;
    size_t off_xh, off_xl, off_yh, off_yl, off_partial, bytes;
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

GkmPlan gkm_plan(uint32_t ncur, uint32_t leff, uint32_t n) {
    GkmPlan p{};
    p.leff = leff;
    p.ncorr = ncur * ncur;
    p.narr = 3 * p.ncorr + 2;
    p.ncx = 3 * ncur;
    // t0 axis: rows of 32, MFMA steps of 16 rows, chunks of GKM_CS steps
    const uint64_t nsteps = ((uint64_t)n + 511) / 512;
    p.nchunks = (nsteps + GKM_CS - 1) / GKM_CS;
    p.PP = p.nchunks * GKM_CS * 16;
    // lag axis: tiles of 32 (plus one for the negative diagonals)
    p.ntiles = (leff + 30) / 32 + 1;
    p.ngroups = (p.ntiles + 16 * GKM_Q - 1) / (16 * GKM_Q) * 16;
    const uint64_t qmax = p.ngroups / 16 * 16 * GKM_Q;  // tiles covered, >= ntiles
    // highest Y row read: 16 (st + Q - 1) + 8 + 7 + qoff, st < nchunks * CS, qoff <= qmax - 16 Q + 8
    p.YP = (p.PP + 16 * GKM_Q + qmax + 16 + 7) / 8 * 8;
    if (p.YP >= (1ull << 31) || p.nchunks > 65535)
        fail(ANL_ERR_ARG, "series too long for the MFMA form (use the FP64 form)");
    const size_t xn = (size_t)p.ncx * 32 * p.PP, yn = (size_t)8 * p.ncx * 32 * p.YP;
    p.off_xh = 0;
    p.off_xl = p.off_xh + align256(xn * 2);
    p.off_yh = p.off_xl + align256(xn * 2);
    p.off_yl = p.off_yh + align256(yn * 2);
    p.off_partial = p.off_yl + align256(yn * 2);
    p.bytes = p.off_partial + align256((size_t)p.ncorr * p.ntiles * p.nchunks * 63 * sizeof(double));
    return p;
}

// Enqueues the lag products of one block on s: d [leff][narr] (device) gets
// C(l) in its first ncur^2 columns and zeros elsewhere.
void gkm_enqueue(anl_series *sr, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t n, uint32_t n_seg,
                 uint64_t primo, const GkmPlan &p, char *work, double *d, hipStream_t s) {
    auto *xh = (unsigned short *)(work + p.off_xh);
    auto *xl = (unsigned short *)(work + p.off_xl);
    auto *yh = (unsigned short *)(work + p.off_yh);
    auto *yl = (unsigned short *)(work + p.off_yl);
    auto *partial = (double *)(work + p.off_partial);
    GkIdx idx{};
    for (uint32_t c = 0; c < ncur; c++) idx.v[c] = idx_j[c];
    auto grid1 = [](uint64_t n_, uint32_t y) { return dim3((unsigned)((n_ + 255) / 256), y); };
    const uint64_t outn = (uint64_t)p.leff * p.narr;
    const uint64_t avail = sr->nrows - primo;
    hipLaunchKernelGGL(gkm_zero_kernel, grid1(outn, 1), dim3(256), 0, s, d, outn);
    ANL_HIP(hipGetLastError());
    hipLaunchKernelGGL(gkm_split_y_kernel, grid1(32 * p.YP, 8 * p.ncx), dim3(256), 0, s, sr->data, sr->ncols, idx,
                       primo, avail, (uint32_t)p.YP, p.ncx, yh, yl);
    ANL_HIP(hipGetLastError());
    GkmArgs ka{xh, xl, yh, yl, ncur, (uint32_t)p.PP, (uint32_t)p.YP, (uint32_t)p.ntiles, (uint32_t)p.nchunks, partial};
    const uint32_t seg = n / n_seg;
    for (uint32_t iseg = 0; iseg < n_seg; iseg++) {
        const uint32_t lo = seg * iseg, hi = seg * (iseg + 1);
        const uint32_t first = lo + lo % skip;  // greenkubo...cpp:208
        const uint32_t cont = first < hi ? (hi - first + skip - 1) / skip : 0;
        hipLaunchKernelGGL(gkm_split_x_kernel, grid1(32 * p.PP, p.ncx), dim3(256), 0, s, sr->data, sr->ncols, idx,
                           primo, (uint32_t)p.PP, lo, hi, first, skip, xh, xl);
        ANL_HIP(hipGetLastError());
        hipLaunchKernelGGL(gkm_tile_kernel, dim3((unsigned)p.ngroups, p.ncorr, (unsigned)p.nchunks), dim3(64), 0, s,
                           ka);
        ANL_HIP(hipGetLastError());
        const uint64_t nw = (uint64_t)p.leff * p.ncorr;
        hipLaunchKernelGGL(gkm_reduce_kernel, grid1(nw, 1), dim3(256), 0, s, partial, (uint32_t)p.ntiles,
                           (uint32_t)p.nchunks, p.ncorr, p.leff, p.narr, (double)cont, iseg, d);
        ANL_HIP(hipGetLastError());
    }
}

}  // namespace

namespace anl {

size_t gk_mfma_scratch_bytes(uint32_t ncur, uint32_t leff, uint32_t n) { return gkm_plan(ncur, leff, n).bytes; }

void gk_mfma_enqueue_blocks(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t leff,
                            uint32_t n_seg, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, double *d,
                            hipStream_t st) {
    size_t need = 0;
    for (uint32_t b = 0; b < nblocks; b++) need = std::max(need, gkm_plan(ncur, leff, n[b]).bytes);
    char *work = (char *)series_workspace(s, need);
    const uint64_t stride = (uint64_t)leff * (3 * ncur * ncur + 2);
    // blocks run one after the other on st, sharing the scratch
    for (uint32_t b = 0; b < nblocks; b++)
        gkm_enqueue(s, idx_j, ncur, skip, n[b], n_seg, primo[b], gkm_plan(ncur, leff, n[b]), work, d + b * stride, st);
}

}  // namespace anl
