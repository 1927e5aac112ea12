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
// Layouts in HBM (bf16 bit patterns, u16), in k-blocks of 8:
//   X[cx][p/8][r][p%8]        = X_cx(32p + r),          p < PP      (cx = 3c + x)
//   Y[c8][cx][p'/8][s][p'%8]  = Y_cx(32(p' + c8) + s),  p' < YP, c8 < 8
// so that the 8 k-values one lane feeds to an MFMA are one aligned 16-byte
// load for every tile offset q (the eight c8 copies absorb q mod 8), and the
// 32 rows of one k-block are 512 contiguous bytes: a wavefront's operand load
// touches 8 whole cache lines.
// A wavefront owns tiles q0 + 16j (j < Q) of one current pair and one K
// chunk of CS steps: tile q0 + 16j at step st reads the same Y vector as tile
// q0 at step st + j, so Y operands live in a rolling register window and each
// step loads one new Y vector and one X vector for Q tiles.  WG wavefronts of
// consecutive tile groups form a workgroup and share the X operands.  The
// shape (Q, CS, WG) is a launch parameter of the series, picked by
// anl_gk_autotune: Q and WG never change results (every tile sums its K
// steps in the same order); CS sets where the FP32 sums are handed to FP64,
// so it changes the rounding within the MFMA form's error bound.
#include <algorithm>
#include <vector>

#include "anl_internal.h"

namespace anl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t GKM_DY = 2;     // Y operands loaded this many steps before first use
constexpr uint32_t GKM_DX = 1;     // X operands likewise
constexpr uint32_t GKM_MAXB = 64;  // blocks per batch of launches
constexpr uint32_t GKM_SP = 32;    // rows p of one split workgroup

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

__host__ __device__ inline uint64_t align256(uint64_t b) { return (b + 255) / 256 * 256; }

// Byte offsets of the four operand arrays inside one block's region.
struct GkmOff {
    uint64_t xh, xl, yh, yl, end;
};
__host__ __device__ inline GkmOff gkm_off(uint32_t ncx, uint32_t PP, uint32_t YP) {
    const uint64_t xn = (uint64_t)ncx * 32 * PP, yn = 8ull * ncx * 32 * YP;
    GkmOff o;
    o.xh = 0;
    o.xl = align256(2 * xn);
    o.yh = o.xl + align256(2 * xn);
    o.yl = o.yh + align256(2 * yn);
    o.end = o.yl + align256(2 * yn);
    return o;
}

// One batch of blocks: block b reads series rows primo[b] ..., averages over
// n[b] t0, owns the operand region at work + base[b] (PP[b] t0 rows of 32 on
// the X side, YP[b] on the Y side), the partial sums at partial + poff[b] and
// MFMA chunks z0[b] .. z0[b+1] of the tile grid.
struct GkmBatch {
    const double *data;
    char *work;
    double *partial;
    double *out;  // block b at out + b * stride, [leff][narr]
    uint64_t stride;
    uint32_t ncols, ncur, ncx, ntiles, leff, narr, skip, n_seg, nblk;
    uint32_t idx[4];
    uint64_t primo[GKM_MAXB], base[GKM_MAXB], poff[GKM_MAXB], nrows;
    uint32_t n[GKM_MAXB], PP[GKM_MAXB], YP[GKM_MAXB], z0[GKM_MAXB + 1];
};

// Operand split of one block (blockIdx.z) and one group of GKM_SP rows p
// (blockIdx.x): O[c8][cx][s][p] = V(32(p + c8) + s), c8 < nshift, where V is
// the series value of column cx on the t0 side of segment iseg (X: zero
// outside the selected t0 of the segment) or the lagged side (Y: zero past
// the end of the series).  Rows of the series are read coalesced into LDS
// (already split), then written out as 16-byte runs of 8 consecutive p.
template <bool YSIDE>
__global__ __launch_bounds__(256) void gkm_split_kernel(const GkmBatch a, uint32_t iseg) {
    constexpr uint32_t NSH = YSIDE ? 8 : 1, NU = GKM_SP + NSH - 1, UP = NU + 1;
    extern __shared__ unsigned short sh[];  // hi [cx][s][UP], then lo
    const uint32_t b = blockIdx.z, P0 = blockIdx.x * GKM_SP, t = threadIdx.x;
    const uint32_t P = YSIDE ? a.YP[b] : a.PP[b];
    if (P0 >= P) return;  // uniform across the workgroup
    const uint32_t ncx = a.ncx;
    unsigned short *shl = sh + ncx * 32 * UP;
    const uint64_t primo = a.primo[b];
    uint64_t lo = 0, hi = 0, first = 0;
    if (!YSIDE) {
        const uint32_t seg = a.n[b] / a.n_seg;
        lo = (uint64_t)seg * iseg;
        hi = (uint64_t)seg * (iseg + 1);
        first = lo + lo % a.skip;  // greenkubo...cpp:208
    }
    const uint64_t avail = a.nrows - primo;
    for (uint32_t e = t; e < NU * 32 * ncx; e += blockDim.x) {
        const uint32_t tl = e / ncx, cx = e - tl * ncx;
        const uint64_t tt = 32ull * P0 + tl;
        bool use;
        if (YSIDE) use = tt < avail;
        else use = tt >= first && tt < hi && (tt - first) % a.skip == 0;
        double v = 0.0;
        if (use) v = a.data[(primo + tt) * a.ncols + a.idx[cx / 3] + cx % 3];
        unsigned short h, l;
        split_bf16(v, h, l);
        const uint32_t o = (cx * 32 + tl % 32) * UP + tl / 32;
        sh[o] = h;
        shl[o] = l;
    }
    __syncthreads();
    const GkmOff off = gkm_off(ncx, a.PP[b], a.YP[b]);
    unsigned short *dh = (unsigned short *)(a.work + a.base[b] + (YSIDE ? off.yh : off.xh));
    unsigned short *dl = (unsigned short *)(a.work + a.base[b] + (YSIDE ? off.yl : off.xl));
    const uint32_t nrun = NSH * ncx * 32 * (GKM_SP / 8);
    for (uint32_t e = t; e < nrun; e += blockDim.x) {
        // consecutive threads: consecutive rows s of one k-block (contiguous)
        const uint32_t sr = e % 32, k = (e / 32) % (GKM_SP / 8), rb = e / (32 * (GKM_SP / 8));  // rb = c8 * ncx + cx
        const uint32_t c8 = rb / ncx, cs = (rb - c8 * ncx) * 32 + sr;
        const uint32_t p = P0 + 8 * k;
        if (p >= P) continue;  // P is a multiple of 8
        const unsigned short *sa = sh + cs * UP + 8 * k + c8, *sb = shl + cs * UP + 8 * k + c8;
        uint4 wh, wl;
        wh.x = sa[0] | (uint32_t)sa[1] << 16;
        wh.y = sa[2] | (uint32_t)sa[3] << 16;
        wh.z = sa[4] | (uint32_t)sa[5] << 16;
        wh.w = sa[6] | (uint32_t)sa[7] << 16;
        wl.x = sb[0] | (uint32_t)sb[1] << 16;
        wl.y = sb[2] | (uint32_t)sb[3] << 16;
        wl.z = sb[4] | (uint32_t)sb[5] << 16;
        wl.w = sb[6] | (uint32_t)sb[7] << 16;
        const uint64_t o = (((uint64_t)rb * (P / 8) + p / 8) * 32 + sr) * 8;
        *reinterpret_cast<uint4 *>(dh + o) = wh;
        *reinterpret_cast<uint4 *>(dl + o) = wl;
    }
}

__device__ __forceinline__ bf16x8 ld8(const unsigned short *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

// The K loop of one wavefront: 3 components x CS steps of Q tiles,
// three split products each (four on tile 0 when LAG0).  The workgroup's
// wavefronts share one X operand block per component: it is copied into LDS
// (xs: [hi/lo][step][lane] 16-byte vectors, contiguous in HBM) while the Y
// window preloads, and read from LDS one step ahead.
template <uint32_t Q, uint32_t CS, uint32_t WG, bool LAG0>
__device__ __forceinline__ void gkm_accumulate(f32x16 (&acc)[Q], uint4 *xs, const unsigned short *__restrict__ xh,
                                               const unsigned short *__restrict__ xl,
                                               const unsigned short *__restrict__ yh,
                                               const unsigned short *__restrict__ yl, uint32_t PP, uint32_t YP,
                                               uint32_t ca, uint32_t cb, uint32_t ncx, uint32_t r, uint32_t h,
                                               uint32_t c8, uint32_t qoff, uint32_t st0) {
    const uint32_t lane = h * 32 + r, tid = threadIdx.x;
    constexpr uint32_t XN = CS * 64;  // vectors of one X block (hi or lo)
    for (uint32_t x = 0; x < 3; x++) {
        // k-block of step st: 2 st + h (X), 2 st + qoff / 8 + h (Y); 512 elements per step
        const uint64_t xb = ((uint64_t)(3 * ca + x) * (PP / 8) + 2 * st0) * 32 * 8;
        const uint64_t yo = ((((uint64_t)c8 * ncx + 3 * cb + x) * (YP / 8) + qoff / 8 + h) * 32 + r) * 8;
        const unsigned short *YH = yh + yo, *YL = yl + yo;
        // Y(v) (the operand vector of step v) lives in window slot v % YW and
        // is loaded GKM_DY steps before its first use (by tile Q - 1);
        // X(v) in register slot v % XW, read from LDS GKM_DX steps ahead
        constexpr uint32_t YW = Q + GKM_DY, XW = GKM_DX + 1;
        bf16x8 wh[YW], wl[YW], xhv[XW], xlv[XW];
#pragma unroll
        for (uint32_t v = 0; v + 1 < YW; v++) {
            wh[v] = ld8(YH + 512 * (st0 + v));
            wl[v] = ld8(YL + 512 * (st0 + v));
        }
        __syncthreads();  // every wavefront is done with the previous X block
        {
            const uint4 *XHb = reinterpret_cast<const uint4 *>(xh + xb), *XLb = reinterpret_cast<const uint4 *>(xl + xb);
#pragma unroll
            for (uint32_t k = 0; k < XN / (64 * WG); k++) {
                const uint4 vh = XHb[tid + 64 * WG * k], vl = XLb[tid + 64 * WG * k];
                xs[tid + 64 * WG * k] = vh;
                xs[XN + tid + 64 * WG * k] = vl;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t v = 0; v + 1 < XW; v++) {
            xhv[v] = __builtin_bit_cast(bf16x8, xs[v * 64 + lane]);
            xlv[v] = __builtin_bit_cast(bf16x8, xs[XN + v * 64 + lane]);
        }
#pragma unroll
        for (uint32_t sv = 0; sv < CS; sv++) {  // unrolled: every slot index is a constant
            const uint32_t sn = st0 + sv;
            if (sv + YW - 1 < CS + Q - 1) {  // Y rows up to step CS + Q - 2 are used
                wh[(sv + YW - 1) % YW] = ld8(YH + 512 * (sn + YW - 1));
                wl[(sv + YW - 1) % YW] = ld8(YL + 512 * (sn + YW - 1));
            }
            if (sv + XW - 1 < CS) {
                xhv[(sv + XW - 1) % XW] = __builtin_bit_cast(bf16x8, xs[(sv + XW - 1) * 64 + lane]);
                xlv[(sv + XW - 1) % XW] = __builtin_bit_cast(bf16x8, xs[XN + (sv + XW - 1) * 64 + lane]);
            }
            // keep this step's loads ahead of its MFMAs (the scheduler would
            // otherwise sink them next to their first use)
            __builtin_amdgcn_sched_barrier(0);
            const bf16x8 ah = xhv[sv % XW], al = xlv[sv % XW];
#pragma unroll
            for (uint32_t j = 0; j < Q; j++) {
                const uint32_t w = (sv + j) % YW;
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, wh[w], acc[j], 0, 0, 0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, wl[w], acc[j], 0, 0, 0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, wh[w], acc[j], 0, 0, 0);
                // tile 0 holds lag 0, where lo*lo' = lo^2 >= 0 for a == b
                // does not average out over t0: add it there
                if (LAG0 && j == 0)
                    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, wl[w], acc[0], 0, 0, 0);
            }
        }
    }
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Operand layouts of v_mfma_f32_32x32x16_bf16: lane l feeds row l % 32 of A
// and column l % 32 of B, each with the 8 k-values of slot l / 32 (the same k
// slots in A and B); result item e of lane l is row 8(e/4) + 4(l/32) + e%4,
// column l % 32.  Grid: x = WG consecutive tile groups (one per
// wavefront), y = current pair, z = (block, chunk).  The 63 diagonal sums of
// a tile are pairwise FP64 sums in a fixed order, so results do not depend
// on the launch.  At least 2 wavefronts per SIMD (up to 256 VGPRs each).
template <uint32_t Q, uint32_t CS, uint32_t WG>
__global__ __launch_bounds__(64 * WG, 2) void gkm_tile_kernel(const GkmBatch a) {
    static_assert(CS % WG == 0, "the X block is copied in WG equal parts");
    // X blocks during the K loop, then one 32 x 33 FP64 tile per wavefront
    constexpr uint32_t TILE_D = WG * 32 * 33, XBLK_D = 2 * CS * 64 * sizeof(uint4) / sizeof(double);
    __shared__ double smem[TILE_D > XBLK_D ? TILE_D : XBLK_D];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * WG + wv, pair = blockIdx.y, z = blockIdx.z;
    uint32_t b = 0;
    while (z >= a.z0[b + 1]) b++;  // uniform
    const uint32_t chunk = z - a.z0[b], nchunks = a.z0[b + 1] - a.z0[b];
    const uint32_t PP = a.PP[b], YP = a.YP[b];
    const GkmOff off = gkm_off(a.ncx, PP, YP);
    const char *reg = a.work + a.base[b];
    const unsigned short *xh = (const unsigned short *)(reg + off.xh), *xl = (const unsigned short *)(reg + off.xl);
    const unsigned short *yh = (const unsigned short *)(reg + off.yh), *yl = (const unsigned short *)(reg + off.yl);
    const uint32_t qbase = (g / 16) * 16 * Q, q0 = g % 16;
    const uint32_t ca = pair / a.ncur, cb = pair % a.ncur, ncx = a.ncx;
    const uint32_t lane = threadIdx.x & 63u, r = lane & 31u, h = lane >> 5;
    const uint32_t c8 = q0 & 7u, qoff = qbase + (q0 & ~7u);
    const uint32_t st0 = chunk * CS;
    uint4 *xs = reinterpret_cast<uint4 *>(smem);
    f32x16 acc[Q];
#pragma unroll
    for (uint32_t j = 0; j < Q; j++) acc[j] = f32x16{};
    // the wavefront holding tile 0 (lag 0) adds the fourth product; the branch
    // sits outside the MFMA loop so that its loads schedule freely (both
    // paths pass the same workgroup barriers)
    if (g == 0)
        gkm_accumulate<Q, CS, WG, true>(acc, xs, xh, xl, yh, yl, PP, YP, ca, cb, ncx, r, h, c8, qoff, st0);
    else
        gkm_accumulate<Q, CS, WG, false>(acc, xs, xh, xl, yh, yl, PP, YP, ca, cb, ncx, r, h, c8, qoff, st0);
    __syncthreads();  // X blocks are dead: the tiles below reuse their LDS
    // diagonal sums of each tile, in FP64, into this chunk's partial; each
    // wavefront works in its own tile, so only wavefront-level syncs
    double(*tile)[33] = reinterpret_cast<double(*)[33]>(smem + wv * 32 * 33);
    double *part = a.partial + a.poff[b];
#pragma unroll
    for (uint32_t j = 0; j < Q; j++) {
        const uint32_t q = qbase + q0 + 16 * j;
        if (q >= a.ntiles) break;  // uniform across the wavefront
#pragma unroll
        for (int e = 0; e < 16; e++) tile[8 * (e / 4) + 4 * h + (e % 4)][r] = (double)acc[j][e];
        wave_sync_lds();
        if (lane < 63) {
            // diagonal d = s - r of this lane: all 32 rows read at once (rows
            // off the tile read a clamped cell and add 0), pairwise sum
            const int d = (int)lane - 31;
            double x[32];
#pragma unroll
            for (int rr = 0; rr < 32; rr++) {
                const int c = rr + d;
                const double v = tile[rr][min(max(c, 0), 31)];
                x[rr] = (c >= 0 && c < 32) ? v : 0.0;
            }
#pragma unroll
            for (int w = 16; w > 0; w >>= 1)
#pragma unroll
                for (int i = 0; i < w; i++) x[i] += x[i + w];
            part[(((uint64_t)pair * a.ntiles + q) * nchunks + chunk) * 63 + lane] = x[0];
        }
        wave_sync_lds();
    }
}

// Segment mean of product/3 for lag l and pair (diagonal l - 32q of tiles
// q = floor(l/32), floor(l/32) + 1; chunks in order), folded into the running
// mean over segments (:230-237); the first segment also clears the other
// columns of the row.  Grid: x = work item, y = block.  The first
// npairs * leff items are (pair column, lag), consecutive threads taking
// consecutive lags of one column, so that their partial-sum reads (63
// diagonals per tile and chunk) are contiguous; the rest clear the other
// columns.
__global__ __launch_bounds__(256) void gkm_reduce_kernel(const GkmBatch a, uint32_t iseg) {
    const uint32_t b = blockIdx.y;
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.leff * a.narr) return;
    const uint32_t npairs = a.ncur * a.ncur, npw = npairs * a.leff;
    if (w >= npw) {
        if (iseg == 0) {
            const uint32_t r = w - npw, nother = a.narr - npairs;
            a.out[b * a.stride + (uint64_t)(r / nother) * a.narr + npairs + r % nother] = 0.0;
        }
        return;
    }
    const uint32_t col = w / a.leff, l = w % a.leff;
    double *v = a.out + b * a.stride + (uint64_t)l * a.narr + col;
    const uint32_t nchunks = a.z0[b + 1] - a.z0[b];
    // four running sums over the chunks (c mod 4), added in a fixed order:
    // independent chains keep four loads in flight
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t q = l / 32; q <= l / 32 + 1 && q < a.ntiles; q++) {
        const int d = (int)l - 32 * (int)q;
        if (d < -31 || d > 31) continue;
        const double *p = a.partial + a.poff[b] + ((uint64_t)col * a.ntiles + q) * nchunks * 63 + (d + 31);
        uint32_t ch = 0;
        for (; ch + 4 <= nchunks; ch += 4) {
#pragma unroll
            for (int k = 0; k < 4; k++) s4[k] += p[(uint64_t)(ch + k) * 63];
        }
        for (; ch < nchunks; ch++) s4[ch & 3u] += p[(uint64_t)ch * 63];
    }
    const double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    const uint32_t seg = a.n[b] / a.n_seg;
    const uint32_t lo = seg * iseg, hi = seg * (iseg + 1);
    const uint32_t first = lo + lo % a.skip;  // greenkubo...cpp:208
    const uint32_t cont = first < hi ? (hi - first + a.skip - 1) / a.skip : 0;
    const double jj = cont > 0 ? (s / 3.0) / (double)cont : 0.0;
    const double old = iseg == 0 ? 0.0 : *v;
    *v = old + (jj - old) / (double)(iseg + 1);
}

}  // namespace anl

using namespace anl;

namespace {

// Sizes of one block's operands, from its t0 count n and the lag count leff.
struct GkmSize {
    uint32_t nchunks, PP, YP, ntiles, ngroups;
    uint64_t region, partial;  // bytes of the operand region, doubles of partial sums
};

// Launch shape of the tile kernel (see gkm_tile_kernel): tiles per
// wavefront, MFMA steps per FP32 chunk, wavefronts per workgroup.
struct GkmShape {
    uint32_t q, cs, wg;
};

GkmShape gkm_shape_of(const anl_series *s) { return GkmShape{s->gk_q, s->gk_cs, s->gk_wg}; }

GkmSize gkm_size(uint32_t ncur, uint32_t leff, uint32_t n, const GkmShape &sh) {
    GkmSize z{};
    const uint64_t Q = sh.q, CS = sh.cs;
    // t0 axis: rows of 32, MFMA steps of 16 rows, chunks of CS steps
    const uint64_t nsteps = ((uint64_t)n + 511) / 512;
    const uint64_t nchunks = (nsteps + CS - 1) / CS;
    const uint64_t PP = nchunks * CS * 16;
    // lag axis: tiles of 32 (plus one for the negative diagonals)
    const uint64_t ntiles = (leff + 30) / 32 + 1;
    const uint64_t ngroups = (ntiles + 16 * Q - 1) / (16 * Q) * 16;
    const uint64_t qmax = ngroups / 16 * 16 * Q;  // tiles covered, >= ntiles
    // highest Y row read: 16 (st + Q - 1) + 8 + 7 + qoff, st < nchunks * CS, qoff <= qmax - 16 Q + 8;
    // rounded up to whole split workgroups
    const uint64_t YP = (PP + 16 * Q + qmax + 16 + 7) / GKM_SP * GKM_SP;
    if (YP >= (1ull << 31) || nchunks > 65535)
        fail(ANL_ERR_ARG, "series too long for the MFMA form (use the FP64 form)");
    z.nchunks = (uint32_t)nchunks;
    z.PP = (uint32_t)PP;
    z.YP = (uint32_t)YP;
    z.ntiles = (uint32_t)ntiles;
    z.ngroups = (uint32_t)ngroups;
    z.region = gkm_off(3 * ncur, z.PP, z.YP).end;
    z.partial = (uint64_t)ncur * ncur * ntiles * nchunks * 63;
    return z;
}

// Scratch of a batch of blocks sharing one set of launches.
constexpr uint64_t GKM_BATCH_BYTES = 4ull << 30;

uint64_t gkm_need(const GkmSize &z) { return z.region + align256(z.partial * sizeof(double)); }

// Blocks [b0, b1) of one batch: at most GKM_MAXB blocks, 65,535 chunks and
// GKM_BATCH_BYTES of scratch (a single block may exceed the latter).
uint32_t gkm_batch_end(uint32_t ncur, uint32_t leff, const uint32_t *n, uint32_t nblocks, uint32_t b0,
                       const GkmShape &sh, uint64_t *bytes) {
    uint64_t tot = 0, chunks = 0;
    uint32_t b = b0;
    for (; b < nblocks && b - b0 < GKM_MAXB; b++) {
        const GkmSize z = gkm_size(ncur, leff, n[b], sh);
        if (b > b0 && (tot + gkm_need(z) > GKM_BATCH_BYTES || chunks + z.nchunks > 65535)) break;
        tot += gkm_need(z);
        chunks += z.nchunks;
    }
    *bytes = tot;
    return b;
}

// Enqueues the lag products of blocks [b0, b1) on s; block b's output
// d + b * stride [leff][narr] gets C(l) in its first ncur^2 columns and zeros
// elsewhere.
void gkm_enqueue_batch(anl_series *sr, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t leff,
                       uint32_t n_seg, const uint64_t *primo, const uint32_t *n, uint32_t b0, uint32_t b1, char *work,
                       double *d, uint64_t stride, const GkmShape &sh, hipStream_t s) {
    GkmBatch a{};
    a.data = sr->data;
    a.work = work;
    a.ncols = sr->ncols;
    a.nrows = sr->nrows;
    a.ncur = ncur;
    a.ncx = 3 * ncur;
    a.leff = leff;
    a.narr = 3 * ncur * ncur + 2;
    a.skip = skip;
    a.n_seg = n_seg;
    a.nblk = b1 - b0;
    a.out = d + b0 * stride;
    a.stride = stride;
    for (uint32_t c = 0; c < ncur; c++) a.idx[c] = idx_j[c];
    uint64_t region = 0, partial = 0;
    uint32_t maxPP = 0, maxYP = 0, ngroups = 0;
    a.z0[0] = 0;
    for (uint32_t k = 0; k < a.nblk; k++) {
        const GkmSize z = gkm_size(ncur, leff, n[b0 + k], sh);
        a.primo[k] = primo[b0 + k];
        a.n[k] = n[b0 + k];
        a.PP[k] = z.PP;
        a.YP[k] = z.YP;
        a.base[k] = region;
        a.poff[k] = partial;
        a.z0[k + 1] = a.z0[k] + z.nchunks;
        a.ntiles = z.ntiles;
        ngroups = z.ngroups;
        region += z.region;
        partial += align256(z.partial * sizeof(double)) / sizeof(double);
        maxPP = std::max(maxPP, z.PP);
        maxYP = std::max(maxYP, z.YP);
    }
    a.partial = (double *)(work + region);
    const uint32_t ncorr = ncur * ncur, ncx = 3 * ncur;
    const size_t lds_x = 2 * (size_t)ncx * 32 * (GKM_SP + 1) * sizeof(unsigned short);
    const size_t lds_y = 2 * (size_t)ncx * 32 * (GKM_SP + 8) * sizeof(unsigned short);
    hipLaunchKernelGGL(gkm_split_kernel<true>, dim3(maxYP / GKM_SP, 1, a.nblk), dim3(256), lds_y, s, a, 0u);
    ANL_HIP(hipGetLastError());
    const dim3 rgrid((unsigned)(((uint64_t)leff * a.narr + 255) / 256), a.nblk);
    for (uint32_t iseg = 0; iseg < n_seg; iseg++) {
        hipLaunchKernelGGL(gkm_split_kernel<false>, dim3(maxPP / GKM_SP, 1, a.nblk), dim3(256), lds_x, s, a, iseg);
        ANL_HIP(hipGetLastError());
        const dim3 tgrid(ngroups / sh.wg, ncorr, a.z0[a.nblk]);
#define ANL_GKM_TILE(Q, CS, WG) hipLaunchKernelGGL((gkm_tile_kernel<Q, CS, WG>), tgrid, dim3(64 * WG), 0, s, a)
        switch (sh.q * 100 + sh.cs * 10 + sh.wg) {
            case 4 * 100 + 8 * 10 + 2: ANL_GKM_TILE(4, 8, 2); break;
            case 4 * 100 + 8 * 10 + 4: ANL_GKM_TILE(4, 8, 4); break;
            case 4 * 100 + 16 * 10 + 2: ANL_GKM_TILE(4, 16, 2); break;
            case 4 * 100 + 16 * 10 + 4: ANL_GKM_TILE(4, 16, 4); break;
            case 8 * 100 + 8 * 10 + 2: ANL_GKM_TILE(8, 8, 2); break;
            case 8 * 100 + 8 * 10 + 4: ANL_GKM_TILE(8, 8, 4); break;
            case 8 * 100 + 16 * 10 + 2: ANL_GKM_TILE(8, 16, 2); break;
            case 8 * 100 + 16 * 10 + 4: ANL_GKM_TILE(8, 16, 4); break;
            default: fail(ANL_ERR_ARG, "unsupported Green-Kubo MFMA launch shape");
        }
#undef ANL_GKM_TILE
        ANL_HIP(hipGetLastError());
        hipLaunchKernelGGL(gkm_reduce_kernel, rgrid, dim3(256), 0, s, a, iseg);
        ANL_HIP(hipGetLastError());
    }
}

}  // namespace

namespace anl {

bool gkm_shape_ok(uint32_t q, uint32_t cs, uint32_t wg) {
    return (q == 4 || q == 8) && (cs == 8 || cs == 16) && (wg == 2 || wg == 4);
}

size_t gk_mfma_scratch_bytes(const anl_series *s, uint32_t ncur, uint32_t leff, uint32_t n) {
    uint64_t bytes = 0;
    gkm_batch_end(ncur, leff, &n, 1, 0, gkm_shape_of(s), &bytes);
    return bytes;
}

void gk_mfma_enqueue_blocks(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t leff,
                            uint32_t n_seg, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, double *d,
                            hipStream_t st) {
    const GkmShape sh = gkm_shape_of(s);
    uint64_t need = 0;
    for (uint32_t b0 = 0; b0 < nblocks;) {
        uint64_t bytes = 0;
        b0 = gkm_batch_end(ncur, leff, n, nblocks, b0, sh, &bytes);
        need = std::max(need, bytes);
    }
    char *work = (char *)series_workspace(s, need);
    const uint64_t stride = (uint64_t)leff * (3 * ncur * ncur + 2);
    // batches run one after the other on st, sharing the scratch
    for (uint32_t b0 = 0; b0 < nblocks;) {
        uint64_t bytes = 0;
        const uint32_t b1 = gkm_batch_end(ncur, leff, n, nblocks, b0, sh, &bytes);
        gkm_enqueue_batch(s, idx_j, ncur, skip, leff, n_seg, primo, n, b0, b1, work, d, stride, sh, st);
        b0 = b1;
    }
}

}  // namespace anl
