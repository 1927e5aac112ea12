// Spherical-harmonic density correlations: SphericalCorrelations<lmax>
// (sphericalcorrelations.cpp:62-270) over SphericalBase::calc
// (sphericalbase.cpp:19-66, no neighbour list), FP64 throughout.
//
// For every frame f and atom i the reference builds the density
//   S_f[i][type_j][bin][l,m] = sum over the other atoms j in the bin of the
//   pair range (type_i, type_j) of Y_lm(x_i - x_j)   (real harmonics),
// and the output is, per lag dt, the mean over t1 = 0, skip, .. < n of the
// mean over the atoms i of each type of sum_m S_{t1}[i] S_{t1+dt}[i].
//
// Per chunk of t1 values (chunks bound the HBM held by the densities):
//  1. sph_density_kernel: one wavefront per (frame, central atom).  The lanes
//     take 64 other atoms at a time (whole 64-slot blocks are skipped by a
//     bounding-sphere test in orthorhombic cells): FP64 minimum image in the
//     reference's loop order, distance bin from exact host-computed edges of
//     the pair's range, ballot of the atoms in range.  The atoms in range
//     join a queue held one per lane, in slot order; 64/(lmax+1) at a time
//     they are expanded by lanes (slot, m), lane m running the reference's
//     Legendre recursion of column m and the Chebyshev recursion of
//     cos/sin(m phi), and added into the wave's density block in LDS in queue
//     order.  The block is written once, as
//     S[frame][type_j][bin][l][atom by type][m].
//  2. sph_gram_mfma_kernel: for every problem (type_i, type_j, bin, l) the
//     lag products sum_{i of type_i} sum_m S_{t1} S_{t2} are a banded Gram
//     matrix over frames with K = (atoms of type_i) x (2l+1) contiguous in
//     HBM.  A 256-thread workgroup takes a block of 64 frames t2 and 32 of
//     the t1 rows whose band reaches it, on FP64 MFMA (16 x 16 x 4), K staged
//     through LDS 32 at a time; entries with 0 <= t2 - t1 < leff go to
//     G[problem][t1 index][t2 - t1].
//  3. sph_reduce_kernel: R[problem][dt] += sum over t1 of G, in t1 order.
// The host divides R by (number of t1) x (atoms of type_i) and writes the
// reference layout [dt][type_i][type_j][bin][k] with k holding l = lmax - k.
// Every sum has a fixed order: results are deterministic.
//
// Steinhardt<lmax> (steinhardt.cpp:90-236) reuses step 1 on frames
// primo, primo + skip, ..; sph_steinhardt_kernel then turns each (atom,
// type_j, bin) density into q_l and either a histogram count (64-bit atomics,
// exact) or the q_1..q_lmax trajectory.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

#include "min_image.h"

constexpr int SPH_WAVES = 4;           // wavefronts (central atoms) per density workgroup
constexpr uint32_t SPH_WAVE_LDS = 2048;  // doubles of LDS per wavefront (16 KiB): density block + staging
constexpr int GT = 64;                 // t1 values per chunk are rounded to multiples of this

// doubles of a wavefront's staging area: the harmonic vectors of 64/(lmax+1)
// atoms, and at least 64 queue entries of 4 doubles
__host__ __device__ inline uint32_t stage_doubles(uint32_t lmax) {
    const uint32_t h = 64 / (lmax + 1) * (lmax + 1) * (lmax + 1);
    return h > 256 ? h : 256;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SphDensArgs {
    const double4 *pos;        // [F][N] slots
    const double *box;         // [F][9]
    const int32_t *type_slot;  // [N]
    const int32_t *aprime;     // [N] index of the slot among the atoms sorted by type
    const double *thr;         // [nt*nt][nbin+1] bin edges of each pair range (in d2)
    const double *coef;        // [(lmax+1)^2]: coefficient of (l, |m|) at l*(lmax+1)+|m|
    double *S;                 // [nfr][FS]
    uint64_t N, FS, f0;        // atoms, doubles per frame, first frame
    uint64_t fstride;          // frame fl of the launch is f0 + fl * fstride
    uint32_t nt, nbin, lmax;
    uint32_t nbc;              // bins per launch column (blockIdx.y)
    uint32_t ngroups;          // ceil(N / SPH_WAVES)
    double d2lo, d2hi;         // min over type pairs of T[0], max of T[nbin]
    double rcut;               // > sqrt(d2hi), with a rounding margin
    const double4 *block_rad;  // [nfr][ceil(N/64)] {centre, radius} of each 64-slot block, or null (triclinic)
};

// bounding sphere of each 64-slot block of each frame (orthorhombic cells):
// centre = first atom + mean minimum-image offset of the block's atoms, radius
// = largest minimum-image distance from the centre, rounded up by a relative
// 1e-12; one wavefront per (frame, block), out[.] = {centre, radius}
__global__ __launch_bounds__(64) void sph_block_radius_kernel(const double4 *pos, const double *box, uint64_t N,
                                                              uint64_t f0, uint64_t fstride, double4 *out) {
    const uint64_t nblk = (N + 63) / 64;
    const uint64_t fl = blockIdx.x / nblk, blk = blockIdx.x - fl * nblk;
    const uint64_t f = f0 + fl * fstride;
    const double4 *P = pos + f * N;
    const DevCell c = load_cell(box + f * 9);
    const uint64_t j = blk * 64 + threadIdx.x;
    const uint64_t nin = min((uint64_t)64, N - blk * 64);
    const double4 r = P[blk * 64];
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if (j < N) {
        const double4 pj = P[j];
        x0 = pj.x - r.x;
        x1 = pj.y - r.y;
        x2 = pj.z - r.z;
        min_image_ref<false>(x0, x1, x2, c);
    }
    double s0 = x0, s1 = x1, s2 = x2;
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    const double c0 = r.x + s0 / (double)nin, c1 = r.y + s1 / (double)nin, c2 = r.z + s2 / (double)nin;
    double d = 0.0;
    if (j < N) {
        const double4 pj = P[j];
        double y0 = pj.x - c0, y1 = pj.y - c1, y2 = pj.z - c2;
        min_image_ref<false>(y0, y1, y2, c);
        d = sqrt(d2_of(y0, y1, y2));
    }
    for (int o = 32; o > 0; o >>= 1) d = fmax(d, __shfl_xor(d, o));
    if (threadIdx.x == 0) out[blockIdx.x] = make_double4(c0, c1, c2, d * (1.0 + 1e-12) + 1e-12);
}

// Lane m's share of the real harmonics of direction (x, y, z)
// The direction of x (specialfunctions.h:11-22): cos theta, cos phi and
// sin phi replace x0, x1, x2, computed once per neighbour and shared by its m
// lanes.
__device__ __forceinline__ void direction(double &x0, double &x1, double &x2) {
    const double rxy = sqrt(x0 * x0 + x1 * x1);
    const double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    const double ct = x2 / r, sp = x1 / rxy, cp = x0 / rxy;
    x0 = ct;
    x1 = cp;
    x2 = sp;
}

// (specialfunctions.h:11-22, 262-360): P_l^m for l = m..lmax and
// cos/sin(m phi), written to y[l*l + l + m] (cos) and y[l*l + l - m] (sin);
// cf[l] is the coefficient of (l, m).  The loops run to the largest lmax with
// the lane's own range as a predicate, so cf stays in registers.  The
// recursion multiplies by 1/(l-m), which depends on l and m only, where the
// reference divides by l-m: the division leaves the dependency chain of the
// recursion, and each step changes by at most one rounding.
__device__ __forceinline__ void put_harmonics(double *y, double ct, double cp, double sp, int m, int lmax,
                                              const double (&cf)[ANL_SPH_LMAX_MAX + 1]) {
    const double st = sqrt(1 - ct * ct);
    double pmm = 1.0;
#pragma unroll
    for (int l = 1; l <= ANL_SPH_LMAX_MAX; ++l)
        if (l <= m) pmm = (double)(-(2 * l - 1)) * st * pmm;
    double cm = m > 0 ? cp : 1.0, sm = m > 0 ? sp : 0.0;
    {
        double c2 = 1.0, s2 = 0.0, c1 = cp, s1 = sp;
#pragma unroll
        for (int k = 2; k <= ANL_SPH_LMAX_MAX; ++k)
            if (k <= m) {
                const double cn = 2 * cp * c1 - c2;
                const double sn = 2 * cp * s1 - s2;
                c2 = c1;
                s2 = s1;
                c1 = cn;
                s1 = sn;
                cm = c1;
                sm = s1;
            }
    }
    double p1 = 0.0, p2 = 0.0;  // P_{l-1}^m, P_{l-2}^m
#pragma unroll
    for (int l = 0; l <= ANL_SPH_LMAX_MAX; ++l) {
        if (l >= m && l <= lmax) {
            double p;
            if (l == m) p = pmm;
            else if (l == m + 1) p = pmm * ct * (2 * m + 1);
            else p = (p1 * ct * (2 * l - 1) - p2 * (l + m - 1)) * (1.0 / (double)(l - m));
            const double v = p * cf[l];
            y[l * l + l + m] = v * cm;
            if (m > 0) y[l * l + l - m] = v * sm;
            p2 = p1;
            p1 = p;
        }
    }
}

// Expands the nb oldest queue entries (entry k in lane k: direction, as
// direction() leaves it, and type << 16 | bin - b0): lane (slot, m) puts entry
// slot's harmonics into its staging block, the blocks are added into the
// density block acc in queue order (each entry by one lane, in order), and the
// queue moves down by nb.
__device__ __forceinline__ void expand_queue(uint32_t nb, double &qx0, double &qx1, double &qx2, uint32_t &qcode,
                                             uint32_t &qn, double *acc, double *stage, uint32_t nbc, uint32_t L2,
                                             uint32_t lane, uint32_t slot, uint32_t m, int lmax,
                                             const double (&cf)[ANL_SPH_LMAX_MAX + 1]) {
    const int src = slot < nb ? (int)slot : 0;
    const double y0 = __shfl(qx0, src), y1 = __shfl(qx1, src), y2 = __shfl(qx2, src);
    if (slot < nb) put_harmonics(stage + slot * L2, y0, y1, y2, (int)m, lmax, cf);
    wave_sync_lds();
    for (uint32_t k = 0; k < nb; ++k) {
        const uint32_t code = (uint32_t)__shfl((int)qcode, (int)k);
        double *blk = acc + ((code >> 16) * nbc + (code & 0xFFFFu)) * L2;
        for (uint32_t e = lane; e < L2; e += 64) blk[e] += stage[k * L2 + e];
    }
    wave_sync_lds();
    const int up = (int)min(lane + nb, 63u);
    qx0 = __shfl(qx0, up);
    qx1 = __shfl(qx1, up);
    qx2 = __shfl(qx2, up);
    qcode = (uint32_t)__shfl((int)qcode, up);
    qn -= nb;
}

// Writes atom i's density block acc (bins b0 .. b0 + nbc - 1 of every type)
// to S[frame fl][type_j][bin][l][atom by type][m].
__device__ __forceinline__ void store_density(const SphDensArgs &a, const double *acc, uint32_t fl, uint64_t i,
                                              uint32_t b0, uint32_t nbc, uint32_t lane) {
    const uint32_t L2 = (a.lmax + 1) * (a.lmax + 1);
    double *Sf = a.S + (uint64_t)fl * a.FS;
    const uint64_t ap = (uint64_t)a.aprime[i];
    // blocks (type_j, bin) in a uniform loop, the (l, m) entries of a block
    // across the lanes: no per-element division
    for (uint32_t jt = 0; jt < a.nt; ++jt)
        for (uint32_t bl = 0; bl < nbc && b0 + bl < a.nbin; ++bl) {
            const double *src = acc + (jt * nbc + bl) * L2;
            const uint64_t boff = (uint64_t)(jt * a.nbin + b0 + bl) * L2;
            for (uint32_t e = lane; e < L2; e += 64) {
                uint32_t l = (uint32_t)sqrtf((float)e);
                while (l * l > e) --l;
                while ((l + 1) * (l + 1) <= e) ++l;
                Sf[(boff + l * l) * a.N + ap * (2 * l + 1) + (e - l * l)] = src[e];
            }
        }
}

template <bool TRI>
__global__ __launch_bounds__(256, 3) void sph_density_kernel(SphDensArgs a) {
    extern __shared__ double sph_lds[];
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t L1 = a.lmax + 1, L2 = L1 * L1;
    const uint32_t accn = a.nt * a.nbc * L2;
    const uint32_t nc = 64 / L1;                       // candidates expanded at once
    const uint32_t slot = lane / L1, m = lane - slot * L1;
    double *acc = sph_lds + wv * (accn + stage_doubles(a.lmax));
    double *stage = acc + accn;                        // [nc][L2] harmonics, or [64][4] new queue entries
    double cf[ANL_SPH_LMAX_MAX + 1];                   // coefficients of (l, m) of this lane's m
#pragma unroll
    for (int l = 0; l <= ANL_SPH_LMAX_MAX; ++l)
        cf[l] = (l >= (int)m && l <= (int)a.lmax) ? a.coef[l * L1 + m] : 0.0;
    const uint32_t fl = blockIdx.x / a.ngroups;
    const uint64_t i = (uint64_t)(blockIdx.x % a.ngroups) * SPH_WAVES + wv;
    if (i >= a.N) return;  // the whole wavefront; no workgroup barrier below
    const uint32_t b0 = blockIdx.y * a.nbc;
    const uint32_t b1 = min(b0 + a.nbc, a.nbin);
    for (uint32_t k = lane; k < accn; k += 64) acc[k] = 0.0;
    wave_sync_lds();

    const uint64_t f = a.f0 + fl * a.fstride;
    const double4 *P = a.pos + f * a.N;
    const DevCell c = load_cell(a.box + f * 9);
    const double4 pi = P[i];
    const int ti = a.type_slot[i];
    const double *thr_i = a.thr + (size_t)ti * a.nt * (a.nbin + 1);
    const double4 *rad = a.block_rad ? a.block_rad + (uint64_t)fl * ((a.N + 63) / 64) : nullptr;
    // queue of atoms in range, in slot order: entry k in lane k (minimum-image
    // vector and type << 16 | bin - b0); qn entries, wavefront-uniform
    double qx0 = 0.0, qx1 = 0.0, qx2 = 0.0;
    uint32_t qcode = 0, qn = 0;
    const uint64_t nblk = (a.N + 63) / 64;
    // Blocks of 64 slots, 64 blocks at a time: lane b decides whether block b
    // can hold an atom within rcut of atom i.  In orthorhombic cells the
    // minimum image is the nearest one, a metric, so a block whose bounding
    // sphere's centre is farther than rcut + its radius cannot.  The atoms in
    // range of each near block join the queue.  The queue is drained when a
    // block's atoms do not fit and at the end: every entry's direction is
    // computed in one pass over the lanes, then the entries are expanded nc
    // at a time.  One expansion site keeps the code compact.
    uint64_t g = 0;
    unsigned long long blocks = 0;
    bool pending = false, finished = false, in = false;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    uint32_t code = 0, cnt = 0;
    unsigned long long mask = 0;
    for (;;) {
        if (!pending && !finished) {
            while (blocks == 0 && g < nblk) {
                bool near = g + lane < nblk;
                if (rad && near) {
                    const double4 bc = rad[g + lane];
                    double y0 = pi.x - bc.x, y1 = pi.y - bc.y, y2 = pi.z - bc.z;
                    min_image_ref<false>(y0, y1, y2, c);
                    near = !(sqrt(d2_of(y0, y1, y2)) > a.rcut + bc.w);
                }
                blocks = __ballot(near);
                if (blocks == 0) g += 64;
            }
            if (blocks == 0) {
                finished = true;
            } else {
                const uint64_t jb = (g + (uint64_t)(__ffsll(blocks) - 1)) * 64;
                blocks &= blocks - 1;
                if (blocks == 0) g += 64;
                const uint64_t j = jb + lane;
                in = false;
                x0 = x1 = x2 = 0.0;
                code = 0;
                if (j < a.N && j != i) {
                    const double4 pj = P[j];
                    x0 = pi.x - pj.x;
                    x1 = pi.y - pj.y;
                    x2 = pi.z - pj.z;
                    min_image_ref<TRI>(x0, x1, x2, c);
                    const double d2 = d2_of(x0, x1, x2);
                    if (d2 >= a.d2lo && d2 < a.d2hi) {  // outer bounds over every type pair
                        const uint32_t tj = (uint32_t)a.type_slot[j];
                        const double *T = thr_i + (size_t)tj * (a.nbin + 1);
                        if (d2 >= T[b0] && d2 < T[b1]) {
                            in = true;
                            uint32_t b = b0;
                            while (b + 1 < b1 && d2 >= T[b + 1]) ++b;
                            code = tj << 16 | (b - b0);  // b - b0 < nbc <= SPH_WAVE_LDS
                        }
                    }
                }
                mask = __ballot(in);
                cnt = (uint32_t)__popcll(mask);
                pending = cnt > 0;
            }
        }
        if (pending && qn + cnt <= 64) {
            // append in slot order: compact through LDS, lanes qn..qn+cnt-1 take them
            if (in) {
                double *e = stage + 4 * __popcll(mask & ((1ull << lane) - 1));
                e[0] = x0;
                e[1] = x1;
                e[2] = x2;
                e[3] = (double)code;
            }
            wave_sync_lds();
            if (lane >= qn && lane < qn + cnt) {
                const double *e = stage + 4 * (lane - qn);
                qx0 = e[0];
                qx1 = e[1];
                qx2 = e[2];
                qcode = (uint32_t)e[3];
            }
            wave_sync_lds();
            qn += cnt;
            pending = false;
        }
        if (qn > 0 && (pending || finished)) {
            if (lane < qn) direction(qx0, qx1, qx2);
            while (qn > 0)
                expand_queue(min(qn, nc), qx0, qx1, qx2, qcode, qn, acc, stage, a.nbc, L2, lane, slot, m,
                             (int)a.lmax, cf);
        }
        if (finished) break;
    }
    wave_sync_lds();
    store_density(a, acc, fl, i, b0, a.nbc, lane);
}

constexpr uint32_t SANN_MAXT = 16;      // atom types of a neighbour list
constexpr uint32_t SANN_MAXTOT = 1024;  // sum over types of the list capacities

struct SannArgs {
    uint32_t nn[SANN_MAXT];     // capacity of the list of each neighbour type
    uint32_t off[SANN_MAXT];    // its first entry
    double cut2[SANN_MAXT];     // squared cutoff of each neighbour type
    uint32_t total, capmax;     // sum and maximum of nn (total rounded up to even)
    int *overflow;              // set when the reference would throw "Too many neighbours in shell!"
    int32_t *nbr;               // [nfr][N][total] chosen neighbours (type-sorted index) from off[t], or null
    const int32_t *orig;        // [N] original atom index of each slot
};

// SphericalBase::calc with a neighbour list (neighbour.cpp:8-101,
// sphericalbase.cpp:67-93): one wavefront per (frame, atom i).  For every
// type t the atoms j != i of type t with d2 <= cut2[t] are listed in LDS
// (distance, slot), sorted by (distance, slot) with a rank sort, and the
// first SANN-count of them (3, then more while the sum of the first m
// distances exceeds r_m (m - 2)) are expanded into bin 0; the other bins are
// zero.  The neighbour count is the l = 0 entry over Y_00.
template <bool TRI>
__global__ __launch_bounds__(64) void sph_density_sann_kernel(SphDensArgs a, SannArgs s) {
    extern __shared__ double sann_lds[];
    const uint32_t lane = threadIdx.x;
    const uint32_t L1 = a.lmax + 1, L2 = L1 * L1;
    const uint32_t nc = 64 / L1;
    const uint32_t slot = lane / L1, m = lane - slot * L1;
    double *acc = sann_lds;                           // [nt][L2], bin 0
    double *stage = acc + a.nt * L2;
    double *ru = stage + stage_doubles(a.lmax);       // [total] distances, in list order
    double *rs = ru + s.total;                        // [capmax] sorted distances of one type
    int *ju = (int *)(rs + s.capmax);                 // [total] slots, in list order
    int *js = ju + s.total;                           // [capmax] sorted slots of one type
    int *cnt = js + s.capmax;                         // [nt] list lengths (may exceed capacity)
    double cf[ANL_SPH_LMAX_MAX + 1];
#pragma unroll
    for (int l = 0; l <= ANL_SPH_LMAX_MAX; ++l)
        cf[l] = (l >= (int)m && l <= (int)a.lmax) ? a.coef[l * L1 + m] : 0.0;
    const uint32_t fl = (uint32_t)(blockIdx.x / a.N);
    const uint64_t i = blockIdx.x - (uint64_t)fl * a.N;
    for (uint32_t k = lane; k < a.nt * L2; k += 64) acc[k] = 0.0;
    if (lane < a.nt) cnt[lane] = 0;
    wave_sync_lds();
    const uint64_t f = a.f0 + fl * a.fstride;
    const double4 *P = a.pos + f * a.N;
    const DevCell c = load_cell(a.box + f * 9);
    const double4 pi = P[i];
    for (uint64_t jb = 0; jb < a.N; jb += 64) {
        const uint64_t j = jb + lane;
        bool ok = false;
        uint32_t tj = 0;
        double d2 = 0.0;
        if (j < a.N && j != i) {
            const double4 pj = P[j];
            double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
            min_image_ref<TRI>(x0, x1, x2, c);
            d2 = d2_of(x0, x1, x2);
            tj = (uint32_t)a.type_slot[j];
            ok = d2 <= s.cut2[tj];
        }
        for (uint32_t t = 0; t < a.nt; ++t) {
            const unsigned long long mask = __ballot(ok && tj == t);
            if (!mask) continue;
            const uint32_t base = (uint32_t)cnt[t], add = (uint32_t)__popcll(mask);
            if (ok && tj == t) {
                const uint32_t pos = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
                if (pos < s.nn[t]) {
                    ru[s.off[t] + pos] = sqrt(d2);
                    ju[s.off[t] + pos] = (int)j;
                }
            }
            wave_sync_lds();
            if (lane == 0) {
                cnt[t] = (int)(base + add);
                if (base + add > s.nn[t]) atomicOr(s.overflow, 1);
            }
            wave_sync_lds();
        }
    }
    // neighbour.cpp:19-44 meets the pairs of atom i in increasing order of
    // the partner's index and throws when a pair within either type's cutoff
    // finds i's list for the partner's type already full.  Beyond a list that
    // overflows (above), that happens when the list of type t is exactly full
    // and an atom j of type t with cut2[t] < d2 <= cut2[type i] comes after
    // the list's last member in index order.  Only cutoffs that differ by
    // type can give it, so this pass is rare.
    {
        const uint32_t ti = (uint32_t)a.type_slot[i];
        for (uint32_t t = 0; t < a.nt; ++t) {
            if (!((uint32_t)cnt[t] == s.nn[t] && s.cut2[ti] > s.cut2[t])) continue;  // uniform
            int last = -1;
            for (uint32_t k = lane; k < s.nn[t]; k += 64) last = max(last, s.orig[ju[s.off[t] + k]]);
            for (int o = 32; o > 0; o >>= 1) last = max(last, __shfl_xor(last, o, 64));
            bool late = false;
            for (uint64_t jb = 0; jb < a.N; jb += 64) {
                const uint64_t j = jb + lane;
                if (j < a.N && j != i && (uint32_t)a.type_slot[j] == t && s.orig[j] > last) {
                    const double4 pj = P[j];
                    double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                    min_image_ref<TRI>(x0, x1, x2, c);
                    const double d2 = d2_of(x0, x1, x2);
                    late = late || (d2 > s.cut2[t] && d2 <= s.cut2[ti]);
                }
            }
            if (__ballot(late) && lane == 0) atomicOr(s.overflow, 1);
        }
    }
    double qx0 = 0.0, qx1 = 0.0, qx2 = 0.0;
    uint32_t qcode = 0, qn = 0;
    for (uint32_t t = 0; t < a.nt; ++t) {
        const uint32_t n = min((uint32_t)cnt[t], s.nn[t]);
        const double *r = ru + s.off[t];
        const int *jj = ju + s.off[t];
        for (uint32_t k = lane; k < n; k += 64) {
            const double rk = r[k];
            const int jk = jj[k];
            uint32_t rank = 0;
            for (uint32_t q = 0; q < n; ++q) rank += (r[q] < rk || (r[q] == rk && jj[q] < jk)) ? 1u : 0u;
            rs[rank] = rk;
            js[rank] = jk;
        }
        wave_sync_lds();
        uint32_t ms = 0;  // neighbour.cpp:85-101
        if (n >= 3) {
            ms = 3;
            double sum = rs[0] + rs[1] + rs[2];
            while (ms < n) {
                if (sum <= rs[ms] * (double)(ms - 2)) break;
                sum += rs[ms];
                ms++;
            }
        }
        if (s.nbr) {
            int32_t *o = s.nbr + ((uint64_t)fl * a.N + (uint64_t)a.aprime[i]) * s.total + s.off[t];
            for (uint32_t k = lane; k < ms; k += 64) o[k] = a.aprime[js[k]];
        }
        for (uint32_t c0 = 0; c0 < ms; c0 += 64) {
            qn = min(64u, ms - c0);
            if (lane < qn) {
                const double4 pj = P[js[c0 + lane]];
                qx0 = pi.x - pj.x;
                qx1 = pi.y - pj.y;
                qx2 = pi.z - pj.z;
                min_image_ref<TRI>(qx0, qx1, qx2, c);
                qcode = t << 16;
            }
            if (lane < qn) direction(qx0, qx1, qx2);
            while (qn > 0)
                expand_queue(min(qn, nc), qx0, qx1, qx2, qcode, qn, acc, stage, 1, L2, lane, slot, m,
                             (int)a.lmax, cf);
        }
        wave_sync_lds();  // rs, js are reused by the next type
    }
    wave_sync_lds();
    store_density(a, acc, fl, i, 0, 1, lane);
    // bins 1.. hold nothing with a neighbour list
    double *Sf = a.S + (uint64_t)fl * a.FS;
    const uint64_t ap = (uint64_t)a.aprime[i];
    for (uint32_t t = 0; t < a.nt; ++t)
        for (uint32_t b = 1; b < a.nbin; ++b)
            for (uint32_t e = lane; e < L2; e += 64) {
                uint32_t l = (uint32_t)sqrtf((float)e);
                while (l * l > e) --l;
                while ((l + 1) * (l + 1) <= e) ++l;
                Sf[((uint64_t)(t * a.nbin + b) * L2 + l * l) * a.N + ap * (2 * l + 1) + (e - l * l)] = 0.0;
            }
}

struct SphGramArgs {
    const double *S;          // chunk densities [nfr][FS]
    double *G;                // [nprob][n0c][leff]
    const uint32_t *tstart;   // [nt+1] first type-sorted index of each type
    uint64_t N, FS;
    uint32_t nt, nbin, lmax;
    uint32_t n0c, skip, leff, nfr;
    uint32_t ncol;            // row sub-blocks per column block
};

// Banded Gram tiles on the matrix cores (v_mfma_f64_16x16x4f64), by column
// block: a workgroup takes CB = 64 consecutive frames t2 of one problem and
// 32 of the t1 rows whose lag band [t1, t1 + leff) meets them, so each frame
// of the chunk is read by ceil(rows / 32) workgroups and a t1 row by the
// (leff + 63) / 64 column blocks it spans, instead of every row tile reading
// every column of its band.  Four wavefronts each compute 16 rows x 32
// columns: per step of 4 K values a lane feeds one A value and two B values
// (A[row l%16][k l/16], B[k l/16][col l%16]) and the results land as
// D[l/16 + 4i][l%16].  K is staged through LDS 32 values at a time,
// row-major with one pad double per row, the next stage's values loaded into
// registers during this stage's MFMAs.  Every entry's K sum runs in the
// same order whatever the tiling, so chunking the t1 rows changes nothing.
constexpr int GCB = 64;   // frames per column block
constexpr int GRB = 32;   // t1 rows per workgroup
constexpr int GKM = 32;   // K values per LDS stage
typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void sph_gram_mfma_kernel(SphGramArgs a) {
    __shared__ double As[GRB][GKM + 1];
    __shared__ double Bs[GCB][GKM + 1];
    const uint32_t L1 = a.lmax + 1, L2 = L1 * L1;
    uint32_t y = blockIdx.y;
    const uint32_t l = y % L1;
    y /= L1;
    const uint32_t bin = y % a.nbin;
    y /= a.nbin;
    const uint32_t jt = y % a.nt, ti = y / a.nt;
    const uint32_t cb = blockIdx.x / a.ncol, rs = blockIdx.x - cb * a.ncol;  // ncol: row sub-blocks per column block
    const uint64_t c0 = (uint64_t)cb * GCB;
    if (c0 >= a.nfr) return;  // the whole workgroup
    // rows e whose band meets [c0, c0 + GCB): e skip <= c0 + GCB - 1 and e skip + leff - 1 >= c0
    const uint64_t e_lo = c0 + 1 > a.leff ? (c0 + 1 - a.leff + a.skip - 1) / a.skip : 0;
    const uint64_t e_hi = std::min<uint64_t>(a.n0c - 1, (c0 + GCB - 1) / a.skip);
    const uint64_t e0 = e_lo + (uint64_t)rs * GRB;
    if (e0 > e_hi) return;

    const uint32_t w = 2 * l + 1;
    const uint64_t K = (uint64_t)(a.tstart[ti + 1] - a.tstart[ti]) * w;
    const uint64_t base = ((uint64_t)(jt * a.nbin + bin) * L2 + l * l) * a.N + (uint64_t)a.tstart[ti] * w;

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // loaders: A row tid / 8, 4 K values; B frame tid / 4, 8 K values
    const uint32_t ar = tid >> 3, ak = (tid & 7u) * 4u;
    const uint32_t bf = tid >> 2, bk = (tid & 3u) * 8u;
    const uint64_t ea = e0 + ar;
    const bool va = ea <= e_hi;
    const uint64_t fb = c0 + bf;
    const bool vb = fb < a.nfr;
    const double *Ap = a.S + (va ? ea * a.skip : 0) * a.FS + base;
    const double *Bp = a.S + (vb ? fb : 0) * a.FS + base;
    // compute: wavefront (wr, wc) holds rows 16 wr.., columns 32 wc..
    const uint32_t wr = wave >> 1, wc = wave & 1u;
    const uint32_t lr = lane & 15u, lk = lane >> 4;

    f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    // the next stage's K values are loaded into registers while this stage's
    // MFMAs run
    double ra[4], rb[8];
    auto load = [&](uint64_t kc) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t k = kc + ak + u;
            ra[u] = (va && k < K) ? Ap[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t k = kc + bk + u;
            rb[u] = (vb && k < K) ? Bp[k] : 0.0;
        }
    };
    load(0);
    for (uint64_t kc = 0; kc < K; kc += GKM) {
        __syncthreads();  // the previous stage's reads are done
#pragma unroll
        for (int u = 0; u < 4; ++u) As[ar][ak + u] = ra[u];
#pragma unroll
        for (int u = 0; u < 8; ++u) Bs[bf][bk + u] = rb[u];
        __syncthreads();
        if (kc + GKM < K) load(kc + GKM);
#pragma unroll
        for (int kk = 0; kk < GKM; kk += 4) {
            const double av = As[16 * wr + lr][kk + lk];
            const double b0 = Bs[32 * wc + lr][kk + lk];
            const double b1 = Bs[32 * wc + 16 + lr][kk + lk];
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b1, acc1, 0, 0, 0);
        }
    }
    double *Gp = a.G + (uint64_t)blockIdx.y * a.n0c * a.leff;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint64_t e = e0 + 16 * wr + lk + 4 * i;
        if (e > e_hi) continue;
        const uint64_t t1 = e * a.skip;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint64_t t2 = c0 + 32 * wc + 16 * h + lr;
            if (t2 >= a.nfr || t2 < t1 || t2 - t1 >= a.leff) continue;
            Gp[e * a.leff + (t2 - t1)] = h ? acc1[i] : acc0[i];
        }
    }
}

// R[p][dt] += sum over the chunk's t1 (in order) of G[p][t1][dt]
__global__ __launch_bounds__(256) void sph_reduce_kernel(const double *G, double *R, uint32_t nprob, uint32_t n0c,
                                                         uint32_t leff) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (uint64_t)nprob * leff) return;
    const uint64_t p = k / leff, dt = k - p * leff;
    const double *g = G + p * n0c * leff + dt;
    double s = 0.0;
    for (uint32_t e = 0; e < n0c; ++e) s += g[(uint64_t)e * leff];
    R[k] = R[k] + s;
}

constexpr uint32_t STEIN_MAXL = 16;  // l values of one histogram

struct SteinArgs {
    const double *S;             // chunk densities [nfr][FS]
    const int32_t *type_ap;      // [N] type of each atom in type-sorted order
    const int32_t *orig_ap;      // [N] original index of each atom in type-sorted order
    unsigned long long *hist;    // [nbin][nt][nt][hs] counts, or null
    double *q;                   // [nbin][nfs][N][nt][lmax] q_l, l = 1..lmax, or null
    uint64_t N, FS, hs;
    uint32_t nt, nbin, lmax, nl, nbin_st, nfr, fidx0, nfs;
    uint32_t lmask;              // bit l set: q_l is needed
    double c00;                  // Y_00 of any direction: the density's l = 0 entry is (count) * c00
    uint32_t l_list[STEIN_MAXL];
    const int32_t *nbr;          // averaged order parameter: [nfr][N][nbr_stride] SANN neighbours, or null
    uint32_t nbr_stride;
    uint32_t nbr_off[SANN_MAXT];  // first neighbour of each type
};

// steinhardt.cpp:166-225: one thread per (frame of the chunk, type_j, bin,
// atom), atoms fastest.  q_l = sqrt(sum_m S_lm^2 / n / n * 4 pi / (2l+1)), the
// m sum in the reference's order (m = 0, -1..-l, 1..l), n the neighbour count
// recovered exactly from the l = 0 entry
__global__ __launch_bounds__(256) void sph_steinhardt_kernel(SteinArgs a) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (uint64_t)a.nfr * a.nt * a.nbin * a.N) return;
    uint64_t r = k;
    const uint64_t ap = r % a.N;
    r /= a.N;
    const uint32_t b = (uint32_t)(r % a.nbin);
    r /= a.nbin;
    const uint32_t jt = (uint32_t)(r % a.nt), fl = (uint32_t)(r / a.nt);
    const uint32_t L1 = a.lmax + 1, L2 = L1 * L1;
    const double *Sb = a.S + (uint64_t)fl * a.FS + (uint64_t)(jt * a.nbin + b) * L2 * a.N;
    const double na = rint(Sb[ap] / a.c00);
    const double pi = 3.14159265358979323846264338327950288419716939937510;
    const uint32_t ti = (uint32_t)a.type_ap[ap];
    // averaged order parameter (steinhardt.cpp:128-173): qbar = S(i)/n(i) for
    // the atom's own type, plus S(k)/n(k) of its SANN neighbours k in order
    const int32_t *nb = a.nbr ? a.nbr + ((uint64_t)fl * a.N + ap) * a.nbr_stride + a.nbr_off[jt] : nullptr;
    const uint32_t nk = nb ? min((uint32_t)na, a.nbr_stride - a.nbr_off[jt]) : 0;
    double q[ANL_SPH_LMAX_MAX + 1];
#pragma unroll
    for (int l = 0; l <= ANL_SPH_LMAX_MAX; ++l) {
        q[l] = 0.0;
        if (l <= (int)a.lmax && na > 0 && ((a.lmask >> l) & 1u)) {
            const uint64_t lb = (uint64_t)(l * l) * a.N;
            const double *s = Sb + lb + ap * (2 * l + 1);
            double v;
            if (!nb) {
                v = s[l] * s[l];
                for (int mm = 1; mm <= l; ++mm) v += s[l - mm] * s[l - mm];
                for (int mm = 1; mm <= l; ++mm) v += s[l + mm] * s[l + mm];
            } else {
                // qbar of m = 0, then m = -1..-l, then m = 1..l, squared and summed in that order
                v = 0.0;
                for (int o = 0; o <= 2 * l; ++o) {
                    const int mi = o == 0 ? l : (o <= l ? l - o : o);  // index of (l, m) in the m-run
                    double qb = ti == jt ? s[mi] / na : 0.0;
                    for (uint32_t k = 0; k < nk; ++k) {
                        const uint64_t b2 = (uint64_t)nb[k];
                        const double nn2 = rint(Sb[b2] / a.c00);
                        qb += Sb[lb + b2 * (2 * l + 1) + mi] / nn2;
                    }
                    v = o == 0 ? qb * qb : v + qb * qb;
                }
            }
            q[l] = sqrt(v / na / na * 4 * pi / (2 * l + 1));
        }
    }
    if (a.hist) {
        if (!(na > 0)) return;
        uint64_t hidx = 0, stride = a.hs;
        for (uint32_t kk = 0; kk < a.nl; ++kk) {
            double qk = 0.0;
#pragma unroll
            for (int l = 0; l <= ANL_SPH_LMAX_MAX; ++l)
                if ((uint32_t)l == a.l_list[kk]) qk = q[l];
            stride /= a.nbin_st;
            // beyond the last cell, and NaN (a neighbour without neighbours of
            // type_j in the averaged form), go to the last cell, as the size_t
            // conversion and clamp of steinhardt.cpp:199-202 do on x86-64
            const float x = floorf((float)(qk * a.nbin_st));
            const uint64_t jh = x < (float)a.nbin_st ? (uint64_t)x : a.nbin_st - 1;
            hidx += jh * stride;
        }
        // the lanes of a wavefront that hit the same cell add once, with their
        // number (q concentrates in few cells, where atomics would queue)
        const unsigned long long key = ((uint64_t)(b * a.nt + ti) * a.nt + jt) * a.hs + hidx;
        const int lane = (int)(threadIdx.x & 63);
        unsigned long long left = __ballot(1);
        while (left) {
            const int leader = __ffsll(left) - 1;
            const unsigned long long lk = __shfl(key, leader);
            const unsigned long long same = __ballot(key == lk);
            if (lane == leader) atomicAdd(&a.hist[lk], (unsigned long long)__popcll(same));
            left &= ~same;
        }
    } else {
        double *o = a.q + ((((uint64_t)b * a.nfs + a.fidx0 + fl) * a.N + (uint64_t)a.orig_ap[ap]) * a.nt + jt) * a.lmax;
#pragma unroll
        for (int l = 1; l <= ANL_SPH_LMAX_MAX; ++l)
            if (l <= (int)a.lmax) o[l - 1] = q[l];
    }
}

__global__ __launch_bounds__(256) void sph_zero_kernel(double *p, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

// specialfunctions.h:33-49: sqrt((2l+1)/(4 pi)) for m = 0, else
// sqrt((2l+1)/(4 pi) * 2 / v) (-1)^m, v = (l-m+1)(l-m+2)...(l+m) multiplied
// up in that order from 1.
static double sh_coeff(int l, int m) {
    const double pi = 3.14159265358979323846264338327950288419716939937510;
    if (m == 0) return std::sqrt((double)(2 * l + 1) / (4 * pi));
    double value = 1.0;
    for (int v = l - m + 1; v <= l + m; v++) value = value * v;
    return std::sqrt((double)(2 * l + 1) / (4 * pi) * 2 / value) * std::pow(-1.0, m);
}

}  // namespace anl

using namespace anl;

struct anl_sph {
    anl_traj *tr = nullptr;
    uint32_t lmax = 0, nbin = 0, nt = 0;
    std::vector<double> rminmax;      // [nt*nt][2]
    std::vector<int32_t> aprime;      // [N] type-sorted index of each slot
    std::vector<uint32_t> tstart;     // [nt+1]
    double *d_thr = nullptr;          // [nt*nt][nbin+1]
    double *d_coef = nullptr;         // tables [(l'+1)^2] for l' = 0..lmax, in that order (coef_offset)
    int32_t *d_aprime = nullptr;
    uint32_t *d_tstart = nullptr;
    int32_t *d_type_ap = nullptr;     // [N] type of each type-sorted atom
    int32_t *d_orig_ap = nullptr;     // [N] original index of each type-sorted atom
    double c00 = 0.0;                 // Y_00
    uint64_t budget = 0;              // bytes of densities held at once (set at creation)
    double d2lo = 0.0, d2hi = 0.0;    // min over type pairs of T[0], max of T[nbin]
    double rcut = 0.0;                // distance bound of the block screen
    bool block_screen = false;        // orthorhombic cell and finite rcut
    bool sann = false;                // neighbour lists (SANN) instead of the radial ranges
    SannArgs sann_args{};
    int *d_overflow = nullptr;        // set by the list kernel when a list is full
    void *work = nullptr;             // densities, G, R (grow-only)
    size_t work_cap = 0;
};

namespace {

// doubles of one frame's densities expanded up to l = lm (default lmax)
uint64_t frame_doubles(const anl_sph *h, uint32_t lm = ~0u) {
    if (lm > h->lmax) lm = h->lmax;
    const uint64_t L2 = (uint64_t)(lm + 1) * (lm + 1);
    return h->tr->N * h->nt * h->nbin * L2;
}

// first coefficient of the table for expansions up to l = lm
uint64_t coef_offset(uint32_t lm) {
    uint64_t o = 0;
    for (uint32_t k = 0; k < lm; k++) o += (uint64_t)(k + 1) * (k + 1);
    return o;
}

uint64_t a256(uint64_t b) { return (b + 255) / 256 * 256; }

void *workspace(anl_sph *h, size_t bytes) {
    if (bytes <= h->work_cap) return h->work;
    ANL_HIP(hipDeviceSynchronize());
    if (h->work) ANL_HIP(hipFree(h->work));
    h->work = nullptr;
    h->work_cap = 0;
    ANL_HIP(hipMalloc(&h->work, bytes));
    h->work_cap = bytes;
    return h->work;
}

// bins per density launch column: a wavefront's density block for them plus
// the staging of 64/(lmax+1) harmonic vectors fits its share of LDS
uint32_t bins_per_launch(uint32_t nt, uint32_t nbin, uint32_t lmax) {
    const uint32_t per_bin = nt * (lmax + 1) * (lmax + 1);
    const uint32_t room = SPH_WAVE_LDS - stage_doubles(lmax);
    return std::min(nbin, room / per_bin);
}

// densities of frames [f0, f0 + nfr) of the device trajectory into S
// bytes of block bounding spheres for nfr frames (none for triclinic cells)
uint64_t radius_bytes(const anl_sph *h, uint64_t nfr) {
    return h->block_screen ? nfr * ((h->tr->N + 63) / 64) * sizeof(double4) : 0;
}

// densities of frames f0 + k * fstride, k < nfr, of the device trajectory into
// S, expanded up to l = lm (default lmax; the values of l <= lm do not depend
// on it); rad holds radius_bytes(h, nfr)
void launch_density(anl_sph *h, uint64_t f0, uint64_t fstride, uint32_t nfr, double *S, double4 *rad,
                    hipStream_t s, int32_t *nbr = nullptr, uint32_t lm = ~0u) {
    if (lm > h->lmax) lm = h->lmax;
    SphDensArgs a{};
    a.fstride = fstride;
    if (h->sann) {
        a.pos = h->tr->pos;
        a.box = h->tr->box;
        a.type_slot = h->tr->type_slot;
        a.aprime = h->d_aprime;
        a.coef = h->d_coef + coef_offset(lm);
        a.S = S;
        a.N = h->tr->N;
        a.FS = frame_doubles(h, lm);
        a.f0 = f0;
        a.nt = h->nt;
        a.nbin = h->nbin;
        a.lmax = lm;
        a.nbc = 1;
        const uint64_t gx = (uint64_t)nfr * a.N;
        if (gx > 0x7fffffffull) fail(ANL_ERR_ARG, "too many (frame, atom) pairs in one density launch");
        SannArgs sa = h->sann_args;
        sa.nbr = nbr;  // the chosen neighbours, for the averaged order parameter
        const size_t lds = (size_t)(h->nt * (lm + 1) * (lm + 1) + stage_doubles(lm) + sa.total +
                                    sa.capmax) * sizeof(double) +
                           (size_t)(sa.total + sa.capmax + SANN_MAXT) * sizeof(int);
        ANL_HIP(hipMemsetAsync(h->d_overflow, 0, sizeof(int), s));
        if (h->tr->triclinic)
            hipLaunchKernelGGL(sph_density_sann_kernel<true>, dim3((uint32_t)gx), dim3(64), lds, s, a, sa);
        else
            hipLaunchKernelGGL(sph_density_sann_kernel<false>, dim3((uint32_t)gx), dim3(64), lds, s, a, sa);
        ANL_HIP(hipGetLastError());
        int over = 0;
        ANL_HIP(hipMemcpyAsync(&over, h->d_overflow, sizeof(int), hipMemcpyDeviceToHost, s));
        ANL_HIP(hipStreamSynchronize(s));
        if (over) fail(ANL_ERR_ARG, "Too many neighbours in shell!");  // neighbour.cpp:44-48
        return;
    }
    a.d2lo = h->d2lo;
    a.d2hi = h->d2hi;
    a.rcut = h->rcut;
    a.block_rad = h->block_screen ? rad : nullptr;
    if (a.block_rad) {
        const uint64_t nb = (uint64_t)nfr * ((h->tr->N + 63) / 64);
        if (nb > 0x7fffffffull) fail(ANL_ERR_ARG, "too many (frame, atom block) pairs in one launch");
        hipLaunchKernelGGL(sph_block_radius_kernel, dim3((uint32_t)nb), dim3(64), 0, s, h->tr->pos, h->tr->box,
                           h->tr->N, f0, fstride, rad);
        ANL_HIP(hipGetLastError());
    }
    a.pos = h->tr->pos;
    a.box = h->tr->box;
    a.type_slot = h->tr->type_slot;
    a.aprime = h->d_aprime;
    a.thr = h->d_thr;
    a.coef = h->d_coef + coef_offset(lm);
    a.S = S;
    a.N = h->tr->N;
    a.FS = frame_doubles(h, lm);
    a.f0 = f0;
    a.nt = h->nt;
    a.nbin = h->nbin;
    a.lmax = lm;
    a.nbc = bins_per_launch(h->nt, h->nbin, lm);
    a.ngroups = (uint32_t)((a.N + SPH_WAVES - 1) / SPH_WAVES);
    const uint32_t accn = a.nt * a.nbc * (a.lmax + 1) * (a.lmax + 1);
    const size_t lds = (size_t)SPH_WAVES * (accn + stage_doubles(a.lmax)) * sizeof(double);
    const uint64_t gx = (uint64_t)nfr * a.ngroups;
    if (gx > 0x7fffffffull) fail(ANL_ERR_ARG, "too many (frame, atom) pairs in one density launch");
    if ((h->nbin + a.nbc - 1) / a.nbc > 65535) fail(ANL_ERR_ARG, "too many distance bins");
    dim3 grid((uint32_t)gx, (h->nbin + a.nbc - 1) / a.nbc);
    if (h->tr->triclinic)
        hipLaunchKernelGGL(sph_density_kernel<true>, grid, dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL(sph_density_kernel<false>, grid, dim3(256), lds, s, a);
    ANL_HIP(hipGetLastError());
}

void sph_calculate(anl_sph *h, uint64_t primo, uint32_t n, uint32_t tmax, uint32_t skip, double *out_host) {
    anl_traj *tr = h->tr;
    if (skip == 0) skip = 1;
    const uint32_t leff = (n < tmax || tmax == 0) ? n : tmax;
    if ((uint64_t)leff + n + primo > tr->F)
        fail(ANL_ERR_TOO_SHORT,
             "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
             "the size of the average or the lenght of the correlation function");
    const uint32_t L1 = h->lmax + 1, nt = h->nt;
    const size_t fin = (size_t)nt * nt * h->nbin * L1;
    if (n == 0 || leff == 0) return;
    const uint64_t n0 = (n + skip - 1) / skip;  // t1 = 0, skip, .. < n
    const uint64_t fbytes = frame_doubles(h) * sizeof(double);
    const uint64_t maxfr = std::max<uint64_t>(1, h->budget / std::max<uint64_t>(fbytes, 1));
    if (leff > maxfr) fail(ANL_ERR_ARG, "the densities of one lag window do not fit the density budget");
    // t1 values per chunk: as many as the budget holds, spread evenly over
    // the chunks, in whole row tiles where that still fits
    const uint64_t rc_max = std::min<uint64_t>(n0, (maxfr - leff) / skip + 1);
    const uint64_t nchunks = (n0 + rc_max - 1) / rc_max;
    uint64_t rc = (n0 + nchunks - 1) / nchunks;
    if (rc > (uint64_t)GT) rc = std::min<uint64_t>(rc_max, (rc + GT - 1) / GT * GT);
    const uint32_t nprob = (uint32_t)fin;
    if (nprob > 65535) fail(ANL_ERR_ARG, "too many (type, type, bin, l) channels for one launch (at most 65535)");
    const uint64_t nfr_max = (rc - 1) * skip + leff;
    const uint64_t offS = 0, offG = a256(nfr_max * fbytes), offR = offG + a256((uint64_t)nprob * rc * leff * 8);
    const uint64_t offQ = offR + a256((uint64_t)nprob * leff * 8);
    char *w = (char *)workspace(h, offQ + a256(radius_bytes(h, nfr_max)));
    double *S = (double *)(w + offS), *G = (double *)(w + offG), *R = (double *)(w + offR);
    double4 *rad = (double4 *)(w + offQ);
    hipStream_t s = nullptr;
    const uint64_t nR = (uint64_t)nprob * leff;
    hipLaunchKernelGGL(sph_zero_kernel, dim3((uint32_t)((nR + 255) / 256)), dim3(256), 0, s, R, nR);
    ANL_HIP(hipGetLastError());
    for (uint64_t e0 = 0; e0 < n0; e0 += rc) {
        const uint32_t n0c = (uint32_t)std::min<uint64_t>(rc, n0 - e0);
        const uint64_t t1_first = e0 * skip;
        const uint32_t nfr = (uint32_t)((uint64_t)(n0c - 1) * skip + leff);
        launch_density(h, primo + t1_first, 1, nfr, S, rad, s);
        SphGramArgs g{};
        g.S = S;
        g.G = G;
        g.tstart = h->d_tstart;
        g.N = tr->N;
        g.FS = frame_doubles(h);
        g.nt = nt;
        g.nbin = h->nbin;
        g.lmax = h->lmax;
        g.n0c = n0c;
        g.skip = skip;
        g.leff = leff;
        g.nfr = nfr;
        // column blocks of 64 frames, each with the row sub-blocks of 32 t1
        // values whose lag band meets it
        const uint64_t rows_max = std::min<uint64_t>(n0c, (uint64_t)(GCB - 1 + leff - 1) / skip + 1);
        g.ncol = (uint32_t)((rows_max + GRB - 1) / GRB);
        const uint64_t ncb = (nfr + GCB - 1) / GCB;
        if (ncb * g.ncol > 0x7fffffffull) fail(ANL_ERR_ARG, "too many Gram tiles in one launch");
        hipLaunchKernelGGL(sph_gram_mfma_kernel, dim3((uint32_t)(ncb * g.ncol), nprob), dim3(256), 0, s, g);
        ANL_HIP(hipGetLastError());
        hipLaunchKernelGGL(sph_reduce_kernel, dim3((uint32_t)((nR + 255) / 256)), dim3(256), 0, s, G, R, nprob,
                           n0c, leff);
        ANL_HIP(hipGetLastError());
    }
    std::vector<double> Rh(nR);
    ANL_HIP(hipMemcpy(Rh.data(), R, nR * sizeof(double), hipMemcpyDeviceToHost));
    // R[((ti*nt + tj)*nbin + b)*L1 + l][dt] -> out[dt][ti][tj][b][lmax - l]
    for (uint32_t dt = 0; dt < leff; dt++)
        for (uint32_t p = 0; p < nprob; p++) {
            const uint32_t l = p % L1, q = p / L1, ti = q / (nt * h->nbin);
            const double na = (double)(h->tstart[ti + 1] - h->tstart[ti]);
            out_host[(size_t)dt * fin + (size_t)q * L1 + (h->lmax - l)] = Rh[(size_t)p * leff + dt] / na / (double)n0;
        }
}

// steinhardt.cpp:47-236 (no neighbour list, not averaged): frames primo + i,
// i = 0, skip, .. < n - n % skip, in chunks of frames that fit the budget
void steinhardt_calculate(anl_sph *h, uint64_t primo, uint32_t n, uint32_t skip, const uint32_t *l_list, uint32_t nl,
                          uint32_t nbin_st, int do_hist, int averaged, double *out_host) {
    anl_traj *tr = h->tr;
    if (skip == 0) skip = 1;
    if (averaged && !h->sann)  // steinhardt.cpp:30-32
        fail(ANL_ERR_ARG, "You must use a neighbour list when computing a averaged steinhardt order parameter! ");
    if ((uint64_t)n + primo > tr->F)
        fail(ANL_ERR_TOO_SHORT,
             "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
             "the size of the average ");
    if (nl > STEIN_MAXL) fail(ANL_ERR_ARG, "at most 16 l values per histogram");
    for (uint32_t k = 0; k < nl; k++)
        if (l_list[k] > h->lmax) fail(ANL_ERR_ARG, "Wrong value of l!");
    if (nl > 0 && nbin_st == 0) fail(ANL_ERR_ARG, "the histogram size must be positive");
    uint64_t hs = 1;
    for (uint32_t k = 0; k < nl; k++) {
        hs *= nbin_st;
        if (hs > (1ull << 31)) fail(ANL_ERR_ARG, "the histogram has more than 2^31 cells per (bin, type, type)");
    }
    const uint32_t nt = h->nt, nbin = h->nbin;
    const uint64_t N = tr->N, nfs = n / skip;
    const uint64_t len = do_hist ? (uint64_t)nbin * nt * nt * hs : (uint64_t)nbin * nfs * N * nt * h->lmax;
    std::fill(out_host, out_host + len, 0.0);
    if (nfs == 0 || len == 0) return;
    // the histogram reads the listed l only (and l = 0 for the neighbour
    // count): the densities are expanded up to the largest of them
    uint32_t lm = h->lmax;
    if (do_hist) {
        lm = 0;
        for (uint32_t k = 0; k < nl; k++) lm = std::max(lm, l_list[k]);
    }
    const uint64_t fbytes = frame_doubles(h, lm) * sizeof(double);
    const uint64_t fpc = std::min<uint64_t>(nfs, std::max<uint64_t>(1, h->budget / std::max<uint64_t>(fbytes, 1)));
    const uint64_t nbr_bytes = averaged ? fpc * N * h->sann_args.total * sizeof(int32_t) : 0;
    const uint64_t offR = a256(fpc * fbytes), offO = offR + a256(radius_bytes(h, fpc));
    const uint64_t offN = offO + a256(len * 8);
    char *w = (char *)workspace(h, offN + a256(nbr_bytes));
    double *S = (double *)w;
    double4 *rad = (double4 *)(w + offR);
    double *outd = (double *)(w + offO);
    int32_t *nbr = averaged ? (int32_t *)(w + offN) : nullptr;
    hipStream_t s = nullptr;
    hipLaunchKernelGGL(sph_zero_kernel, dim3((uint32_t)((len + 255) / 256)), dim3(256), 0, s, outd, len);
    ANL_HIP(hipGetLastError());
    SteinArgs a{};
    a.S = S;
    a.type_ap = h->d_type_ap;
    a.orig_ap = h->d_orig_ap;
    a.hist = do_hist ? (unsigned long long *)outd : nullptr;  // zero bits are the integer 0
    a.q = do_hist ? nullptr : outd;
    a.N = N;
    a.FS = frame_doubles(h, lm);
    a.hs = hs;
    a.nt = nt;
    a.nbin = nbin;
    a.lmax = lm;  // with the histogram off, lm = lmax: q_1..q_lmax are written
    a.lmask = do_hist ? 0u : ~0u;
    if (do_hist)
        for (uint32_t k = 0; k < nl; k++) a.lmask |= 1u << l_list[k];
    a.nl = nl;
    a.nbin_st = nbin_st;
    a.nfs = (uint32_t)nfs;
    a.c00 = h->c00;
    for (uint32_t k = 0; k < nl; k++) a.l_list[k] = l_list[k];
    a.nbr = nbr;
    a.nbr_stride = h->sann_args.total;
    for (uint32_t t = 0; t < nt && t < SANN_MAXT; t++) a.nbr_off[t] = h->sann_args.off[t];
    for (uint64_t f0 = 0; f0 < nfs; f0 += fpc) {
        const uint32_t nfc = (uint32_t)std::min<uint64_t>(fpc, nfs - f0);
        launch_density(h, primo + f0 * skip, skip, nfc, S, rad, s, nbr, lm);
        a.nfr = nfc;
        a.fidx0 = (uint32_t)f0;
        const uint64_t nth = (uint64_t)nfc * nt * nbin * N;
        if ((nth + 255) / 256 > 0x7fffffffull) fail(ANL_ERR_ARG, "too many (frame, atom) pairs in one launch");
        hipLaunchKernelGGL(sph_steinhardt_kernel, dim3((uint32_t)((nth + 255) / 256)), dim3(256), 0, s, a);
        ANL_HIP(hipGetLastError());
    }
    if (do_hist) {
        std::vector<unsigned long long> c(len);
        ANL_HIP(hipMemcpy(c.data(), outd, len * 8, hipMemcpyDeviceToHost));
        const double incr = 1.0 / (double)(int)nfs;  // steinhardt.cpp:92-93
        for (uint64_t k = 0; k < len; k++) out_host[k] = (double)c[k] * incr;
    } else {
        ANL_HIP(hipMemcpy(out_host, outd, len * 8, hipMemcpyDeviceToHost));
    }
}

}  // namespace

extern "C" {

int anl_sph_create(anl_traj *tr, uint32_t lmax, const double *rminmax, uint32_t nbin, anl_sph **out) {
    return guard([&] {
        if (!tr || !rminmax || !out) fail(ANL_ERR_ARG, "null pointer passed to anl_sph_create");
        if (lmax > ANL_SPH_LMAX_MAX) fail(ANL_ERR_ARG, "lmax must be at most 10");
        if (nbin == 0) fail(ANL_ERR_ARG, "nbin must be positive");
        const uint32_t nt = tr->nt;
        for (uint32_t p = 0; p < nt * nt; p++) {
            const double lo = rminmax[2 * p], hi = rminmax[2 * p + 1];
            if (!std::isfinite(lo) || !std::isfinite(hi) || !(hi > lo))
                fail(ANL_ERR_ARG, "every radial range must be finite with rmax > rmin");
        }
        if (bins_per_launch(nt, nbin, lmax) == 0)
            fail(ANL_ERR_ARG, "too many atom types for the per-atom density block");
        ensure_device();
        anl_sph *h = new anl_sph();
        h->tr = tr;
        h->lmax = lmax;
        h->nbin = nbin;
        h->nt = nt;
        h->rminmax.assign(rminmax, rminmax + 2 * nt * nt);
        std::vector<double> thr((size_t)nt * nt * (nbin + 1));
        for (uint32_t p = 0; p < nt * nt; p++) {
            const double lo = rminmax[2 * p], hi = rminmax[2 * p + 1];
            const double dr = (hi - lo) / (double)nbin;  // sphericalbase.h:43-47
            std::vector<double> T = bin_edges(lo, hi, nbin, dr);
            std::copy(T.begin(), T.end(), thr.begin() + (size_t)p * (nbin + 1));
        }
        h->d2lo = HUGE_VAL;
        h->d2hi = 0.0;
        for (uint32_t p = 0; p < nt * nt; p++) {
            h->d2lo = std::min(h->d2lo, thr[(size_t)p * (nbin + 1)]);
            h->d2hi = std::max(h->d2hi, thr[(size_t)p * (nbin + 1) + nbin]);
        }
        h->rcut = std::sqrt(h->d2hi) * (1.0 + 1e-12) + 1e-9;
        h->block_screen = !tr->triclinic && std::isfinite(h->rcut);
        // coefficient tables [l * (l'+1) + m] for expansions up to every l' <= lmax
        std::vector<double> coef(coef_offset(lmax + 1), 0.0);
        for (uint32_t lp = 0; lp <= lmax; lp++)
            for (uint32_t l = 0; l <= lp; l++)
                for (uint32_t m = 0; m <= l; m++) coef[coef_offset(lp) + l * (lp + 1) + m] = sh_coeff((int)l, (int)m);
        h->c00 = coef[0];
        h->aprime.assign(tr->N, 0);
        h->tstart.assign(nt + 1, 0);
        uint32_t k = 0;
        for (uint32_t t = 0; t < nt; t++) {
            h->tstart[t] = k;
            for (uint64_t sl = 0; sl < tr->N; sl++)
                if ((uint32_t)tr->type_of_slot[sl] == t) h->aprime[sl] = (int32_t)k++;
        }
        h->tstart[nt] = k;
        size_t free_b = 0, total_b = 0;
        ANL_HIP(hipMemGetInfo(&free_b, &total_b));
        h->budget = std::min<uint64_t>(64ull << 30, std::max<uint64_t>(free_b / 2, 1ull << 30));
        try {
            ANL_HIP(hipMalloc(&h->d_thr, thr.size() * sizeof(double)));
            ANL_HIP(hipMemcpy(h->d_thr, thr.data(), thr.size() * sizeof(double), hipMemcpyHostToDevice));
            ANL_HIP(hipMalloc(&h->d_coef, coef.size() * sizeof(double)));
            ANL_HIP(hipMemcpy(h->d_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
            ANL_HIP(hipMalloc(&h->d_aprime, tr->N * sizeof(int32_t)));
            ANL_HIP(hipMemcpy(h->d_aprime, h->aprime.data(), tr->N * sizeof(int32_t), hipMemcpyHostToDevice));
            ANL_HIP(hipMalloc(&h->d_tstart, (nt + 1) * sizeof(uint32_t)));
            ANL_HIP(hipMemcpy(h->d_tstart, h->tstart.data(), (nt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
            std::vector<int32_t> orig(tr->N), type_ap(tr->N), orig_ap(tr->N);
            ANL_HIP(hipMemcpy(orig.data(), tr->orig_slot, tr->N * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (uint64_t sl = 0; sl < tr->N; sl++) {
                type_ap[h->aprime[sl]] = tr->type_of_slot[sl];
                orig_ap[h->aprime[sl]] = orig[sl];
            }
            ANL_HIP(hipMalloc(&h->d_type_ap, tr->N * sizeof(int32_t)));
            ANL_HIP(hipMemcpy(h->d_type_ap, type_ap.data(), tr->N * sizeof(int32_t), hipMemcpyHostToDevice));
            ANL_HIP(hipMalloc(&h->d_orig_ap, tr->N * sizeof(int32_t)));
            ANL_HIP(hipMemcpy(h->d_orig_ap, orig_ap.data(), tr->N * sizeof(int32_t), hipMemcpyHostToDevice));
        } catch (...) {
            anl_sph_destroy(h);
            throw;
        }
        *out = h;
    });
}

int anl_sph_destroy(anl_sph *h) {
    if (!h) return ANL_OK;
    (void)hipDeviceSynchronize();
    if (h->d_thr) (void)hipFree(h->d_thr);
    if (h->d_coef) (void)hipFree(h->d_coef);
    if (h->d_aprime) (void)hipFree(h->d_aprime);
    if (h->d_tstart) (void)hipFree(h->d_tstart);
    if (h->d_type_ap) (void)hipFree(h->d_type_ap);
    if (h->d_orig_ap) (void)hipFree(h->d_orig_ap);
    if (h->d_overflow) (void)hipFree(h->d_overflow);
    if (h->work) (void)hipFree(h->work);
    delete h;
    return ANL_OK;
}

int anl_sph_set_neighbours(anl_sph *h, const uint32_t *nneigh, const double *cutoff2) {
    return guard([&] {
        if (!h || !nneigh || !cutoff2) fail(ANL_ERR_ARG, "null pointer passed to anl_sph_set_neighbours");
        if (h->nt > SANN_MAXT) fail(ANL_ERR_ARG, "neighbour lists support at most 16 atom types");
        SannArgs sa{};
        uint32_t tot = 0;
        for (uint32_t t = 0; t < h->nt; t++) {
            if (nneigh[t] == 0) fail(ANL_ERR_ARG, "every neighbour list needs a positive capacity");
            if (!std::isfinite(cutoff2[t])) fail(ANL_ERR_ARG, "neighbour cutoffs must be finite");
            sa.nn[t] = nneigh[t];
            sa.off[t] = tot;
            sa.cut2[t] = cutoff2[t];
            sa.capmax = std::max(sa.capmax, nneigh[t]);
            tot += nneigh[t];
            if (tot > SANN_MAXTOT) fail(ANL_ERR_ARG, "the neighbour lists hold at most 1024 atoms in total");
        }
        sa.total = (tot + 1) / 2 * 2;
        if (!h->d_overflow) ANL_HIP(hipMalloc(&h->d_overflow, sizeof(int)));
        sa.overflow = h->d_overflow;
        sa.orig = h->tr->orig_slot;
        h->sann_args = sa;
        h->sann = true;
    });
}

int anl_sph_set_budget(anl_sph *h, uint64_t bytes) {
    return guard([&] {
        if (!h) fail(ANL_ERR_ARG, "null pointer passed to anl_sph_set_budget");
        if (bytes == 0) fail(ANL_ERR_ARG, "the density budget must be positive");
        h->budget = bytes;
    });
}

int anl_sph_snapshot(anl_sph *h, uint64_t frame, double *out_host) {
    return guard([&] {
        if (!h || !out_host) fail(ANL_ERR_ARG, "null pointer passed to anl_sph_snapshot");
        if (frame >= h->tr->F) fail(ANL_ERR_ARG, "frame index out of range");
        const uint64_t FS = frame_doubles(h);
        char *w = (char *)workspace(h, a256(FS * sizeof(double)) + a256(radius_bytes(h, 1)));
        double *S = (double *)w;
        double4 *rad = (double4 *)(w + a256(FS * sizeof(double)));
        launch_density(h, frame, 1, 1, S, rad, nullptr);
        std::vector<double> Sh(FS);
        ANL_HIP(hipMemcpy(Sh.data(), S, FS * sizeof(double), hipMemcpyDeviceToHost));
        // S[tj][b][l][a'][m + l] -> out[atom][tj][b][reference (l, m) order]
        const uint32_t L1 = h->lmax + 1, L2 = L1 * L1, nt = h->nt, nbin = h->nbin;
        const uint64_t N = h->tr->N;
        std::vector<int32_t> orig(N);
        ANL_HIP(hipMemcpy(orig.data(), h->tr->orig_slot, N * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (uint64_t sl = 0; sl < N; sl++) {
            const uint64_t ap = (uint64_t)h->aprime[sl], atom = (uint64_t)orig[sl];
            for (uint32_t tj = 0; tj < nt; tj++)
                for (uint32_t b = 0; b < nbin; b++) {
                    double *dst = out_host + ((atom * nt + tj) * nbin + b) * L2;
                    for (uint32_t l = 0; l <= h->lmax; l++) {
                        const double *src = Sh.data() + ((uint64_t)(tj * nbin + b) * L2 + l * l) * N + ap * (2 * l + 1);
                        const uint32_t off = L2 - (l + 1) * (l + 1);  // l = lmax first
                        for (uint32_t m = 1; m <= l; m++) dst[off + m - 1] = src[l - m];
                        for (uint32_t m = 0; m <= l; m++) dst[off + l + m] = src[l + m];
                    }
                }
        }
    });
}

int anl_steinhardt_calculate(anl_sph *h, uint64_t primo, uint32_t n, uint32_t skip, const uint32_t *l_list,
                             uint32_t nl, uint32_t nbin_st, int do_histogram, int averaged, double *out_host) {
    return guard([&] {
        if (!h || !out_host || (nl > 0 && !l_list)) fail(ANL_ERR_ARG, "null pointer passed to anl_steinhardt_calculate");
        steinhardt_calculate(h, primo, n, skip, l_list, nl, nbin_st, do_histogram, averaged, out_host);
    });
}

int anl_sph_calculate(anl_sph *h, uint64_t primo, uint32_t n, uint32_t tmax, uint32_t skip, double *out_host) {
    return guard([&] {
        if (!h || !out_host) fail(ANL_ERR_ARG, "null pointer passed to anl_sph_calculate");
        sph_calculate(h, primo, n, tmax, skip, out_host);
    });
}

}  // extern "C"
