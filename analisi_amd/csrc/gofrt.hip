// g(r,t): van Hove histogram of the minimum-image distance between atom i at
// frame primo+t0 and atom j at frame primo+t0+t, for every lag t (step
// every), every t0 (step skip) and every ordered pair (i, j), i == j included
// (gofrt.cpp:96-122).  Counts are exact integers.
//
// Two arithmetic forms produce the same counts:
//  * FP64 plain  -- every pair goes through the reference's FP64 sequence:
//    x = xi - xj, iterative min image z->y->x with tilt carry
//    (basetrajectory.h:224-268), d2 = 0 + x^2 + y^2 + z^2 without FMA
//    (:200-219), gate rmin2 <= d2 <= rmax2, bin floorf((float)((sqrt(d2) -
//    rmin)/dr)) (gofrt.cpp:104-118).
//  * SCREENED   -- every pair is first evaluated in 16-bit fixed-point
//    fractional coordinates (packed 2-wide integer ops); only pairs whose
//    screened distance can be <= rmax (plus a proven margin) become
//    candidates, which are queued per wavefront and evaluated with the exact
//    FP64 sequence above.  Rejecting a pair in the screen is proven safe when
//    the cell is constant and rmax < half of every cell width.
//
// The bin is never computed with device sqrt/division: the host tabulates
// T[k] = smallest double d2 whose reference bin is >= k (the reference bin is a
// monotone step function of d2), so on the device
//   accept  <=>  lo <= d2 <= rmax2  and  d2 < T[nbin]
//   bin     =    max{k : T[k] <= d2}
// which reproduces the reference's counts bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

// ---------------------------------------------------------------------------
// host: exact bin edges
// ---------------------------------------------------------------------------

// gofrt.cpp:114-115, evaluated on the host with IEEE double sqrt/sub/div and
// the double->float round-to-nearest cast before floorf.
static int ref_bin(double d2, double rmin, double dr) {
    volatile double d = std::sqrt(d2);
    volatile double q = (d - rmin) / dr;
    float f = (float)q;
    return (int)std::floor(f);
}

static uint64_t dbits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}
static double bitsd(uint64_t u) {
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

// T[k] for k = 0..nbin: smallest non-negative double with ref_bin >= k, by
// bisection on the (monotone) bit patterns of non-negative doubles.
std::vector<double> bin_edges(double rmin, double rmax, uint32_t nbin, double dr) {
    std::vector<double> T(nbin + 1);
    double top = (std::fabs(rmax) + std::fabs(rmin) + 4.0 * std::fabs(dr) + 1.0);
    top = top * top;
    for (uint32_t k = 0; k <= nbin; k++) {
        if (ref_bin(0.0, rmin, dr) >= (int)k) {
            T[k] = 0.0;
            continue;
        }
        if (ref_bin(top, rmin, dr) < (int)k) {
            T[k] = HUGE_VAL;
            continue;
        }
        uint64_t lo = 0, hi = dbits(top);  // ref_bin(lo) < k <= ref_bin(hi)
        while (hi - lo > 1) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (ref_bin(bitsd(mid), rmin, dr) >= (int)k) hi = mid;
            else lo = mid;
        }
        T[k] = bitsd(hi);
    }
    return T;
}

// ---------------------------------------------------------------------------
// host: screen preconditions and threshold
// ---------------------------------------------------------------------------

// Smallest eigenvalue of the symmetric 3x3 metric G = A^T A (cyclic Jacobi).
static void eig_sym3(const double G[3][3], double *lmin, double *lmax) {
    double a[3][3];
    std::memcpy(a, G, sizeof(a));
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    *lmin = std::min(a[0][0], std::min(a[1][1], a[2][2]));
    *lmax = std::max(a[0][0], std::max(a[1][1], a[2][2]));
}

// The screen works on q = 16-bit wrapped differences of fractional
// coordinates.  It may reject a pair only if the reference cannot count it:
//  (1) constant cell, so both frames share one fractional frame;
//  (2) rmax < w_k (1/2 - 2^-15) for every cell width w_k: then any image with
//      d <= rmax is the unique shortest one, the reference's image equals it,
//      and its fractional components lie strictly inside (-1/2, 1/2), where
//      16-bit wrapping picks the same image;
//  (3) |A delta| >= sqrt(lambda_min(A^T A)) |delta|, and rounding each 16-bit
//      coordinate moves each component of q by at most 1 unit, so
//      d <= rmax  =>  |q| <= rmax 2^16 / sqrt(lambda_min) + sqrt(3).
static Screen make_screen(const Cell &c, bool box_const, uint32_t nt, double rmax) {
    Screen s{false, 0, "", 0.5, 0.0, 0.0, 0.0, 0.0};
    if (!box_const) {
        s.why_not = "the cell changes along the trajectory";
        return s;
    }
    if (nt > 255) {  // a candidate's type rides in 8 bits
        s.why_not = "more than 255 atom types";
        return s;
    }
    double A[3][3] = {{c.L[0], c.xy, c.xz}, {0, c.L[1], c.yz}, {0, 0, c.L[2]}};
    // inverse of the upper-triangular A
    double Ai[3][3] = {{0}};
    Ai[0][0] = 1 / A[0][0];
    Ai[1][1] = 1 / A[1][1];
    Ai[2][2] = 1 / A[2][2];
    Ai[0][1] = -A[0][1] / (A[0][0] * A[1][1]);
    Ai[1][2] = -A[1][2] / (A[1][1] * A[2][2]);
    Ai[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / (A[0][0] * A[1][1] * A[2][2]);
    double wmin = HUGE_VAL;
    for (int k = 0; k < 3; k++) {
        double n = std::sqrt(Ai[k][0] * Ai[k][0] + Ai[k][1] * Ai[k][1] + Ai[k][2] * Ai[k][2]);
        wmin = std::min(wmin, 1.0 / n);
    }
    double r = std::fabs(rmax) * (1 + 1e-9) + 1e-12;
    if (!(r < wmin * (0.5 - 4.0 / 65536.0))) {
        s.why_not = "rmax is not below half of the smallest cell width";
        return s;
    }
    double G[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) G[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
    double lam, lam_max;
    eig_sym3(G, &lam, &lam_max);
    lam *= 1 - 1e-9;
    if (!(lam > 0)) {
        s.why_not = "degenerate cell";
        return s;
    }
    s.smin = std::sqrt(lam);
    double R = r * 65536.0 / std::sqrt(lam) + std::sqrt(3.0) * 1.0001 + 1e-3;
    double R2 = std::ceil(R * R * (1 + 1e-9));
    if (!(R2 < 4294967295.0)) {
        s.why_not = "screen threshold out of range";
        return s;
    }
    s.valid = true;
    s.threshold = (uint32_t)R2;
    s.row_sum = 0;
    for (int i = 0; i < 3; i++)
        s.row_sum = std::max(s.row_sum, std::fabs(A[i][0]) + std::fabs(A[i][1]) + std::fabs(A[i][2]));
    s.wmin = wmin;
    s.fro = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s.fro += A[i][j] * A[i][j];
    s.fro = std::sqrt(s.fro);
    // |A|^T |A| and its largest eigenvalue (rounded up): the spectral norm of
    // |A| bounds || |A| |x| || by anorm |x|
    double H[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            H[i][j] = std::fabs(A[0][i]) * std::fabs(A[0][j]) + std::fabs(A[1][i]) * std::fabs(A[1][j]) +
                      std::fabs(A[2][i]) * std::fabs(A[2][j]);
    double hmin, hmax;
    eig_sym3(H, &hmin, &hmax);
    s.anorm = std::sqrt(hmax * (1 + 1e-9));
    return s;
}

// Window (in bins) of the FP32 candidate evaluation (eval_candidate), with
// u = 2^-24.  The pair vector is A' float(D), D the wrapped 32-bit integer
// difference of the fractions and A' the FP32 cell constants times 2^-32.
// Per Cartesian component k the int->float conversions, the rounded cell
// constants, the product and the FMAs of a row round the terms A_km D_m at
// most five times (three in the orthogonal form), so the computed vector is
// off by e_k <= c u (|A| |D|)_k with c = 6 (4) covering the second-order
// terms.  In the 2-norm, || |A| |D| || <= anorm |D| <= anorm |A D| / smin
// = K' |A D| (anorm: spectral norm of |A|; smin: smallest singular value of
// A), and |A D| <= dmax = rmax + 1 for the pairs whose bin matters, so
// |e| <= c u K' dmax.  The 32-bit quantisation of both atoms adds
// sqrt(3) row_sum 2^-31.  On d the three-term sum of squares (3u on d^2) and
// v_sqrt_f32 (1 ulp, 2u) add 3.5u d, taken as 4u dmax.  On u =
// (d - rmin)/dr: that over dr, plus the FP32 rmin (u rmin/dr), the
// subtraction, the FP32 1/dr, the product and the reference's own rounding
// of u to float (floorf of a double, gofrt.cpp:114-115), u umax each (5u
// umax, umax = nbin + 1).  Beyond the window around nbin every pair is
// rejected, as the reference does.  The window is four times the bound.
static double du_tail(double rmin, double dr, uint32_t nbin) {
    const double e24 = std::ldexp(1.0, -24);
    return 5.0 * e24 * (nbin + 1.0) + e24 * std::fabs(rmin) / dr;
}
static float du_window(const Screen &s, bool tri, double rmin, double rmax, double dr, uint32_t nbin) {
    const double e24 = std::ldexp(1.0, -24);
    const double dmax = std::fabs(rmax) + 1.0;
    const double Kp = s.anorm / s.smin;
    const double e_d = (tri ? 6.0 : 4.0) * e24 * Kp * dmax + std::sqrt(3.0) * s.row_sum * std::ldexp(1.0, -31) +
                       4.0 * e24 * dmax;
    const double du = e_d / dr + du_tail(rmin, dr, nbin);
    return (float)(4.0 * du * (1 + 1e-6) + 1e-6);
}

// Window of the offset evaluation (modes 2 and 4, eval_candidate_off), which
// subtracts FP32 Cartesian offsets of both atoms from the wavefront's centre
// c: P = A' float(W_i), Q = A' float(W_j), W = 32-bit fraction - c (exact,
// wrapped).  Per component, the conversions and the up to three products and
// sums of a row cost <= 4u (|A| |W|)_k for each atom (u = 2^-24); the FP32
// cell constants are common to P and Q, so their rounding acts on the
// difference: <= u (|A| |D|)_k; the subtraction adds u |d_k|.  In the 2-norm
// (K' = anorm / smin as in du_window), with |P| <= rho + dq and
// |Q| <= rho + dq + d for the wavefront radius rho and a pair at distance
// d <= dmax = rmax + 1:
//   |e| <= u K' (8 (rho + dq) + 5 dmax) + u dmax + sqrt(3) row_sum 2^-31,
// the last term the 32-bit quantisation; the sum, square root, u and the
// factor 4 follow du_window.  The window is m0 + m1 (rho + dq) bins,
// computed per wavefront from its own radius.
static void du_window_offsets(const Screen &s, double rmin, double rmax, double dr, uint32_t nbin, float *m0,
                              float *m1) {
    const double e24 = std::ldexp(1.0, -24);
    const double Kp = s.anorm / s.smin;
    const double dmax = std::fabs(rmax) + 1.0;
    const double c0 = e24 * Kp * 5.0 * dmax + e24 * dmax + std::sqrt(3.0) * s.row_sum * std::ldexp(1.0, -31) +
                      4.0 * e24 * dmax;
    const double c1 = e24 * Kp * 8.0;
    *m0 = (float)((4.0 * (c0 / dr + du_tail(rmin, dr, nbin)) + 1e-6) * (1 + 1e-6));
    *m1 = (float)(4.0 * c1 / dr * (1 + 1e-6));
}

// Bound (A, per Cartesian component) of the rounding in the reference's FP64
// pair vector (basetrajectory.h:200-268): the difference of two positions
// and every minimum-image step (z, then y, then x) round once, each by at
// most 2^-53 of a magnitude below M, and a component sees at most
// 1 + nz + ny + nx of them.  pmax bounds |position components|, so
// |xi - xj| <= 2 pmax; a z step moves y by yz and x by xz, a y step x by xy.
static double ref_vector_rounding(const Cell &c, double pmax) {
    const double u = std::ldexp(1.0, -53);
    const double d0 = 2.0 * pmax;
    const double nz = std::ceil(d0 / c.L[2]) + 1.0;
    const double m1 = d0 + nz * std::fabs(c.yz);
    const double ny = std::ceil(m1 / c.L[1]) + 1.0;
    const double m0 = d0 + nz * std::fabs(c.xz) + ny * std::fabs(c.xy);
    const double nx = std::ceil(m0 / c.L[0]) + 1.0;
    const double M = std::max(m0, m1) + std::max(c.L[0], std::max(c.L[1], c.L[2]));
    return u * M * (1.0 + nz + ny + nx);
}

// Margin (A^2) of the FP64 evaluation of a window candidate from its wrapped
// 32-bit fractional difference (recheck_candidate), with u = 2^-53.  Per
// Cartesian component of the pair vector:
//  - quantisation: an atom's fraction is rint(2^32 frac(s)) of its FP64
//    fractional coordinate s, off by <= q = 2^-33 + e_s, e_s bounding the
//    rounding of s (a few operations on magnitudes up to (pmax + |lo|) over a
//    cell length, tilts included); two atoms, row_sum 2 q;
//  - this evaluation: three roundings of sums of terms <= row_sum / 2;
//  - the reference's own sequence (ref_vector_rounding).
// With e = sqrt(3) e_comp the norm of the vector error and d <= dmax =
// rmax + 1, |d2 - d2_ref| <= (2 dmax + e) e + 8 u dmax^2, the last term the
// rounding of both squared sums.  Pairs beyond dmax are rejected by both
// (e is far below 1 A).  The margin is four times the bound.
static double d2_window_fp64(const Screen &s, const Cell &c, double pmax, double rmax) {
    const double u = std::ldexp(1.0, -53);
    const double lmin = std::min(c.L[0], std::min(c.L[1], c.L[2]));
    const double tilt = std::max(std::fabs(c.xy), std::max(std::fabs(c.xz), std::fabs(c.yz))) / lmin;
    const double lo = std::max(std::fabs(c.lo[0]), std::max(std::fabs(c.lo[1]), std::fabs(c.lo[2])));
    const double e_s = 16.0 * u * (1.0 + tilt) * (1.0 + tilt) * ((pmax + lo) / lmin + 1.0);
    const double q = std::ldexp(1.0, -33) + e_s;
    const double e_comp = s.row_sum * 2.0 * q + 4.0 * u * s.row_sum + ref_vector_rounding(c, pmax);
    const double e = std::sqrt(3.0) * e_comp;
    const double dmax = std::fabs(rmax) + 1.0;
    return 4.0 * ((2.0 * dmax + e) * e + 8.0 * u * dmax * dmax) * (1 + 1e-6);
}

// Cluster screen.  A wavefront's 64 i atoms in frame fi get a common centre
// c (16-bit fractional) and Cartesian offsets p_i = A wrap16(f_i - c); j atoms
// get q_j = A wrap16(f_j - c).  If rho = max |p_i| satisfies
//   rho + rmax + dq < w_min (1/2 - 2^-15)            (rho < rho_cap)
// then for every pair the reference can count (d <= rmax) the vector
// q_j - p_i is the reference's minimum-image vector up to the 16-bit
// quantisation (|error| <= dq): the fractional components of
// (f_j + n - c) stay strictly inside (-1/2, 1/2), so wrapping relative to c
// picks the reference's image.  The squared distance
//   |q - p|^2 = |p|^2 + |q|^2 - 2 p.q
// of 32 x 32 pairs is then one MFMA with rows (-2p, |p|^2, 1) and columns
// (q, 1, |q|^2) in FP16, accumulated in FP32.  Rounding p, q, |p|^2, |q|^2 to
// FP16 (2^-11 relative) and the FP32 sum change the result by at most
// 2^-10 (|p| + |q|)^2, and |q| <= rho + rmax + dq for counted pairs, so
//   D < ((rmax + dq)^2 + 2^-10 (2 rho + rmax + dq)^2) (1 + 2^-12) + 1e-4
// holds for every pair the reference counts.  Passing pairs go to the same
// candidate queue as the 16-bit screen's.
//
// FP16 range.  Inactive rows carry |p|^2 = 60000 and columns past N carry
// |q|^2 = 60000, so that their D >= 60000 never passes; and a counted pair's
// |q|^2 must be finite in FP16 (< 65504).  Both hold when
//   (rho + rmax + dq)^2 <= CLUSTER_Q2_MAX = 2^14,
// since then the threshold is below 2^14 (1 + 2^-8)(1 + 2^-12) + 1e-4, far
// under the sentinel.  rho_cap enforces it; a cell or rmax that leaves no room
// turns the cluster screen off (rho_cap = 0) and the 16-bit screen runs.
constexpr double CLUSTER_Q2_MAX = 16384.0;
constexpr float CLUSTER_SENTINEL = 60000.0f;

struct ClusterParams {
    double dq;       // bound of a pair vector's error from 16-bit quantisation of both atoms
    double rmaxf;    // rmax as the kernel compares it
    double rho_cap;  // largest usable wavefront radius, 0 = cluster screen off
};

// dq: each fractional component within 2^-17 + 2^-33 per atom, times
// |A|_2 <= |A|_F, sqrt(3) components.  rho_cap: rho + rmax + dq below
// w_min (1/2 - 2^-15) (the image argument) and below sqrt(CLUSTER_Q2_MAX)
// (the FP16 range), less 1e-3 A for the FP32 evaluation of rho.
static ClusterParams cluster_params(const Screen &scr, double rmax) {
    ClusterParams p;
    p.dq = 2.0 * std::sqrt(3.0) * scr.fro * (std::ldexp(1.0, -17) + std::ldexp(1.0, -33)) * 1.001 * 1.0001;
    p.rmaxf = std::fabs(rmax) * (1 + 1e-6);
    const double image_cap = scr.wmin * (0.5 - std::ldexp(1.0, -15)) - p.rmaxf - p.dq - 1e-3;
    const double range_cap = std::sqrt(CLUSTER_Q2_MAX) - p.rmaxf - p.dq - 1e-3;
    const double cap = std::min(image_cap, range_cap);
    p.rho_cap = scr.valid && cap > 0 ? cap : 0.0;
    return p;
}

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------

constexpr int WG = 256;       // threads per workgroup = i atoms per tile

// Launch statistics (device counters, zeroed by every launch).  A pair is
// "evaluated" when some distance test looks at that very (i, j) pair: the
// FP64 sequence, the 16-bit screen or an FP16 tile of the cluster screen.
// Pairs of 32-column halves that the cluster screen skips whole (no j of the
// half near the wavefront) are pruned, not evaluated.
enum Stat {
    STAT_CANDIDATES = 0,  // pairs passed to the FP32 evaluation
    STAT_RECHECKS,        // candidates re-evaluated in FP64 from their fractions (FP32 result near a bin edge, rmin or rmax)
    STAT_CLUSTERED,       // (wavefront, t0) units that used the cluster screen
    STAT_UNITS,           // all (wavefront, t0) units of a cluster-screen launch
    STAT_PAIRS_TILE,      // ordered pairs evaluated by FP16 tiles
    STAT_PAIRS_16,        // ordered pairs evaluated by the 16-bit screen
    STAT_PAIRS_FP64,      // ordered pairs evaluated by the plain FP64 kernel
    STAT_PAIRS_ALL,       // all ordered pairs (i active, any j) of the launch
    STAT_EXACT,           // rechecks evaluated with the reference's FP64 sequence (still near an edge)
    NSTAT
};
// FP64 window constants stored after the bin edges: the cell matrix entries
// times 2^-32 and the margin eps_d2 (d2_window_fp64)
enum Fp64K { FP64K_X = 0, FP64K_XY, FP64K_XZ, FP64K_Y, FP64K_YZ, FP64K_Z, FP64K_EPS, NFP64K };
constexpr int QCAP = 192;     // per-wavefront candidate queue entries

struct BlockInfo {
    uint64_t primo;
    uint32_t leff;
    uint32_t pad;
};

struct WorkItem {
    uint32_t block;
    uint32_t k0, k1;  // t0 indices [k0, k1): t0 = k * skip
    uint32_t pad;
};

struct GofrtArgs {
    const double4 *pos;
    const uint2 *frac;
    const double4 *pos_scr;      // screened form: positions in each frame's screen order
    const double *box;
    const int32_t *type_slot;
    const int32_t *orig_slot;
    unsigned long long *out;     // [nblocks][leff_max][nch][nbin]
    unsigned long long *ncand;   // [NSTAT] launch statistics, see the Stat enum
    const double *thr;           // [nbin+1] bin edges, then the FP64 window constants (FP64K_*)
    const int32_t *chan;         // [nt*nt]
    const WorkItem *work;
    const BlockInfo *blocks;
    uint64_t N;
    uint64_t out_block_stride;   // leff_max * nch * nbin
    uint32_t nt, nbin, nch, half, skip, every;
    double lo, rmax2, Tn;
    float rminf, inv_drf;
    uint32_t threshold;
    int32_t a0, a1;              // original-index range of i atoms
    uint32_t lag0;               // first lag (lag = lag0 + lag index * every)
    uint32_t ntiles, nlag, nwork, per_xcd;  // 1-D grid decoding, see decode_wg (per_xcd 0: no remap)
    // screened form: 32-bit fractions and the FP32 evaluation of candidates
    const uint4 *frac32;
    unsigned long long *nrecheck;
    float cx, cxy, cxz, cy, cyz, cz;  // cell matrix entries times 2^-32
    float du_margin;                  // bins: |u - nearest integer| below this -> FP64
    float du_m0, du_m1;               // mode 2: window m0 + m1 (rho + dq) of the offset evaluation
    // cluster screen (see cluster_threshold): cell entries times 2^-16,
    // largest usable cluster radius, quantisation bound, rmax
    float hx, hxy, hxz, hy, hyz, hz;
    float rho_cap, dq, rmaxf;
    // bounding spheres of 32-slot groups of each frame ([F][ngroups]) and
    // 2^16 / (smallest cell width), the fractional bound of a group's extent
    const uint4 *jgroups;
    uint32_t ngroups;
    float inv_w16;
    float smin16;  // sqrt(smallest eigenvalue of A^T A) times 2^-16
};

// The grid is one-dimensional.  Workgroups are dispatched round-robin over
// the 8 XCDs (hardware id h runs on XCD h % 8), so logical workgroup
// L = (h % 8) * per_xcd + h / 8 puts consecutive logical ids -- the i tiles of
// one (lag, t0 range), which read the same j frames -- on one XCD and its L2.
// The map is a bijection on [0, 8 * per_xcd); ids >= the real count idle.
struct WgId {
    uint32_t tile, lagi, work;
    bool valid;
};
__device__ __forceinline__ WgId decode_wg(uint32_t ntiles, uint32_t nlag, uint32_t nwork, uint32_t per_xcd) {
    const uint32_t h = blockIdx.x;
    const uint32_t L = per_xcd ? (h & 7u) * per_xcd + (h >> 3) : h;
    WgId w;
    w.valid = L < ntiles * nlag * nwork;
    w.tile = L % ntiles;
    const uint32_t rest = L / ntiles;
    w.lagi = rest % nlag;
    w.work = rest / nlag;
    return w;
}

#include "min_image.h"

__device__ __forceinline__ int bin_of(double d2, const double *T, int nbin, float rminf, float inv_drf) {
    // a first guess (any rounding of it is corrected by the edge table)
    float u = (__builtin_amdgcn_sqrtf((float)d2) - rminf) * inv_drf;
    int b = (int)u;
    b = b < 0 ? 0 : (b > nbin - 1 ? nbin - 1 : b);
    while (b > 0 && d2 < T[b]) --b;
    while (b < nbin - 1 && d2 >= T[b + 1]) ++b;
    return b;
}

__device__ __forceinline__ void lds_setup(const GofrtArgs &a, unsigned *hist, double *thr, int *chan) {
    const int nh = (int)(a.nch * a.nbin);
    for (int k = threadIdx.x; k < nh; k += WG) hist[k] = 0u;
    for (int k = threadIdx.x; k <= (int)a.nbin; k += WG) thr[k] = a.thr[k];
    for (int k = threadIdx.x; k < (int)(a.nt * a.nt); k += WG) chan[k] = a.chan[k];
}

// Screened form: the workgroup's histogram rows are the ordered type pairs of
// distinct atoms (row ti nt + tj) and the types of an atom paired with itself
// (row nt^2 + ti): nt (nt + 1) rows, as many as the reference's channels, so
// that a candidate's row is arithmetic (hidx) and the flush folds the rows
// into the channels: rowch[row] = (channel) * nbin, the rows (ti, tj) and
// (tj, ti) sharing the channel of the pair and row nt^2 + t the self channel
// half + channel(t, t).
__device__ __forceinline__ void lds_setup_rows(const GofrtArgs &a, unsigned *hist, double *thr, int *rowch) {
    const int nh = (int)(a.nch * a.nbin);
    for (int k = threadIdx.x; k < nh; k += WG) hist[k] = 0u;
    for (int k = threadIdx.x; k <= (int)a.nbin; k += WG) thr[k] = a.thr[k];
    const int nt2 = (int)(a.nt * a.nt);
    for (int k = threadIdx.x; k < nt2 + (int)a.nt; k += WG)
        rowch[k] = (k < nt2 ? a.chan[k] : a.chan[(k - nt2) * ((int)a.nt + 1)] + (int)a.half) * (int)a.nbin;
}
// x y + z in one 24-bit multiply-add (x, y < 2^24; the compiler otherwise
// picks a 64-bit multiply-add or a multiply, shifts and a three-way add)
__device__ __forceinline__ uint32_t mad24(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(x), "s"(y), "v"(z));
    return r;
}
// the histogram index of a pair from the .w words (type | slot << 8) of its
// atoms: row ti nt + tj (see lds_setup_rows) times nbin plus its bin bb.
// Every pair takes its distinct row, an atom paired with itself included:
// the screened kernel moves each self pair to its self row afterwards
// (move_self_pair), which keeps the per-candidate index free of the test.
__device__ __forceinline__ uint32_t hidx(uint32_t wi, uint32_t wj, uint32_t nt, uint32_t nbin, uint32_t bb) {
    return mad24(mad24(wi & 0xFFu, nt, wj & 0xFFu), nbin, bb);
}

// LDS (address space 3) pointers and the 32-bit LDS address of a pointer
// into the workgroup's shared memory
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) u32x4_t lds_u4_t;
typedef __attribute__((address_space(3))) unsigned lds_u32_t;
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
    return (uint32_t)(size_t)(const __attribute__((address_space(3))) unsigned char *)p;
}

// v_cvt_flr_i32_f32: (int)floor(x) in one instruction (the compiler emits a
// floor and a conversion)
__device__ __forceinline__ int floor_i32(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// The screened form's flush: row r of the workgroup's histogram adds to its
// channel rowch[r] (see lds_setup_rows), from a rotated start as lds_flush.
__device__ __forceinline__ void lds_flush_rows(const GofrtArgs &a, const unsigned *hist, const int *rowch,
                                               uint32_t block, uint32_t t) {
    const int nh = (int)(a.nch * a.nbin);
    unsigned long long *o = a.out + (uint64_t)block * a.out_block_stride + (uint64_t)t * nh;
    const int rot = (int)((blockIdx.x * 977u) % (uint32_t)nh);
    for (int k0 = threadIdx.x; k0 < nh; k0 += WG) {
        const int k = k0 + rot < nh ? k0 + rot : k0 + rot - nh;
        const unsigned v = hist[k];
        const int r = k / (int)a.nbin;
        if (v) atomicAdd(o + rowch[r] + (k - r * (int)a.nbin), (unsigned long long)v);
    }
}

// Workgroups of one (block, lag) flush into the same counters, often at the
// same time (they start and end together): each starts at its own offset,
// so that their atomics meet different addresses.
__device__ __forceinline__ void lds_flush(const GofrtArgs &a, const unsigned *hist, uint32_t block, uint32_t t) {
    const int nh = (int)(a.nch * a.nbin);
    unsigned long long *o = a.out + (uint64_t)block * a.out_block_stride + (uint64_t)t * nh;
    const int rot = (int)((blockIdx.x * 977u) % (uint32_t)nh);
    for (int k0 = threadIdx.x; k0 < nh; k0 += WG) {
        const int k = k0 + rot < nh ? k0 + rot : k0 + rot - nh;
        unsigned v = hist[k];
        if (v) atomicAdd(o + k, (unsigned long long)v);
    }
}

// grid: x = i tile (256 slots), y = lag index, z = work item (block, t0 range)
template <bool TRI>
__global__ __launch_bounds__(WG) void gofrt_fp64_kernel(const GofrtArgs a) {
    extern __shared__ unsigned char smem[];
    const WgId id = decode_wg(a.ntiles, a.nlag, a.nwork, a.per_xcd);
    if (!id.valid) return;
    const WorkItem w = a.work[id.work];
    const BlockInfo b = a.blocks[w.block];
    const uint32_t t = a.lag0 + id.lagi * a.every;
    if (t >= b.leff) return;
    double *thr = (double *)smem;
    int *chan = (int *)(thr + a.nbin + 1);
    unsigned *hist = (unsigned *)(chan + a.nt * a.nt);
    lds_setup(a, hist, thr, chan);
    __syncthreads();

    const uint64_t N = a.N;
    const uint64_t slot = (uint64_t)id.tile * WG + threadIdx.x;
    bool active = slot < N;
    int ti = 0;
    if (active) {
        int o = a.orig_slot[slot];
        active = o >= a.a0 && o < a.a1;
        ti = a.type_slot[slot];
    }
    const int nbin = (int)a.nbin;
    // every ordered pair of an active i is evaluated: N per t0
    const unsigned long long npairs = (unsigned long long)__popcll(__ballot(active)) * N * (w.k1 - w.k0);
    for (uint32_t k = w.k0; k < w.k1; k++) {
        const uint64_t fi = b.primo + (uint64_t)k * a.skip;
        const uint64_t fj = fi + t;
        const DevCell c = load_cell(a.box + fi * 9);
        const double4 *__restrict__ pj = a.pos + fj * N;
        if (active) {
            const double4 xi = a.pos[fi * N + slot];
            for (uint64_t j = 0; j < N; j++) {
                const double4 xj = pj[j];
                double d2 = d2_ref<TRI>(xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, c);
                if (d2 >= a.lo && d2 <= a.rmax2 && d2 < a.Tn) {
                    int bb = bin_of(d2, thr, nbin, a.rminf, a.inv_drf);
                    int tj = a.type_slot[j];
                    int ch = chan[ti * a.nt + tj] + (slot == j ? (int)a.half : 0);
                    atomicAdd(&hist[ch * nbin + bb], 1u);
                }
            }
        }
    }
    if ((threadIdx.x & 63u) == 0 && a.ncand && npairs) {
        atomicAdd(a.ncand + STAT_PAIRS_FP64, npairs);
        atomicAdd(a.ncand + STAT_PAIRS_ALL, npairs);
    }
    __syncthreads();
    lds_flush(a, hist, w.block, t);
}

typedef short short2_t __attribute__((ext_vector_type(2)));
typedef unsigned short ushort2_t __attribute__((ext_vector_type(2)));

// |q|^2 of the 16-bit wrapped fractional difference of (a,b | c) words:
// packed 2-wide subtractions wrap each component to [-2^15, 2^15), and the
// sum of the three squares (< 3*2^30) is exact as a 32-bit unsigned value.
// The c word's high half is 0 - 0 = 0.  The first dot product takes the
// inline constant 0 as accumulator (VOP3P form), the second accumulates in
// place, so a pair costs 2 subtractions and 2 dot products.
__device__ __forceinline__ uint32_t sq16(uint32_t ab_i, uint32_t ab_j, uint32_t c_i, uint32_t c_j) {
    ushort2_t dab = __builtin_bit_cast(ushort2_t, ab_i) - __builtin_bit_cast(ushort2_t, ab_j);
    ushort2_t dc = __builtin_bit_cast(ushort2_t, c_i) - __builtin_bit_cast(ushort2_t, c_j);
    int s;
    asm("v_dot2_i32_i16 %0, %1, %1, 0" : "=v"(s) : "v"(__builtin_bit_cast(uint32_t, dc)));
    s = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, dab), __builtin_bit_cast(short2_t, dab), s, false);
    return (uint32_t)s;
}

// Loads through the constant address space: uniform addresses become scalar
// (SMEM) loads, so the j atom's data arrives in SGPRs.
typedef const __attribute__((address_space(4))) unsigned long long cu64_t;

__device__ __forceinline__ uint32_t lane_prefix(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// base + number of set bits of m below this lane
__device__ __forceinline__ uint32_t lane_slot(unsigned long long m, uint32_t base) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}

// One queued candidate (entry = j << 8 | i_local).  The distance is first
// evaluated in FP32 from the 32-bit fractional difference (the wrapped
// difference is the minimum image for every pair the reference can count);
// if u = (d - rmin)/dr lies within du_margin of an integer (a bin edge, rmin
// or rmax) the pair is re-evaluated with the reference's exact FP64 sequence.
// du_margin is four times a worst-case bound of |u_fp32 - u_reference|, so
// outside the window both floor to the same bin.
template <bool TRI>
__device__ __forceinline__ bool eval_candidate(const GofrtArgs &a, uint32_t e, const uint4 *fi_lds, const uint4 uj,
                                               unsigned *hist, unsigned *sink) {
    const uint32_t il = e & 0xFFu;
    const uint4 ui = fi_lds[il];  // .w = type | slot << 8 of i; uj likewise of j
    const float fa = (float)(int)(ui.x - uj.x);
    const float fb = (float)(int)(ui.y - uj.y);
    const float fc = (float)(int)(ui.z - uj.z);
    float dx, dy, dz;
    if (TRI) {
        dz = a.cz * fc;
        dy = fmaf(a.cyz, fc, a.cy * fb);
        dx = fmaf(a.cxz, fc, fmaf(a.cxy, fb, a.cx * fa));
    } else {
        dx = a.cx * fa;
        dy = a.cy * fb;
        dz = a.cz * fc;
    }
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float u = (__builtin_amdgcn_sqrtf(d2) - a.rminf) * a.inv_drf;
    const bool need = fabsf(u - rintf(u)) < a.du_margin;
    const int bb = floor_i32(u);
    const bool add = !need && (unsigned)bb < a.nbin;
    atomicAdd(add ? &hist[hidx(ui.w, uj.w, a.nt, a.nbin, (uint32_t)bb)] : sink, 1u);
    return need;
}

// FP32 Cartesian offset of an atom (32-bit fractions f) from a centre c
// (32-bit fractional units): A' float(f - c), the wrapped difference taken
// as a signed integer; .w = w, the atom's part of its pairs' histogram index
// (eval_candidate_off).
template <bool TRI>
__device__ __forceinline__ uint4 offset32(const GofrtArgs &a, const uint4 f, uint32_t ca, uint32_t cb, uint32_t cc,
                                          uint32_t w) {
    const float fa = (float)(int)(f.x - ca);
    const float fb = (float)(int)(f.y - cb);
    const float fc = (float)(int)(f.z - cc);
    float x, y, z;
    if (TRI) {
        z = a.cz * fc;
        y = fmaf(a.cyz, fc, a.cy * fb);
        x = fmaf(a.cxz, fc, fmaf(a.cxy, fb, a.cx * fa));
    } else {
        x = a.cx * fa;
        y = a.cy * fb;
        z = a.cz * fc;
    }
    return make_uint4(__float_as_uint(x), __float_as_uint(y), __float_as_uint(z), w);
}

// Mode 2: a candidate from the offsets of both atoms from the wavefront's
// centre (pi from fi_lds, qj from the column cache).  For every pair the
// reference can count, wrapping relative to the centre picks the
// reference's image (the cluster screen's validity condition), so the pair
// vector is P - Q; margin (bins) bounds the error of this form for the
// wavefront's radius.  Byte offsets from the LDS base lds: the entry holds
// the row's (low half) and the column's (high half); the .w words hold the
// histogram's offset plus 4 ti nt nbin (row) and 4 tj nbin (column), so
// that the count's offset is their sum plus 4 bins.
__device__ __forceinline__ bool eval_candidate_off(const GofrtArgs &a, uint32_t e, uint32_t sink, float margin) {
    const u32x4_t pi = *(const lds_u4_t *)(size_t)(e & 0xFFFFu);
    const u32x4_t qj = *(const lds_u4_t *)(size_t)(e >> 16);
    const float dx = __uint_as_float(pi.x) - __uint_as_float(qj.x);
    const float dy = __uint_as_float(pi.y) - __uint_as_float(qj.y);
    const float dz = __uint_as_float(pi.z) - __uint_as_float(qj.z);
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float u = (__builtin_amdgcn_sqrtf(d2) - a.rminf) * a.inv_drf;
    const bool need = fabsf(u - rintf(u)) < margin;
    const int bb = floor_i32(u);
    const bool add = !need && (unsigned)bb < a.nbin;
    const uint32_t at = add ? pi.w + qj.w + ((uint32_t)bb << 2) : sink;
    __hip_atomic_fetch_add((lds_u32_t *)(size_t)at, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return need;
}

// A pointer into read-only device memory for the recheck's rare path, opaque
// to the optimiser, so that loads through it stay where they are written
// instead of being hoisted out of the pair loop, where their results would
// hold registers across it.  SCALAR: wavefront-uniform scalar loads (constant
// address space); otherwise vector loads from global memory.  Measured on
// 128/300-t0 slices: the triclinic kernels run best with scalar loads (their
// spill fell from 44 to 8 B/lane; 987 -> 931 ms), the cubic ones with vector
// loads (1,663 ms scalar, 1,614 ms vector).
typedef const __attribute__((address_space(4))) double *rare_cptr;
typedef const __attribute__((address_space(1))) double *rare_gptr;
template <bool SCALAR>
__device__ __forceinline__ auto rare_const(const double *p) {
    uint64_t v = (uint64_t)p;
    if constexpr (SCALAR) {
        asm volatile("" : "+s"(v));
        v = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
        return (rare_cptr)v;
    } else {
        asm volatile("" : "+v"(v));
        return (rare_gptr)v;
    }
}

// A candidate whose FP32 distance fell inside the window around a bin edge,
// rmin or rmax, in two FP64 stages.  First the pair vector from the same
// wrapped 32-bit fractional difference in FP64: its squared distance is
// within eps_d2 of the reference's (the 32-bit quantisation of both atoms
// and the rounding of both FP64 sequences; see d2_window_fp64), so if every
// value within eps_d2 of it falls on the same side of rmin, rmax and the
// bin edges, that decision is the reference's.  The rest -- pairs within
// about 1e-7 A of an edge -- take the reference's exact FP64 sequence from
// the positions.  pi and pj are the positions of the two frames in screen
// order and fi32 the t0 frame's fractions from the tile's first slot (the
// order of the entry's indices).  At lag 0 (same_frame) an atom paired with
// itself has the reference's d2 = 0 exactly, as the first stage computes it,
// so its margin is 0.  Returns whether the exact sequence ran.
template <bool TRI>
__device__ __forceinline__ bool recheck_candidate(const GofrtArgs &a, bool same_frame, uint32_t e,
                                                  const uint4 *__restrict__ fi32,
                                                  const uint4 *__restrict__ fj32, const double4 *__restrict__ pi,
                                                  const double4 *__restrict__ pj, const double *cell,
                                                  const double *thr, unsigned *hist) {
    const uint32_t il = e & 0xFFu;
    const uint32_t j = e >> 8;
    const uint4 ui = fi32[il], uj = fj32[j];  // .w = type | slot << 8
    const double fa = (double)(int)(ui.x - uj.x);
    const double fb = (double)(int)(ui.y - uj.y);
    const double fc = (double)(int)(ui.z - uj.z);
    // the constants are read here, on this rare path, rather than held in
    // scalar registers across the whole pair loop
    const auto k = rare_const<TRI>(a.thr + a.nbin + 1);
    double x, y, z;
    if (TRI) {
        z = k[FP64K_Z] * fc;
        y = fma(k[FP64K_YZ], fc, k[FP64K_Y] * fb);
        x = fma(k[FP64K_XZ], fc, fma(k[FP64K_XY], fb, k[FP64K_X] * fa));
    } else {
        x = k[FP64K_X] * fa;
        y = k[FP64K_Y] * fb;
        z = k[FP64K_Z] * fc;
    }
    const double d2 = fma(z, z, fma(y, y, x * x));
    const double eps = same_frame && ui.w == uj.w ? 0.0 : k[FP64K_EPS];
    const double dlo = d2 - eps, dhi = d2 + eps;
    const int nb = (int)a.nbin;
    int bb = -1;  // bin of the pair, -1: not counted
    bool exact = true;
    if (dhi < a.lo || dlo > a.rmax2 || dlo >= a.Tn) {
        exact = false;  // below rmin or beyond rmax for every value in the margin
    } else if (dlo >= a.lo && dhi <= a.rmax2 && dhi < a.Tn) {
        const int b = bin_of(d2, thr, nb, a.rminf, a.inv_drf);
        if ((b == 0 || dlo >= thr[b]) && (b == nb - 1 || dhi < thr[b + 1])) {
            bb = b;
            exact = false;
        }
    }
    if (exact) {
        const double4 xi = pi[il];
        const double4 xj = pj[j];
        // the t0 frame's cell is read here, on this rare path, rather than
        // held in registers across the whole pair loop
        const DevCell c = load_cell(rare_const<TRI>(cell));
        const double d2r = d2_ref<TRI>(xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, c);
        if (d2r >= a.lo && d2r <= a.rmax2 && d2r < a.Tn) bb = bin_of(d2r, thr, nb, a.rminf, a.inv_drf);
    }
    if (bb >= 0) atomicAdd(&hist[hidx(ui.w, uj.w, a.nt, a.nbin, (uint32_t)bb)], 1u);
    return exact;
}

// An atom paired with itself, at the t0 frame fi and the frame fj: every
// candidate path above bins it like any pair, into the distinct row (ti, ti)
// of its type and with the reference's bin (or not at all).  This moves it
// to the self row nt^2 + ti: the bin is that of the reference's FP64
// sequence from the positions (pos, in slot order), the same bin, so the
// count taken from the distinct row is the one the pair put there.  w is
// the atom's .w word (type | slot << 8), cell the t0 frame's cell.
template <bool TRI>
__device__ __forceinline__ void move_self_pair(const GofrtArgs &a, uint64_t fi, uint64_t fj, uint32_t w,
                                               const double *cell, const double *thr, unsigned *hist) {
    const uint32_t sl = w >> 8, ti = w & 0xFFu;
    const double4 xi = a.pos[fi * a.N + sl];
    const double4 xj = a.pos[fj * a.N + sl];
    const DevCell c = load_cell(rare_const<TRI>(cell));
    const double d2 = d2_ref<TRI>(xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, c);
    if (d2 >= a.lo && d2 <= a.rmax2 && d2 < a.Tn) {
        const uint32_t bb = (uint32_t)bin_of(d2, thr, (int)a.nbin, a.rminf, a.inv_drf);
        atomicSub(&hist[(ti * a.nt + ti) * a.nbin + bb], 1u);
        atomicAdd(&hist[(a.nt * a.nt + ti) * a.nbin + bb], 1u);
    }
}

// Per-wavefront queues.  Each ends with QSINK = 64 entries that only take
// the stores of lanes whose test failed (entry sink + lane), so that an
// append is one store by every lane, with no change of the execution mask;
// nothing reads them.  The recheck queue's sink also takes the histogram adds
// of candidates that are not counted.
constexpr int QSINK = 64;
constexpr int SQCAP = 576 + QSINK;       // 16-bit screen: < 64 left + up to 8 masks of 64
constexpr int SQCAP_CL = 1088 + QSINK;   // cluster screen: < 64 left + 16 masks of 64 (one drain per 32 x 32 tile)
constexpr int SQCAP_CLC = 576 + QSINK;   // cluster screen with the LDS column cache: < 64 left + 8 masks (a drain per 8)
constexpr int RQCAP = 128 + QSINK;       // recheck queue: < 64 left + one batch

// c += v for a wavefront-uniform count.  PIN keeps it in scalar registers
// (readfirstlane): that lowers the vector-register spills of the triclinic
// screened kernels and speeds them up (3 % on the 16,384-atom triclinic
// workload), while in the cubic ones it costs time (5 % on the 65,536-atom
// cubic workload), where the compiler is left to place the counts.
template <bool PIN>
__device__ __forceinline__ void uadd(unsigned long long &c, unsigned long long v) {
    if constexpr (PIN) {
        const unsigned long long r = c + v;
        c = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(r >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)r);
    } else {
        c += v;
    }
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int s16(uint32_t x) { return (int)(short)(x & 0xFFFFu); }

// Wavefront sum / max over all 64 lanes (butterfly, every lane gets the result).
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// The cluster screen's distance threshold for a wavefront of radius rho (see
// the kernel); shared by the kernel and the host-side range check below.
__host__ __device__ __forceinline__ float cluster_threshold(float rho, float rmax, float dq) {
    const float a = rmax + dq, b = 2.0f * rho + rmax + dq;
    return (a * a + b * b * 0x1p-10f) * (1.0f + 0x1p-12f) + 1e-4f;
}

// appends: every lane forms its slot and lanes whose test failed store into
// the queue's sink instead; the empty asm keeps the compiler from turning the
// select back into an execution-mask branch.  The slot is mbcnt on the count
// qn, or, in the cubic kernels' cluster screen (QADDR), qa is the LDS byte
// address of the queue's end and qsink this lane's sink entry, so that a
// slot's address is one shift-add on the lane's prefix with no queue base to
// reload (measured: 2 % faster on the cubic workload, 7 % slower on the
// triclinic one).
#define ANL_APPEND_E(P, M, ENTRY)                                           \
    if (M) {                                                                \
        uint32_t slot_ = lane_slot(M, qn);                                  \
        asm volatile("" : "+v"(slot_)); /* every lane: a select, no branch */ \
        q[(P) ? slot_ : (uint32_t)(QC - QSINK) + lane] = (ENTRY);           \
        qn += (uint32_t)__popcll(M);                                        \
    }

#define ANL_APPEND_A(P, M, ENTRY)                                           \
    if (M) {                                                                \
        uint32_t at_ = qa + 4u * lane_prefix(M);                            \
        asm volatile("" : "+v"(at_)); /* every lane: a select, no branch */  \
        *(lds_u32_t *)(size_t)((P) ? at_ : qsink) = (ENTRY);                \
        qa += 4u * (uint32_t)__popcll(M);                                   \
    }

#define ANL_APPEND(P, M, J)                                                 \
    if (M) {                                                                \
        uint32_t slot_ = lane_slot(M, qn);                                  \
        asm volatile("" : "+v"(slot_)); /* every lane: a select, no branch */ \
        q[(P) ? slot_ : (uint32_t)(QC - QSINK) + lane] = ((uint32_t)(J) << 8) | il; \
        qn += (uint32_t)__popcll(M);                                        \
    }

// CMODE, the cluster screen's j loop: 1 the groups left by the bounding-
// sphere test, two per step; 2 the same with each step's columns cached in
// LDS as FP32 offsets for the candidates' evaluation; 3 every 64 consecutive
// j per step, the next step's fractions loaded ahead (no group table); 4 as
// 3 with the column cache of 2.  0 without CLUSTER.
template <bool TRI, int UNROLL, bool CLUSTER, int CMODE>
__global__ __launch_bounds__(WG, CLUSTER ? 4 : 1) void gofrt_screened_kernel(const GofrtArgs a) {
    constexpr bool COLC = CMODE == 2 || CMODE == 4;    // column cache + offset evaluation
    constexpr bool GROUPS = CMODE == 1 || CMODE == 2;  // j groups from the bounding-sphere table
    constexpr bool QADDR = !TRI;                       // cluster queue's end as an LDS address
    extern __shared__ unsigned char smem[];
    const WgId id = decode_wg(a.ntiles, a.nlag, a.nwork, a.per_xcd);
    if (!id.valid) return;
    const WorkItem w = a.work[id.work];
    const BlockInfo b = a.blocks[w.block];
    const uint32_t t = a.lag0 + id.lagi * a.every;
    if (t >= b.leff) return;
    uint4 *fi_lds = (uint4 *)smem;                          // [WG]
    // cluster screen: the 32-bit fractions of the current step's 64 columns,
    // per wavefront, read by the candidate evaluation instead of global memory
    uint4 *colc_all = fi_lds + WG;                          // [WG/64][64] (COLC)
    constexpr int QC = COLC ? SQCAP_CLC : (CLUSTER ? SQCAP_CL : SQCAP);
    uint32_t *queue = (uint32_t *)(fi_lds + WG + (COLC ? WG : 0));  // [WG/64][QC]
    uint32_t *rqueue = queue + (WG / 64) * QC;              // [WG/64][RQCAP]
    double *thr = (double *)(rqueue + (WG / 64) * RQCAP);   // [nbin+1]
    int *rowch = (int *)(thr + a.nbin + 1);                 // [nt*nt*2], the first nt*(nt+1) used
    unsigned *hist = (unsigned *)(rowch + 2 * a.nt * a.nt); // [nch*nbin]
    lds_setup_rows(a, hist, thr, rowch);

    const uint64_t N = a.N;
    const uint32_t lane = threadIdx.x & 63u;
    // per-wavefront queue; its base is wavefront-uniform (kept scalar)
    uint32_t *q = queue + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * QC;
    uint32_t *rq = rqueue + (threadIdx.x >> 6) * RQCAP;
    // this lane's sink entry of the recheck queue also takes the histogram
    // add of a candidate that is not counted, so that the add needs no branch
    unsigned *rsink = rq + (RQCAP - QSINK) + lane;
    // LDS address of fi_lds (the start of the shared memory); the column
    // cache's queue entries hold LDS addresses below 2^16 (fi_lds and the
    // column cache take the first 32 WG bytes)
    const uint32_t lds0 = lds_addr(smem);
    const uint64_t slot0 = (uint64_t)id.tile * WG;
    const uint64_t slot = slot0 + threadIdx.x;
    // an atom range of calc_single_th narrower than all atoms is tested per
    // t0: the fractions of each frame are in that frame's own slot order
    const bool all_atoms = a.a0 <= 0 && (int64_t)a.a1 >= (int64_t)N;
    const uint32_t il = threadIdx.x;
    constexpr int UR = UNROLL;
    const uint32_t NU = (uint32_t)(N - N % UR);
    // wavefront-uniform counts: candidates,
    // window rechecks and their exact evaluations, clustered and all units;
    // all ordered pairs of the active i atoms, pairs evaluated by the FP16
    // tile screen and by the 16-bit screen
    unsigned long long ncand = 0, nre = 0, nex = 0, ncl = 0, nwt = 0;
    unsigned long long nall = 0, ntile = 0, n16 = 0;

    for (uint32_t k = w.k0; k < w.k1; k++) {
        const uint64_t fi = b.primo + (uint64_t)k * a.skip;
        const uint64_t fj = fi + t;
        const double *cellp = a.box + fi * 9;
        const double4 *__restrict__ pi = a.pos_scr + fi * N + slot0;
        const uint4 *__restrict__ fi32 = a.frac32 + fi * N + slot0;
        const double4 *__restrict__ pj = a.pos_scr + fj * N;
        const uint4 *__restrict__ fj32 = a.frac32 + fj * N;
        cu64_t *fjc = (cu64_t *)(a.frac + fj * N);
        __syncthreads();  // previous t0's candidates are done with fi_lds
        uint32_t abi = 0, ci = 0;
        bool active = false;
        if (slot < N) {
            const uint4 f32 = a.frac32[fi * N + slot];
            fi_lds[threadIdx.x] = f32;
            const uint2 f = a.frac[fi * N + slot];
            abi = f.x;
            ci = f.y;
            if (all_atoms) {
                active = true;
            } else {
                const int o = a.orig_slot[f32.w >> 8];
                active = o >= a.a0 && o < a.a1;
            }
        }
        // strict comparison against threshold+1; inactive lanes compare
        // against 0 and never pass
        const uint32_t th = active ? a.threshold + 1u : 0u;
        const unsigned long long nact = (unsigned long long)__popcll(__ballot(active));
        uadd<TRI>(nall, nact * N);
        __syncthreads();
        uint32_t qn = 0, rn = 0;
        // evaluate the queued entry of this lane; candidates inside the edge
        // window go to the recheck queue (entries j << 8 | i_local, to_j
        // converts this queue's entry), evaluated 64 at a time
        auto eval_with = [&](uint32_t e, bool valid, auto ev, auto to_j) {
            const bool need = valid && ev(e);
            const unsigned long long mr = __ballot(need);
            if (mr) {
                rq[need ? rn + lane_prefix(mr) : (uint32_t)(RQCAP - QSINK) + lane] = to_j(e);
                rn += (uint32_t)__popcll(mr);
                if (rn >= 64) {
                    wave_sync_lds();
                    rn -= 64;
                    uadd<TRI>(nre, 64);
                    const bool ex = recheck_candidate<TRI>(a, t == 0, rq[rn + lane], fi32, fj32, pi, pj, cellp, thr, hist);
                    uadd<TRI>(nex, (uint32_t)__popcll(__ballot(ex)));
                    wave_sync_lds();
                }
            }
        };
        // the j atom's fractions from global memory (16-bit screen, final drain)
        auto eval_batch = [&](uint32_t e, bool valid) {
            eval_with(
                e, valid, [&](uint32_t x) { return eval_candidate<TRI>(a, x, fi_lds, fj32[x >> 8], hist, rsink); },
                [](uint32_t x) { return x; });
        };
        bool clustered = false;
        if constexpr (CLUSTER) {
            uadd<TRI>(nwt, 1);
            const unsigned long long act = __ballot(active);
            if (act) {
                // centre: reference atom + mean wrapped offset of the active lanes
                const int ref = __builtin_ctzll(act);
                const uint32_t ab_r = __shfl(abi, ref, 64), c_r = __shfl(ci, ref, 64);
                const int oa = active ? s16(abi - ab_r) : 0;
                const int ob = active ? s16((abi >> 16) - (ab_r >> 16)) : 0;
                const int oc = active ? s16(ci - c_r) : 0;
                const float inv_n = 1.0f / (float)__popcll(act);
                const uint32_t ca = (ab_r + (uint32_t)__float2int_rn((float)wave_sum_i(oa) * inv_n)) & 0xFFFFu;
                const uint32_t cb = ((ab_r >> 16) + (uint32_t)__float2int_rn((float)wave_sum_i(ob) * inv_n)) & 0xFFFFu;
                const uint32_t cc = (c_r + (uint32_t)__float2int_rn((float)wave_sum_i(oc) * inv_n)) & 0xFFFFu;
                const float fa = (float)s16(abi - ca), fb = (float)s16((abi >> 16) - cb), fc = (float)s16(ci - cc);
                float px, py, pz;
                if (TRI) {
                    pz = a.hz * fc;
                    py = fmaf(a.hyz, fc, a.hy * fb);
                    px = fmaf(a.hxz, fc, fmaf(a.hxy, fb, a.hx * fa));
                } else {
                    px = a.hx * fa;
                    py = a.hy * fb;
                    pz = a.hz * fc;
                }
                const float p2 = fmaf(pz, pz, fmaf(py, py, px * px));
                const float rho = sqrtf(wave_max_f(active ? p2 : 0.0f));
                if (rho < a.rho_cap) {
                    clustered = true;
                    uadd<TRI>(ncl, 1);
                    const float thc = cluster_threshold(rho, a.rmaxf, a.dq);
                    // mode 2: candidates subtract FP32 offsets from the centre;
                    // the rows' offsets replace their fractions in fi_lds, which
                    // only this wavefront reads, with .w = the byte offset of the
                    // row's histogram rows (hist + 4 ti nt nbin)
                    float margin_off = 0.0f;
                    if constexpr (COLC) {
                        margin_off = fmaf(a.du_m1, rho + a.dq, a.du_m0);
                        if (slot < N) {
                            const uint4 f = fi_lds[threadIdx.x];
                            const uint32_t hoff = lds_addr(hist);
                            fi_lds[threadIdx.x] = offset32<TRI>(a, f, ca << 16, cb << 16, cc << 16,
                                                                hoff + (f.w & 0xFFu) * a.nt * a.nbin * 4u);
                        }
                        wave_sync_lds();
                    }
                    // j with |q| > rho + rmax + dq is farther than rmax + dq from
                    // every active i (|q - p| >= |q| - |p|), so none of its pairs
                    // can be counted; the bound carries a margin for the FP32
                    // rounding of |q|^2 and rho
                    const float rs = (rho + a.rmaxf + a.dq) * (1.0f + 0x1p-12f) + 1e-3f;
                    const float rs2 = rs * rs;
                    // rows: (-2p, |p|^2, 1) for active lanes, (0, 0, 0, sentinel, 1) otherwise
                    f16x8_t row = {};
                    if (active) {
                        row[0] = (_Float16)(-2.0f * px);
                        row[1] = (_Float16)(-2.0f * py);
                        row[2] = (_Float16)(-2.0f * pz);
                        row[3] = (_Float16)p2;
                    } else {
                        row[3] = (_Float16)CLUSTER_SENTINEL;
                    }
                    row[4] = (_Float16)1.0f;
                    // A of i rows 0..31 (lanes 0-31 own them) and 32..63 (moved
                    // down from lanes 32-63); k slots 8..15 (lanes 32-63) are 0
                    const uint4 rw = __builtin_bit_cast(uint4, row);
                    uint4 up;
                    up.x = __shfl(rw.x, (int)(lane | 32u), 64);
                    up.y = __shfl(rw.y, (int)(lane | 32u), 64);
                    up.z = __shfl(rw.z, (int)(lane | 32u), 64);
                    up.w = __shfl(rw.w, (int)(lane | 32u), 64);
                    const uint4 z4 = make_uint4(0u, 0u, 0u, 0u);
                    const f16x8_t A0 = __builtin_bit_cast(f16x8_t, lane < 32 ? rw : z4);
                    const f16x8_t A1 = __builtin_bit_cast(f16x8_t, lane < 32 ? up : z4);
                    const uint2 *__restrict__ fj16 = a.frac + fj * N;
                    uint4 *colc = colc_all + (threadIdx.x & ~63u);
                    const uint32_t wbase = threadIdx.x & ~63u;
                    const uint32_t half4 = 4u * (lane >> 5);
                    const bool colane = lane < 32;
                    // the queue as LDS byte addresses: its start, its end (the
                    // count is (qa - qlo) / 4) and this lane's sink entry
                    const uint32_t qlo = lds_addr(q);
                    const uint32_t qsink = qlo + 4u * ((uint32_t)(QC - QSINK) + lane);
                    uint32_t qa = qlo + 4u * qn;
                    // One step: lanes 0-31 build the columns of j = jA + lane, lanes
                    // 32-63 those of j = jB + lane - 32 (when hasB).  Lanes 0-31's
                    // columns form the B operand of half 0, lanes 32-63's (moved
                    // down) that of half 1.
                    // g: the lane's column's 16-bit fractions (zero when !jin)
                    auto step = [&](uint32_t jA, uint32_t jB, bool hasB, const uint2 g) {
                        const uint32_t jl = (colane ? jA : jB) + (lane & 31u);
                        const bool jin = (colane || hasB) && jl < (uint32_t)N;
                        const float ga = (float)s16(g.x - ca), gb = (float)s16((g.x >> 16) - cb);
                        const float gc = (float)s16(g.y - cc);
                        float qx, qy, qz;
                        if (TRI) {
                            qz = a.hz * gc;
                            qy = fmaf(a.hyz, gc, a.hy * gb);
                            qx = fmaf(a.hxz, gc, fmaf(a.hxy, gb, a.hx * ga));
                        } else {
                            qx = a.hx * ga;
                            qy = a.hy * gb;
                            qz = a.hz * gc;
                        }
                        const float q2 = fmaf(qz, qz, fmaf(qy, qy, qx * qx));
                        // a half with no j near the cluster skips its MFMAs and compares
                        const unsigned long long nearm = __ballot(jin && q2 <= rs2);
                        if (nearm == 0ull) return;
                        // COLC: this step's columns' 32-bit fractions go to LDS for
                        // the candidate evaluation (the queue is empty here: each
                        // step drains it); otherwise candidates load them from
                        // global memory
                        if constexpr (COLC) {
                            uint4 o = make_uint4(0u, 0u, 0u, 0u);
                            if (jin) {
                                const uint4 f = fj32[jl];
                                o = offset32<TRI>(a, f, ca << 16, cb << 16, cc << 16, (f.w & 0xFFu) * a.nbin * 4u);
                            }
                            colc[lane] = o;
                            wave_sync_lds();
                        }
                        // COLC: entries hold the LDS byte offsets of the row's
                        // offset (low half) and the column's (high half) instead
                        // of j and i; a recheck entry gets j << 8 | i back
                        auto eval_step = [&](uint32_t e, bool valid) {
                            if constexpr (COLC) {
                                eval_with(
                                    e, valid,
                                    [&](uint32_t x) { return eval_candidate_off(a, x, lds_addr(rsink), margin_off); },
                                    [&](uint32_t x) {
                                        const uint32_t c = (((x >> 16) - lds0) >> 4) - (uint32_t)WG - wbase;
                                        return ((c < 32u ? jA + c : jB + (c - 32u)) << 8) | (((x & 0xFFFFu) - lds0) >> 4);
                                    });
                            } else {
                                eval_batch(e, valid);
                            }
                        };
                        // column (q, 1, |q|^2); past N: (0, 0, 0, 1, sentinel)
                        f16x8_t col = {};
                        col[0] = (_Float16)(jin ? qx : 0.0f);
                        col[1] = (_Float16)(jin ? qy : 0.0f);
                        col[2] = (_Float16)(jin ? qz : 0.0f);
                        col[3] = (_Float16)1.0f;
                        col[4] = (_Float16)(jin ? q2 : CLUSTER_SENTINEL);
                        const uint4 cw = __builtin_bit_cast(uint4, col);
                        uint4 cdn;
                        cdn.x = __shfl(cw.x, (int)(lane ^ 32u), 64);
                        cdn.y = __shfl(cw.y, (int)(lane ^ 32u), 64);
                        cdn.z = __shfl(cw.z, (int)(lane ^ 32u), 64);
                        cdn.w = __shfl(cw.w, (int)(lane ^ 32u), 64);
                        const uint4 z4 = make_uint4(0u, 0u, 0u, 0u);
                        const f16x8_t Bj[2] = {__builtin_bit_cast(f16x8_t, colane ? cw : z4),
                                               __builtin_bit_cast(f16x8_t, colane ? cdn : z4)};
#pragma unroll
                        for (int jb = 0; jb < 2; jb++) {
                            if ((uint32_t)(nearm >> (32 * jb)) == 0u) continue;  // uniform
                            const uint32_t jfirst = jb ? jB : jA;
                            uadd<TRI>(ntile, nact * (unsigned long long)min(32u, (uint32_t)N - jfirst));
                            // entry = j << 8 | local i (COLC: the LDS byte offsets
                            // 16 (WG + wbase + column) << 16 | 16 i, column = 32 jb +
                            // lane % 32); item e of tile T is row
                            // 32 T + 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
                            uint32_t col = COLC ? 32u * (uint32_t)jb + (lane & 31u) : jfirst + (lane & 31u);
                            if constexpr (COLC) asm volatile("" : "+v"(col));  // recomputed per step, not held
                            const uint32_t eb = COLC ? ((lds0 + 16u * ((uint32_t)WG + wbase + col)) << 16) |
                                                           (lds0 + 16u * (wbase + half4))
                                                     : (col << 8) | (wbase + half4);
                            // both row tiles' products are issued before either is
                            // compared, so one MFMA's latency hides behind the other's
                            const f32x16_t DT[2] = {
                                __builtin_amdgcn_mfma_f32_32x32x16_f16(A0, Bj[jb], f32x16_t{}, 0, 0, 0),
                                __builtin_amdgcn_mfma_f32_32x32x16_f16(A1, Bj[jb], f32x16_t{}, 0, 0, 0)};
#pragma unroll
                            for (int T = 0; T < 2; T++) {
                                const f32x16_t D = DT[T];
#pragma unroll
                                for (int grp = 0; grp < 2; grp++) {
                                    bool pc[8];
                                    unsigned long long mk[8];
                                    unsigned long long any = 0;
#pragma unroll
                                    for (int u = 0; u < 8; u++) {
                                        pc[u] = D[8 * grp + u] < thc;
                                        mk[u] = __ballot(pc[u]);
                                        any |= mk[u];
                                    }
                                    if (any) {
#pragma unroll
                                        for (int u = 0; u < 8; u++) {
                                            const int e = 8 * grp + u;
                                            const uint32_t roff = 32u * (uint32_t)T + 8u * (uint32_t)(e >> 2) +
                                                                  (uint32_t)(e & 3);
                                            if constexpr (QADDR) {
                                                ANL_APPEND_A(pc[u], mk[u], eb + (COLC ? roff << 4 : roff))
                                            } else {
                                                ANL_APPEND_E(pc[u], mk[u], eb + (COLC ? roff << 4 : roff))
                                            }
                                        }
                                        // with the column cache, one drain site per 8 compares
                                        // (the queue holds < 64 + 8 masks)
                                        if (COLC) {
                                            if constexpr (QADDR) {
                                                while (qa >= qlo + 256u) {
                                                    wave_sync_lds();
                                                    qa -= 256u;
                                                    const uint32_t e = *(lds_u32_t *)(size_t)(qa + 4u * lane);
                                                    uadd<TRI>(ncand, 64);
                                                    eval_step(e, true);
                                                    wave_sync_lds();
                                                }
                                            } else {
                                                while (qn >= 64) {
                                                    wave_sync_lds();
                                                    qn -= 64;
                                                    const uint32_t e = q[qn + lane];
                                                    uadd<TRI>(ncand, 64);
                                                    eval_step(e, true);
                                                    wave_sync_lds();
                                                }
                                            }
                                        }
                                    }
                                }
                                // otherwise one per 32 x 32 tile (< 64 + 16 masks)
                                if (!COLC) {
                                    if constexpr (QADDR) {
                                        while (qa >= qlo + 256u) {
                                            wave_sync_lds();
                                            qa -= 256u;
                                            const uint32_t e = *(lds_u32_t *)(size_t)(qa + 4u * lane);
                                            uadd<TRI>(ncand, 64);
                                            eval_step(e, true);
                                            wave_sync_lds();
                                        }
                                    } else {
                                        while (qn >= 64) {
                                            wave_sync_lds();
                                            qn -= 64;
                                            const uint32_t e = q[qn + lane];
                                            uadd<TRI>(ncand, 64);
                                            eval_step(e, true);
                                            wave_sync_lds();
                                        }
                                    }
                                }
                            }
                        }
                        // the step's last entries (< 64), so that the next step
                        // can replace the cached columns
                        if constexpr (QADDR) {
                            if (COLC && qa != qlo) {
                                wave_sync_lds();
                                const uint32_t n = (qa - qlo) >> 2;
                                const bool valid = lane < n;
                                const uint32_t e = valid ? q[lane] : 0u;
                                uadd<TRI>(ncand, n);
                                eval_step(e, valid);
                                qa = qlo;
                                wave_sync_lds();
                            }
                        } else {
                            if (COLC && qn) {
                                wave_sync_lds();
                                const bool valid = lane < qn;
                                const uint32_t e = valid ? q[lane] : 0u;
                                uadd<TRI>(ncand, qn);
                                eval_step(e, valid);
                                qn = 0;
                                wave_sync_lds();
                            }
                        }
                    };
                    // Groups of 32 screen slots (= the halves) with a bounding
                    // sphere (centre c_g, radius R_g) per frame.  With e_g =
                    // wrap16(c_g - c): when every component satisfies |e_k| +
                    // R_g 2^16 / w_min < 2^15, each member's wrapped offset is
                    // e_g + its offset from c_g, so |q_j| >= |A e_g| - R_g; a group
                    // with |A e_g| > rs + R_g (plus rounding margins) has no j
                    // that the per-j test above would keep, and is skipped whole.
                    // For any group, with or without that condition, each
                    // component of a member's wrapped offset is at least |e_k| -
                    // R_g 2^16 / w_min in magnitude (a member that wraps to the
                    // other image lies at least 2^15 - R_g 2^16 / w_min away), so
                    // |q_j| >= sqrt(lambda_min(A^T A)) |w| with w those lower
                    // bounds; a group whose bound exceeds rs is skipped too.
                    // The groups left are taken two per step, in order.
                    const uint4 *__restrict__ gt = a.jgroups + (uint64_t)fj * a.ngroups;
                    // near mask of the 64 groups from gbase (one lane per group)
                    auto window_mask = [&](uint32_t gbase) -> unsigned long long {
                        const uint32_t gi = gbase + lane;
                        bool maybe = false;
                        if (gi < a.ngroups) {
                            const uint4 gv = gt[gi];
                            const int ea = s16(gv.x - ca), eb_ = s16((gv.x >> 16) - cb), ec = s16(gv.y - cc);
                            const float R = __uint_as_float(gv.z);
                            const float lim = 32766.0f - R * a.inv_w16;
                            const bool inside = fabsf((float)ea) < lim && fabsf((float)eb_) < lim && fabsf((float)ec) < lim;
                            float gx, gy, gz;
                            if (TRI) {
                                gz = a.hz * (float)ec;
                                gy = fmaf(a.hyz, (float)ec, a.hy * (float)eb_);
                                gx = fmaf(a.hxz, (float)ec, fmaf(a.hxy, (float)eb_, a.hx * (float)ea));
                            } else {
                                gx = a.hx * (float)ea;
                                gy = a.hy * (float)eb_;
                                gz = a.hz * (float)ec;
                            }
                            const float g2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
                            const float far_r = (rs + R) * (1.0f + 0x1p-12f) + 1e-3f;
                            // whatever image a member takes, each component of its
                            // wrapped offset is at least |e_k| - (its offset from c_g)
                            const float dl = R * a.inv_w16 + 2.0f;
                            const float la = fmaxf(fabsf((float)ea) - dl, 0.0f), lb = fmaxf(fabsf((float)eb_) - dl, 0.0f);
                            const float lc = fmaxf(fabsf((float)ec) - dl, 0.0f);
                            const float w2 = fmaf(lc, lc, fmaf(lb, lb, la * la)) * (a.smin16 * a.smin16);
                            const float far_w = rs * (1.0f + 0x1p-12f) + 1e-3f;
                            maybe = !((inside && g2 > far_r * far_r) || w2 > far_w * far_w);
                        }
                        return __ballot(maybe);
                    };
                    // the groups left, two per step in order (uniform state)
                    uint32_t gbase = 0;
                    unsigned long long M = GROUPS ? window_mask(0) : 0ull;
                    auto next_pair = [&](uint32_t &jA, uint32_t &jB, bool &hasB) -> bool {
                        while (M == 0ull) {
                            gbase += 64;
                            if (gbase >= a.ngroups) return false;
                            M = window_mask(gbase);
                        }
                        jA = 32u * (gbase + (uint32_t)__builtin_ctzll(M));
                        M &= M - 1ull;
                        hasB = M != 0ull;
                        jB = 0;
                        if (hasB) {
                            jB = 32u * (gbase + (uint32_t)__builtin_ctzll(M));
                            M &= M - 1ull;
                        }
                        return true;
                    };
                    if constexpr (GROUPS) {
                        uint32_t jA = 0, jB = 0;
                        bool hasB = false;
                        while (next_pair(jA, jB, hasB)) {  // uniform
                            const uint32_t jl = (colane ? jA : jB) + (lane & 31u);
                            uint2 g = make_uint2(0u, 0u);
                            if ((colane || hasB) && jl < (uint32_t)N) g = fj16[jl];
                            step(jA, jB, hasB, g);
                        }
                    } else {
                        uint2 gn = make_uint2(0u, 0u);
                        if (lane < (uint32_t)N) gn = fj16[lane];
                        for (uint32_t j0 = 0; j0 < (uint32_t)N; j0 += 64) {
                            const uint2 g = gn;
                            if (j0 + 64 + lane < (uint32_t)N) gn = fj16[j0 + 64 + lane];
                            step(j0, j0 + 32, true, g);
                        }
                    }
                    if constexpr (QADDR) qn = (qa - qlo) >> 2;  // entries left for the final drain
                }
            }
        }
        if (!clustered) {
            uadd<TRI>(n16, nact * N);
            for (uint32_t j = 0; j < NU; j += UR) {
                unsigned long long g[UR];
                bool p[UR];
                unsigned long long m[UR];
    #pragma unroll
                for (int u = 0; u < UR; u++) g[u] = fjc[j + u];
    #pragma unroll
                for (int u = 0; u < UR; u++) p[u] = sq16(abi, (uint32_t)g[u], ci, (uint32_t)(g[u] >> 32)) < th;
                unsigned long long any = 0;
    #pragma unroll
                for (int u = 0; u < UR; u++) {
                    m[u] = __ballot(p[u]);
                    any |= m[u];
                }
                if (any) {
    #pragma unroll
                    for (int u = 0; u < UR; u++) {
                        ANL_APPEND(p[u], m[u], j + u)
                    }
                    while (qn >= 64) {
                        wave_sync_lds();
                        qn -= 64;
                        const uint32_t e = q[qn + lane];
                        uadd<TRI>(ncand, 64);
                        eval_batch(e, true);
                        wave_sync_lds();
                    }
                }
            }
            for (uint32_t j = NU; j < (uint32_t)N; j++) {  // N % UR tail
                const unsigned long long g0 = fjc[j];
                const bool p0 = sq16(abi, (uint32_t)g0, ci, (uint32_t)(g0 >> 32)) < th;
                const unsigned long long m0 = __ballot(p0);
                ANL_APPEND(p0, m0, j)
            }
        }
        // drain: fewer than 64 + (UNROLL-1)*64 entries, 64 at a time, then the rechecks
        wave_sync_lds();
        for (uint32_t base = 0; base < qn; base += 64) {
            const bool valid = base + lane < qn;
            const uint32_t e = valid ? q[base + lane] : 0u;
            uadd<TRI>(ncand, min(qn - base, 64u));
            eval_batch(e, valid);
        }
        wave_sync_lds();
        bool ex = false;
        if (lane < rn) ex = recheck_candidate<TRI>(a, t == 0, rq[lane], fi32, fj32, pi, pj, cellp, thr, hist);
        uadd<TRI>(nex, (uint32_t)__popcll(__ballot(ex)));
        uadd<TRI>(nre, rn);
        // this lane's atom with itself, binned above in its distinct row
        if (active) move_self_pair<TRI>(a, fi, fj, fi32[il].w, cellp, thr, hist);
        wave_sync_lds();
    }
    // per-wavefront totals, one atomic each per wavefront
    if (lane == 0 && a.ncand) {
        atomicAdd(a.ncand + STAT_CANDIDATES, ncand);
        atomicAdd(a.ncand + STAT_RECHECKS, nre);
        if (nex) atomicAdd(a.ncand + STAT_EXACT, nex);
        if (CLUSTER) {
            atomicAdd(a.ncand + STAT_CLUSTERED, ncl);
            atomicAdd(a.ncand + STAT_UNITS, nwt);
        }
        if (nall) atomicAdd(a.ncand + STAT_PAIRS_ALL, nall);
        if (ntile) atomicAdd(a.ncand + STAT_PAIRS_TILE, ntile);
        if (n16) atomicAdd(a.ncand + STAT_PAIRS_16, n16);
    }
    __syncthreads();
    lds_flush_rows(a, hist, rowch, w.block, t);
}
#undef ANL_APPEND
#undef ANL_APPEND_A

}  // namespace anl

// ---------------------------------------------------------------------------
// host state and C ABI
// ---------------------------------------------------------------------------
using namespace anl;

// Launch shape of the screened kernel: unroll of the j loop, XCD-aware
// workgroup order, pair evaluations per workgroup (sets the t0 chunk), and
// whether compact wavefronts use the cluster (MFMA) screen.
struct GofrtTuning {
    int unroll = 8;
    bool xcd = false;
    uint64_t pairs_per_wg = 1ull << 30;
    // cluster (MFMA) screen where a wavefront's atoms are compact enough: 0 off,
    // 1 on (far groups skipped from the bounding-sphere table), 2 as 1 with
    // each step's columns cached in LDS as FP32 offsets, 3 on with every j
    // visited in order (no table), 4 as 3 with the column cache
    int cluster = 2;
};

struct anl_gofrt {
    anl_traj *tr = nullptr;
    GofrtTuning tune;
    double rmin = 0, rmax = 0, dr = 0, rmin2 = 0, rmax2 = 0;
    uint32_t nbin = 0, lmax = 0, skip = 1, every = 1;
    int form = ANL_GOFRT_AUTO;
    uint32_t n = 0, leff = 0;     // from reset()
    std::vector<double> thr;      // T[0..nbin]
    double lo = 0;
    double *d_thr = nullptr;
    int32_t *d_chan = nullptr;
    unsigned long long *d_ncand = nullptr;   // [NSTAT] statistics of the last launch (Stat)
    int last_form = -1;
    // scratch for the work list
    BlockInfo *d_blocks = nullptr;
    WorkItem *d_work = nullptr;
    size_t cap_blocks = 0, cap_work = 0;
};

namespace {

uint32_t nch_of(const anl_traj *tr) { return tr->nt * (tr->nt + 1); }

int choose_form(anl_gofrt *g) {
    if (g->form == ANL_GOFRT_FP64) return ANL_GOFRT_FP64;
    if (g->tr->screen_rmax != g->rmax) {
        g->tr->screen_cache = make_screen(g->tr->cell0, g->tr->box_const, g->tr->nt, g->rmax);
        g->tr->screen_rmax = g->rmax;
    }
    if (g->tr->screen_cache.valid) return ANL_GOFRT_SCREENED;
    if (g->form == ANL_GOFRT_SCREENED)
        fail(ANL_ERR_ARG, "the screened g(r,t) form cannot be used: " + g->tr->screen_cache.why_not);
    return ANL_GOFRT_FP64;
}

struct Launch {
    std::vector<BlockInfo> blocks;
    std::vector<WorkItem> work;
    uint32_t leff_max = 0;
};

// Splits every block's t0 range into chunks: a chunk bounds the LDS counters
// (each t0 adds at most 256*N to one bin: 256*N*chunk < 2^32) and sets the
// work per workgroup (the tuned number of pair evaluations), finer where the
// launch would not fill the chip.
Launch plan(anl_gofrt *g, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, int single_k) {
    Launch L;
    const uint64_t N = g->tr->N;
    uint64_t cap = (0xFFFFFFFFull / ((uint64_t)WG * N));
    if (cap < 1) cap = 1;
    uint64_t target = g->tune.pairs_per_wg / ((uint64_t)WG * N);
    if (target < 1) target = 1;
    const uint64_t chunk = std::min(cap, target);
    // a launch of fewer than 8 workgroups per CU (2,048 in all) leaves SIMDs
    // idle: a small block splits its t0 range finer (the counts are sums of
    // integers, so the split changes no result)
    const uint64_t ntiles = (N + WG - 1) / WG;
    const uint64_t want = (2048 + nblocks - 1) / nblocks;
    for (uint32_t bi = 0; bi < nblocks; bi++) {
        uint32_t leff = (n[bi] < g->lmax || g->lmax == 0) ? n[bi] : g->lmax;
        L.blocks.push_back(BlockInfo{primo[bi], leff, 0});
        L.leff_max = std::max(L.leff_max, leff);
        if (single_k >= 0) {
            L.work.push_back(WorkItem{bi, (uint32_t)single_k, (uint32_t)single_k + 1, 0});
            continue;
        }
        uint32_t n0 = (n[bi] + g->skip - 1) / g->skip;  // t0 = 0, skip, ... < n
        const uint64_t units = ntiles * std::max<uint64_t>(1, (leff + g->every - 1) / g->every);
        uint64_t chunk_b = chunk;
        if (units * ((n0 + chunk - 1) / chunk) < want) {
            const uint64_t items = (want + units - 1) / units;
            chunk_b = std::min(chunk, std::max<uint64_t>(1, (n0 + items - 1) / items));
        }
        for (uint64_t k = 0; k < n0; k += chunk_b)
            L.work.push_back(WorkItem{bi, (uint32_t)k, (uint32_t)std::min<uint64_t>(n0, k + chunk_b), 0});
    }
    return L;
}

void check_lengths(anl_gofrt *g, const uint64_t *primo, const uint32_t *n, uint32_t nblocks) {
    for (uint32_t bi = 0; bi < nblocks; bi++) {
        uint32_t leff = (n[bi] < g->lmax || g->lmax == 0) ? n[bi] : g->lmax;
        // gofrt.cpp:81-83
        if ((uint64_t)leff + n[bi] + primo[bi] > g->tr->F + 1)
            fail(ANL_ERR_TOO_SHORT,
                 "trajectory is too short for this kind of calculation. Select a different starting timestep or "
                 "lower the size of the average or the lenght of the time lag");
    }
}

}  // namespace

// A launch whose host-side work (form choice, fractional coordinates, work
// list upload, arguments) is done: enqueue_prepared() then only fills the
// output and the counters with kernels and launches the pair kernel, so it
// can be captured into a HIP graph and replayed.
struct GofrtPrepared {
    GofrtArgs a{};
    dim3 grid;
    size_t lds = 0;
    int form = ANL_GOFRT_FP64;
    uint64_t out_len = 0;
    bool triclinic = false;
    int unroll = 4;
    int cluster = 0;  // as GofrtTuning::cluster, 0 when the cluster screen cannot be used
    // device work list and block table owned by this launch (block engine)
    BlockInfo *own_blocks = nullptr;
    WorkItem *own_work = nullptr;
};

__global__ void fill_u64_kernel(unsigned long long *__restrict__ p, uint64_t len) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) p[i] = 0ull;
}

static void zero_u64(unsigned long long *p, uint64_t len, hipStream_t s) {
    if (len == 0) return;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, p, len);
    ANL_HIP(hipGetLastError());
}

// own = true: the work list and block table go to buffers owned by the
// prepared launch (the block engine keeps several prepared launches of one
// g alive and replays them later); otherwise to g's scratch, after the device
// has finished any earlier launch that reads it.
GofrtPrepared gofrt_prepare(anl_gofrt *g, const Launch &L, uint64_t out_len, int a0, int a1, hipStream_t s,
                            uint32_t lag0, uint32_t nlag_force, bool own = false) {
    anl_traj *tr = g->tr;
    GofrtPrepared P;
    const int form = choose_form(g);
    if (form == ANL_GOFRT_SCREENED) build_frac16(tr, s);
    BlockInfo *d_blocks = nullptr;
    WorkItem *d_work = nullptr;
    if (own) {
        ANL_HIP(hipMalloc(&P.own_blocks, std::max<size_t>(1, L.blocks.size()) * sizeof(BlockInfo)));
        d_blocks = P.own_blocks;
        hipError_t e = hipMalloc(&P.own_work, std::max<size_t>(1, L.work.size()) * sizeof(WorkItem));
        if (e != hipSuccess) {
            (void)hipFree(P.own_blocks);
            ANL_HIP(e);
        }
        d_work = P.own_work;
    } else {
        ANL_HIP(hipDeviceSynchronize());
        if (L.blocks.size() > g->cap_blocks) {
            (void)hipFree(g->d_blocks);
            g->d_blocks = nullptr;
            g->cap_blocks = 0;
            ANL_HIP(hipMalloc(&g->d_blocks, L.blocks.size() * sizeof(BlockInfo)));
            g->cap_blocks = L.blocks.size();
        }
        if (L.work.size() > g->cap_work) {
            (void)hipFree(g->d_work);
            g->d_work = nullptr;
            g->cap_work = 0;
            ANL_HIP(hipMalloc(&g->d_work, L.work.size() * sizeof(WorkItem)));
            g->cap_work = L.work.size();
        }
        d_blocks = g->d_blocks;
        d_work = g->d_work;
    }
    ANL_HIP(hipMemcpyAsync(d_blocks, L.blocks.data(), L.blocks.size() * sizeof(BlockInfo), hipMemcpyHostToDevice, s));
    ANL_HIP(hipMemcpyAsync(d_work, L.work.data(), L.work.size() * sizeof(WorkItem), hipMemcpyHostToDevice, s));
    ANL_HIP(hipStreamSynchronize(s));  // the host vectors die with L

    GofrtArgs &a = P.a;
    a.pos = tr->pos;
    a.frac = tr->frac16;
    a.pos_scr = tr->pos_scr;
    a.box = tr->box;
    a.type_slot = tr->type_slot;
    a.orig_slot = tr->orig_slot;
    a.out = nullptr;
    a.ncand = g->d_ncand;
    a.thr = g->d_thr;
    a.chan = g->d_chan;
    a.work = d_work;
    a.blocks = d_blocks;
    a.N = tr->N;
    a.nt = tr->nt;
    a.nbin = g->nbin;
    a.nch = nch_of(tr);
    a.half = tr->nt * (tr->nt + 1) / 2;
    a.out_block_stride = (uint64_t)L.leff_max * a.nch * g->nbin;
    a.skip = g->skip;
    a.every = g->every;
    a.lo = g->lo;
    a.rmax2 = g->rmax2;
    a.Tn = g->thr[g->nbin];
    a.rminf = (float)g->rmin;
    a.inv_drf = (float)(1.0 / g->dr);
    a.threshold = form == ANL_GOFRT_SCREENED ? tr->screen_cache.threshold : 0;
    a.a0 = a0;
    a.a1 = a1;
    a.frac32 = tr->frac32;
    a.nrecheck = g->d_ncand + STAT_RECHECKS;
    {
        const double sc = std::ldexp(1.0, -32);
        const Cell &cl = tr->cell0;
        a.cx = (float)(cl.L[0] * sc);
        a.cxy = (float)(cl.xy * sc);
        a.cxz = (float)(cl.xz * sc);
        a.cy = (float)(cl.L[1] * sc);
        a.cyz = (float)(cl.yz * sc);
        a.cz = (float)(cl.L[2] * sc);
        a.du_margin = form == ANL_GOFRT_SCREENED ? du_window(tr->screen_cache, tr->triclinic, g->rmin, g->rmax, g->dr, g->nbin) : 0.0f;
        a.du_m0 = a.du_m1 = 0.0f;
        if (form == ANL_GOFRT_SCREENED)
            du_window_offsets(tr->screen_cache, g->rmin, g->rmax, g->dr, g->nbin, &a.du_m0, &a.du_m1);
        // cluster screen: cell entries times 2^-16, quantisation bound,
        // largest usable wavefront radius (cluster_params)
        const double sh = std::ldexp(1.0, -16);
        a.hx = (float)(cl.L[0] * sh);
        a.hxy = (float)(cl.xy * sh);
        a.hxz = (float)(cl.xz * sh);
        a.hy = (float)(cl.L[1] * sh);
        a.hyz = (float)(cl.yz * sh);
        a.hz = (float)(cl.L[2] * sh);
        const ClusterParams cp = cluster_params(tr->screen_cache, g->rmax);
        a.dq = (float)cp.dq;
        a.rmaxf = (float)cp.rmaxf;
        a.rho_cap = form == ANL_GOFRT_SCREENED ? (float)cp.rho_cap : 0.0f;
        a.jgroups = tr->jgroups;
        a.ngroups = (uint32_t)((tr->N + 31) / 32);
        a.inv_w16 = form == ANL_GOFRT_SCREENED ? (float)(65536.0 / tr->screen_cache.wmin * (1 + 1e-6)) : 0.0f;
        a.smin16 = form == ANL_GOFRT_SCREENED ? (float)(tr->screen_cache.smin * std::ldexp(1.0, -16) * (1 - 1e-6)) : 0.0f;
    }
    const uint32_t ntiles = (uint32_t)((tr->N + WG - 1) / WG);
    const uint32_t nlag = nlag_force ? nlag_force : (L.leff_max + g->every - 1) / g->every;
    a.lag0 = lag0;
    const uint64_t total = (uint64_t)ntiles * nlag * L.work.size();
    const uint64_t per = (total + 7) / 8;
    if (8 * per >= (1ull << 32)) fail(ANL_ERR_ARG, "too many workgroups (reduce the trajectory block)");
    a.ntiles = ntiles;
    a.nlag = nlag;
    a.nwork = (uint32_t)L.work.size();
    const bool remap = g->tune.xcd && form == ANL_GOFRT_SCREENED;
    a.per_xcd = remap ? (uint32_t)per : 0u;
    P.grid = dim3((uint32_t)(remap ? 8 * per : total));
    // channel table (FP64 form) or histogram-offset table (screened form, two
    // entries per type pair), bin-edge table, histogram
    const size_t common = (g->nbin + 1) * sizeof(double) +
                          (size_t)(form == ANL_GOFRT_FP64 ? 1 : 2) * tr->nt * tr->nt * sizeof(int) +
                          (size_t)a.nch * g->nbin * sizeof(unsigned);
    auto screened_lds = [&](int qcap, bool cluster) {
        return (cluster ? 2 : 1) * WG * sizeof(uint4) + (WG / 64) * (qcap + RQCAP) * sizeof(uint32_t) + common;
    };
    // the cluster screen needs the larger queue; without room it is not used
    const bool colc = g->tune.cluster == 2 || g->tune.cluster == 4;
    const int qcl = colc ? SQCAP_CLC : SQCAP_CL;
    const bool cluster_ok = g->tune.cluster != 0 && form == ANL_GOFRT_SCREENED && a.rho_cap > 0 &&
                            screened_lds(qcl, colc) <= 160 * 1024;
    P.lds = form == ANL_GOFRT_FP64 ? common : (cluster_ok ? screened_lds(qcl, colc) : screened_lds(SQCAP, false));
    if (P.lds > 160 * 1024) fail(ANL_ERR_ARG, "histogram does not fit in LDS (reduce nbin or types)");
    P.form = form;
    P.out_len = out_len;
    P.triclinic = tr->triclinic;
    P.unroll = g->tune.unroll;
    P.cluster = cluster_ok ? g->tune.cluster : 0;
    g->last_form = form;
    return P;
}

void gofrt_enqueue(anl_gofrt *g, const GofrtPrepared &P, unsigned long long *out, hipStream_t s) {
    GofrtArgs a = P.a;
    a.out = out;
    zero_u64(out, P.out_len, s);
    zero_u64(g->d_ncand, NSTAT, s);
    if (P.grid.x == 0) return;
    if (P.form == ANL_GOFRT_FP64) {
        if (P.triclinic) hipLaunchKernelGGL(gofrt_fp64_kernel<true>, P.grid, dim3(WG), P.lds, s, a);
        else hipLaunchKernelGGL(gofrt_fp64_kernel<false>, P.grid, dim3(WG), P.lds, s, a);
    } else {
        const bool u8 = P.unroll == 8;
#define ANL_SCREENED(T, U, C, K) hipLaunchKernelGGL((gofrt_screened_kernel<T, U, C, K>), P.grid, dim3(WG), P.lds, s, a)
        if (P.cluster) {  // the 16-bit fallback of non-compact wavefronts takes 8 j per step
            switch (P.cluster * 2 + (P.triclinic ? 1 : 0)) {
                case 2: ANL_SCREENED(false, 8, true, 1); break;
                case 3: ANL_SCREENED(true, 8, true, 1); break;
                case 4: ANL_SCREENED(false, 8, true, 2); break;
                case 5: ANL_SCREENED(true, 8, true, 2); break;
                case 6: ANL_SCREENED(false, 8, true, 3); break;
                case 7: ANL_SCREENED(true, 8, true, 3); break;
                case 8: ANL_SCREENED(false, 8, true, 4); break;
                default: ANL_SCREENED(true, 8, true, 4); break;
            }
        } else {
            if (P.triclinic) {
                if (u8) ANL_SCREENED(true, 8, false, 0);
                else ANL_SCREENED(true, 4, false, 0);
            } else {
                if (u8) ANL_SCREENED(false, 8, false, 0);
                else ANL_SCREENED(false, 4, false, 0);
            }
        }
#undef ANL_SCREENED
    }
    ANL_HIP(hipGetLastError());
}

// Internal API of the block engine (blocks.hip), declared in anl_internal.h.
namespace anl {
void *gofrt_prepare_blocks(anl_gofrt *g, const uint64_t *primo, const uint32_t *n, uint32_t nblocks,
                           uint32_t *leff_max, hipStream_t s) {
    check_lengths(g, primo, n, nblocks);
    Launch L = plan(g, primo, n, nblocks, -1);
    const uint64_t len = (uint64_t)nblocks * L.leff_max * nch_of(g->tr) * g->nbin;
    auto *P = new GofrtPrepared(gofrt_prepare(g, L, len, INT32_MIN, INT32_MAX, s, 0, 0, true));
    if (leff_max) *leff_max = L.leff_max;
    return P;
}
void gofrt_enqueue_prepared(anl_gofrt *g, void *prep, unsigned long long *out, hipStream_t s) {
    gofrt_enqueue(g, *(GofrtPrepared *)prep, out, s);
}
void gofrt_free_prepared(void *prep) {
    auto *P = (GofrtPrepared *)prep;
    if (!P) return;
    (void)hipFree(P->own_blocks);
    (void)hipFree(P->own_work);
    delete P;
}
int gofrt_form_of(void *prep) { return ((GofrtPrepared *)prep)->form; }
uint32_t gofrt_block_leff(const anl_gofrt *g, uint32_t n) { return (n < g->lmax || g->lmax == 0) ? n : g->lmax; }
uint32_t gofrt_nbin(const anl_gofrt *g) { return g->nbin; }
uint32_t gofrt_skip(const anl_gofrt *g) { return g->skip; }
}  // namespace anl

namespace {

void launch(anl_gofrt *g, const Launch &L, unsigned long long *out, uint64_t out_len, int a0, int a1,
            hipStream_t s, uint32_t lag0 = 0, uint32_t nlag_force = 0) {
    GofrtPrepared P = gofrt_prepare(g, L, out_len, a0, a1, s, lag0, nlag_force);
    gofrt_enqueue(g, P, out, s);
}

}  // namespace

extern "C" {

int anl_gofrt_create(anl_traj *tr, double rmin, double rmax, uint32_t nbin, uint32_t tmax, uint32_t skip,
                     uint32_t every, int form, anl_gofrt **out) {
    return guard([&] {
        if (!tr || !out) fail(ANL_ERR_ARG, "null pointer passed to anl_gofrt_create");
        if (nbin == 0) fail(ANL_ERR_ARG, "nbin must be positive");
        if (!(rmax > rmin) || rmin < 0) fail(ANL_ERR_ARG, "need 0 <= rmin < rmax");
        if (form < ANL_GOFRT_FP64 || form > ANL_GOFRT_AUTO) fail(ANL_ERR_ARG, "unknown g(r,t) form");
        ensure_device();
        std::unique_ptr<anl_gofrt, int (*)(anl_gofrt *)> owner(new anl_gofrt(), anl_gofrt_destroy);
        anl_gofrt *g = owner.get();
        g->tr = tr;
        g->rmin = rmin;
        g->rmax = rmax;
        g->nbin = nbin;
        g->lmax = tmax;
        g->skip = skip == 0 ? 1 : skip;     // calculatemultithread.h:43-46
        g->every = every == 0 ? 1 : every;
        g->form = form;
        if (const char *e = std::getenv("ANL_GOFRT_UNROLL")) g->tune.unroll = std::atoi(e) == 8 ? 8 : 4;
        if (const char *e = std::getenv("ANL_GOFRT_XCD")) g->tune.xcd = std::atoi(e) != 0;
        if (const char *e = std::getenv("ANL_GOFRT_CLUSTER")) g->tune.cluster = std::min(4, std::max(0, std::atoi(e)));
        if (const char *e = std::getenv("ANL_GOFRT_PAIRS_PER_WG")) {
            long long v = std::atoll(e);
            if (v > 0) g->tune.pairs_per_wg = (uint64_t)v;
        }
        g->dr = (rmax - rmin) / (double)nbin;  // gofrt.cpp:27-29
        g->rmax2 = rmax * rmax;
        g->rmin2 = rmin * rmin;
        g->thr = bin_edges(rmin, rmax, nbin, g->dr);
        g->lo = std::max(g->rmin2, g->thr[0]);
        {
            // bin edges, then the FP64 window constants of the screened form
            // (zero when the screen does not apply to this cell and rmax)
            std::vector<double> t(g->thr);
            t.resize(nbin + 1 + NFP64K, 0.0);
            const Screen sc = make_screen(tr->cell0, tr->box_const, tr->nt, rmax);
            if (sc.valid) {
                const Cell &cl = tr->cell0;
                const double f = std::ldexp(1.0, -32);
                double *k = t.data() + nbin + 1;
                k[FP64K_X] = cl.L[0] * f;
                k[FP64K_XY] = cl.xy * f;
                k[FP64K_XZ] = cl.xz * f;
                k[FP64K_Y] = cl.L[1] * f;
                k[FP64K_YZ] = cl.yz * f;
                k[FP64K_Z] = cl.L[2] * f;
                k[FP64K_EPS] = d2_window_fp64(sc, cl, tr->pos_max, rmax);
            }
            ANL_HIP(hipMalloc(&g->d_thr, t.size() * sizeof(double)));
            ANL_HIP(hipMemcpy(g->d_thr, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        // gofrt.h:86-104: channel of (type1, type2)
        std::vector<int32_t> chan(tr->nt * tr->nt);
        for (uint32_t a = 0; a < tr->nt; a++)
            for (uint32_t b = 0; b < tr->nt; b++) {
                uint32_t t1 = std::min(a, b), t2 = std::max(a, b);
                chan[a * tr->nt + b] = (int32_t)(tr->nt * (tr->nt + 1) / 2 - (t2 + 1) * (t2 + 2) / 2 + t1);
            }
        ANL_HIP(hipMalloc(&g->d_chan, chan.size() * sizeof(int32_t)));
        ANL_HIP(hipMemcpy(g->d_chan, chan.data(), chan.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        ANL_HIP(hipMalloc(&g->d_ncand, NSTAT * sizeof(unsigned long long)));
        *out = owner.release();
    });
}

int anl_gofrt_destroy(anl_gofrt *g) {
    return guard([&] {
        if (!g) return;
        (void)hipFree(g->d_thr);
        (void)hipFree(g->d_chan);
        (void)hipFree(g->d_ncand);
        (void)hipFree(g->d_blocks);
        (void)hipFree(g->d_work);
        delete g;
    });
}

int anl_gofrt_reset(anl_gofrt *g, uint32_t n, uint32_t *leff_out) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        g->n = n;
        g->leff = (n < g->lmax || g->lmax == 0) ? n : g->lmax;  // gofrt.cpp:56
        if (leff_out) *leff_out = g->leff;
    });
}

int anl_gofrt_calculate_async(anl_gofrt *g, uint64_t primo, uint64_t *counts_dev, void *stream) {
    return guard([&] {
        if (!g || !counts_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_gofrt_calculate_async");
        ensure_device();
        check_lengths(g, &primo, &g->n, 1);
        Launch L = plan(g, &primo, &g->n, 1, -1);
        if (L.leff_max == 0) return;
        uint64_t len = (uint64_t)L.leff_max * nch_of(g->tr) * g->nbin;
        launch(g, L, (unsigned long long *)counts_dev, len, INT32_MIN, INT32_MAX, (hipStream_t)stream);
    });
}

int anl_gofrt_calculate(anl_gofrt *g, uint64_t primo, uint64_t *counts_host) {
    return guard([&] {
        if (!g || !counts_host) fail(ANL_ERR_ARG, "null pointer passed to anl_gofrt_calculate");
        ensure_device();
        check_lengths(g, &primo, &g->n, 1);
        uint64_t len = (uint64_t)g->leff * nch_of(g->tr) * g->nbin;
        if (len == 0) return;
        unsigned long long *d = nullptr;
        ANL_HIP(hipMalloc(&d, len * sizeof(unsigned long long)));
        try {
            Launch L = plan(g, &primo, &g->n, 1, -1);
            launch(g, L, d, len, INT32_MIN, INT32_MAX, 0);
            ANL_HIP(hipMemcpy(counts_host, d, len * sizeof(unsigned long long), hipMemcpyDeviceToHost));

        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        (void)hipFree(d);
    });
}

int anl_gofrt_calc_single_th(anl_gofrt *g, int t, int imedia, int a0, int a1, uint64_t primo,
                             uint64_t *counts_host) {
    return guard([&] {
        if (!g || !counts_host) fail(ANL_ERR_ARG, "null pointer passed to anl_gofrt_calc_single_th");
        if (t < 0 || (uint32_t)t >= g->leff || imedia < 0) fail(ANL_ERR_ARG, "lag or t0 out of range");
        if (primo + (uint64_t)imedia + (uint64_t)t >= g->tr->F) fail(ANL_ERR_TOO_SHORT, "frame out of range");
        ensure_device();
        const uint64_t len = (uint64_t)g->leff * nch_of(g->tr) * g->nbin;
        // one t0 (frame primo+imedia) and one lag t
        Launch L;
        L.leff_max = g->leff;
        L.blocks.push_back(BlockInfo{primo + (uint64_t)imedia, g->leff, 0});
        L.work.push_back(WorkItem{0, 0, 1, 0});
        unsigned long long *d = nullptr;
        ANL_HIP(hipMalloc(&d, len * sizeof(unsigned long long)));
        std::vector<uint64_t> h(len);
        try {
            launch(g, L, d, len, a0, a1, 0, (uint32_t)t, 1);
            ANL_HIP(hipMemcpy(h.data(), d, len * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        (void)hipFree(d);
        for (uint64_t k = 0; k < len; k++) counts_host[k] += h[k];
    });
}

int anl_gofrt_calculate_blocks_async(anl_gofrt *g, const uint64_t *primo, const uint32_t *n, uint32_t nblocks,
                                     uint64_t *counts_dev, void *stream) {
    return guard([&] {
        if (!g || !primo || !n || !counts_dev || nblocks == 0)
            fail(ANL_ERR_ARG, "bad arguments to anl_gofrt_calculate_blocks_async");
        ensure_device();
        check_lengths(g, primo, n, nblocks);
        Launch L = plan(g, primo, n, nblocks, -1);
        uint64_t len = (uint64_t)nblocks * L.leff_max * nch_of(g->tr) * g->nbin;
        launch(g, L, (unsigned long long *)counts_dev, len, INT32_MIN, INT32_MAX, (hipStream_t)stream);
    });
}

int anl_gofrt_last_stats(anl_gofrt *g, int *form_used, uint64_t *candidates, uint64_t *rechecks) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        if (form_used) *form_used = g->last_form;
        unsigned long long c[2] = {0, 0};
        ANL_HIP(hipMemcpy(c, g->d_ncand, sizeof(c), hipMemcpyDeviceToHost));
        if (candidates) *candidates = c[0];
        if (rechecks) *rechecks = c[1];
    });
}

int anl_gofrt_recheck_stats(anl_gofrt *g, uint64_t *window, uint64_t *exact) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        unsigned long long c[NSTAT] = {};
        ANL_HIP(hipMemcpy(c, g->d_ncand, sizeof(c), hipMemcpyDeviceToHost));
        if (window) *window = c[STAT_RECHECKS];
        if (exact) *exact = c[STAT_EXACT];
    });
}

int anl_gofrt_cluster_stats(anl_gofrt *g, uint64_t *clustered, uint64_t *units) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        unsigned long long c[4] = {0, 0, 0, 0};
        ANL_HIP(hipMemcpy(c, g->d_ncand, sizeof(c), hipMemcpyDeviceToHost));
        if (clustered) *clustered = c[2];
        if (units) *units = c[3];
    });
}

int anl_gofrt_pair_stats(anl_gofrt *g, uint64_t *evaluated, uint64_t *all_pairs, uint64_t *tile, uint64_t *screen16,
                         uint64_t *fp64) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        unsigned long long c[NSTAT] = {};
        ANL_HIP(hipMemcpy(c, g->d_ncand, sizeof(c), hipMemcpyDeviceToHost));
        if (evaluated) *evaluated = c[STAT_PAIRS_TILE] + c[STAT_PAIRS_16] + c[STAT_PAIRS_FP64];
        if (all_pairs) *all_pairs = c[STAT_PAIRS_ALL];
        if (tile) *tile = c[STAT_PAIRS_TILE];
        if (screen16) *screen16 = c[STAT_PAIRS_16];
        if (fp64) *fp64 = c[STAT_PAIRS_FP64];
    });
}

int anl_gofrt_screen_info(const double *box9, double rmax, int *valid, uint32_t *threshold, double *rho_cap,
                          double *dq, double *q2_max, double *threshold_max) {
    return guard([&] {
        if (!box9) fail(ANL_ERR_ARG, "null box");
        const Cell c = cell_of(box9);
        const Screen scr = make_screen(c, true, 1, rmax);
        const ClusterParams cp = cluster_params(scr, rmax);
        if (valid) *valid = scr.valid ? 1 : 0;
        if (threshold) *threshold = scr.valid ? scr.threshold : 0;
        if (rho_cap) *rho_cap = cp.rho_cap;
        if (dq) *dq = cp.dq;
        // what the kernel compares for the widest wavefront it clusters: the
        // largest |q|^2 of a pair it must pass, and the distance threshold
        const float rho = (float)cp.rho_cap;
        const float rs = (rho + (float)cp.rmaxf + (float)cp.dq) * (1.0f + 0x1p-12f) + 1e-3f;
        if (q2_max) *q2_max = cp.rho_cap > 0 ? (double)rs * (double)rs : 0.0;
        if (threshold_max)
            *threshold_max = cp.rho_cap > 0 ? (double)cluster_threshold(rho, (float)cp.rmaxf, (float)cp.dq) : 0.0;
    });
}

int anl_gofrt_set_tuning(anl_gofrt *g, int unroll, int xcd, uint64_t pairs_per_wg, int cluster) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        if (unroll != 4 && unroll != 8) fail(ANL_ERR_ARG, "unroll must be 4 or 8");
        if (pairs_per_wg == 0) fail(ANL_ERR_ARG, "pairs_per_wg must be positive");
        g->tune.unroll = unroll;
        g->tune.xcd = xcd != 0;
        g->tune.pairs_per_wg = pairs_per_wg;
        if (cluster < 0 || cluster > 4) fail(ANL_ERR_ARG, "cluster must be 0, 1, 2, 3 or 4");
        g->tune.cluster = cluster;
    });
}

int anl_gofrt_get_tuning(anl_gofrt *g, int *unroll, int *xcd, uint64_t *pairs_per_wg, int *cluster) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        if (unroll) *unroll = g->tune.unroll;
        if (xcd) *xcd = g->tune.xcd ? 1 : 0;
        if (pairs_per_wg) *pairs_per_wg = g->tune.pairs_per_wg;
        if (cluster) *cluster = g->tune.cluster;
    });
}

// Times every candidate launch shape on the first nlag_probe lags of the
// calculation set by reset(n) at primo (one untimed launch, then one timed
// launch per candidate, HIP events on the null stream) and keeps the fastest.
int anl_gofrt_autotune(anl_gofrt *g, uint64_t primo, uint32_t nlag_probe, const int32_t *cand, uint32_t ncand,
                       float *ms_out, uint32_t *best_out) {
    return guard([&] {
        if (!g) fail(ANL_ERR_ARG, "null handle");
        ensure_device();
        check_lengths(g, &primo, &g->n, 1);
        // unroll, xcd, log2 pairs per workgroup, cluster
        static const int32_t defaults[] = {8, 0, 29, 2, 8, 0, 30, 2, 8, 0, 31, 2, 8, 0, 29, 1, 8, 0, 29, 3,
                                           8, 0, 29, 4, 8, 0, 30, 4, 8, 1, 29, 2, 8, 0, 30, 0};
        if (!cand) {
            cand = defaults;
            ncand = sizeof(defaults) / sizeof(defaults[0]) / 4;
        }
        if (ncand == 0) fail(ANL_ERR_ARG, "no candidates");
        for (uint32_t c = 0; c < ncand; c++)
            if ((cand[4 * c] != 4 && cand[4 * c] != 8) || cand[4 * c + 2] < 10 || cand[4 * c + 2] > 40 ||
                cand[4 * c + 3] < 0 || cand[4 * c + 3] > 4)
                fail(ANL_ERR_ARG, "candidate " + std::to_string(c) +
                                      ": unroll 4|8, log2(pairs per workgroup) in [10, 40], cluster 0..4");
        // probe lags: at least nlag_probe, and enough for >= 4096 workgroups
        // (several waves of residency), so that the probe behaves like the run
        const uint32_t nlag_all = (g->leff + g->every - 1) / g->every;
        const uint64_t per_lag = (uint64_t)((g->tr->N + WG - 1) / WG) * plan(g, &primo, &g->n, 1, -1).work.size();
        const uint64_t fill = (4096 + per_lag - 1) / std::max<uint64_t>(per_lag, 1);
        const uint32_t nlag = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(std::max<uint64_t>(nlag_probe, fill), nlag_all));
        const uint64_t len = (uint64_t)g->leff * nch_of(g->tr) * g->nbin;
        if (len == 0) return;
        const GofrtTuning saved = g->tune;
        unsigned long long *d = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ANL_HIP(hipMalloc(&d, len * sizeof(unsigned long long)));
        float best = -1.0f;
        uint32_t besti = 0;
        auto set = [&](uint32_t c) {
            g->tune.unroll = cand[4 * c];
            g->tune.xcd = cand[4 * c + 1] != 0;
            g->tune.pairs_per_wg = 1ull << cand[4 * c + 2];
            g->tune.cluster = cand[4 * c + 3];
        };
        try {
            ANL_HIP(hipEventCreate(&e0));
            ANL_HIP(hipEventCreate(&e1));
            for (uint32_t c = 0; c < ncand; c++) {
                set(c);
                Launch L = plan(g, &primo, &g->n, 1, -1);
                GofrtPrepared P = gofrt_prepare(g, L, len, INT32_MIN, INT32_MAX, 0, 0, nlag);
                gofrt_enqueue(g, P, d, 0);
                ANL_HIP(hipEventRecord(e0, 0));
                gofrt_enqueue(g, P, d, 0);
                ANL_HIP(hipEventRecord(e1, 0));
                ANL_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                ANL_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms_out) ms_out[c] = ms;
                if (best < 0.0f || ms < best) {
                    best = ms;
                    besti = c;
                }
            }
        } catch (...) {
            g->tune = saved;
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipFree(d);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(d);
        set(besti);
        if (best_out) *best_out = besti;
    });
}

}  // extern "C"
