// Internal declarations shared by the HIP translation units of libanalisi_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/analisi_hip.h"

namespace anl {

// Thread-local message behind anl_last_error().
void set_error(const std::string &msg);
void clear_error();

// Status helpers: every C-ABI body runs inside guard(), which converts
// exceptions into status codes and the message.
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string &msg);
void check_hip(hipError_t e, const char *what);
#define ANL_HIP(call) ::anl::check_hip((call), #call)

template <class F>
int guard(F &&f) {
    try {
        clear_error();
        f();
        return ANL_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        set_error(e.what());
        return ANL_ERR_ARG;
    }
}

void ensure_device();

// Cell parameters of one frame in the internal format (basetrajectory.h:21-43):
// origin, half lengths, full lengths (2*half, exact) and tilts xy, xz, yz.
struct Cell {
    double lo[3];
    double lh[3];
    double L[3];
    double xy, xz, yz;
};

// g(r,t) screen parameters (16-bit fractional fixed point; see gofrt.hip).
struct Screen {
    bool valid;            // preconditions hold for the whole trajectory
    uint32_t threshold;    // |q|^2 <= threshold passes
    std::string why_not;   // reason when !valid
    double row_sum;        // max over rows of the cell matrix of sum |A_ik|
    double wmin;           // smallest cell width (distance between opposite faces)
    double fro;            // Frobenius norm of the cell matrix (>= its spectral norm)
    double smin;           // sqrt of the smallest eigenvalue of A^T A (|A x| >= smin |x|)
    double anorm;          // spectral norm of |A| (entry-wise absolute values), rounded up
};

}  // namespace anl

// Device-resident trajectory.  Atoms are stored in "slots" sorted along a
// Morton curve of the frame-0 fractional position, so that a wavefront's 64
// atoms are spatial neighbours (coherent branches in the pair kernels).
// Counts and per-type sums do not depend on this order.
struct anl_traj {
    int device = 0;
    uint64_t F = 0, N = 0;
    uint32_t nt = 0;
    uint32_t stride = 6;
    bool triclinic = false;
    bool box_const = false;             // every frame has bitwise the same box
    double pos_max = 0.0;               // largest |position component| of the trajectory
    double4 *pos = nullptr;             // [F][N] slots, w = 0
    double *box = nullptr;              // [F][9] internal format, tilts 0 if ortho
    int32_t *type_slot = nullptr;       // [N]
    int32_t *orig_slot = nullptr;       // [N] original atom index of a slot
    double *cm = nullptr;               // [F][nt][3]
    uint2 *frac16 = nullptr;            // [F][N] {a | b<<16, c} 16-bit fractions, lazily built
    uint4 *frac32 = nullptr;            // [F][N] {a, b, c, type | slot << 8} 32-bit fractions, lazily built
    int32_t *frac_order = nullptr;      // [F][N] slot held by each screen slot, until the fractions are built
    double4 *pos_scr = nullptr;         // [F][N] positions in each frame's screen order (with the fractions)
    uint4 *jgroups = nullptr;           // [F][ceil(N/32)] bounding spheres of 32 consecutive screen slots:
                                        // {centre a | b << 16, centre c, radius (float bits, A), 0}
    std::vector<double> box_host;       // [F][9]
    std::vector<int32_t> type_of_slot;  // host copy
    std::vector<uint64_t> type_count;   // atoms per type
    anl::Cell cell0;                    // cell of frame 0
    anl::Screen screen_cache;
    double screen_rmax = -1.0;
    uint32_t msd_tile = 16;             // lag tile of the MSD window kernel (anl_msd_set_tuning)
    uint32_t msd_waves = 4;             // its wavefronts per workgroup
};

struct anl_series {
    int device = 0;
    uint64_t nrows = 0;
    uint32_t ncols = 0;
    double *data = nullptr;
    // Green-Kubo scratch (splits, partial sums, output staging), grow-only;
    // calls on one series are stream-ordered on the caller's stream.
    void *work = nullptr;
    size_t work_cap = 0;
    // launch shape of the MFMA lag products: tiles per wavefront, MFMA steps
    // per FP32 chunk, wavefronts per workgroup (anl_gk_set_tuning)
    uint32_t gk_q = 8, gk_cs = 16, gk_wg = 4;
};

namespace anl {
Cell cell_of(const double *box9);
// T[k], k = 0..nbin: the smallest non-negative d2 whose bin
// (int)floorf((float)((sqrt(d2) - rmin) / dr)) is >= k, computed on the host
// with IEEE double arithmetic (gofrt.cpp:114-115, sphericalbase.cpp:52-53).
std::vector<double> bin_edges(double rmin, double rmax, uint32_t nbin, double dr);
void build_frac16(anl_traj *tr, hipStream_t s);

// Prepared launches for the block engine: all host work (checks, plans,
// uploads, lazy data) happens in *_prepare_*; *_enqueue_prepared only
// launches kernels (zero fills included), so it can be captured in a graph.
void *gofrt_prepare_blocks(anl_gofrt *g, const uint64_t *primo, const uint32_t *n, uint32_t nblocks,
                           uint32_t *leff_max, hipStream_t s);
void gofrt_enqueue_prepared(anl_gofrt *g, void *prep, unsigned long long *out, hipStream_t s);
void gofrt_free_prepared(void *prep);
int gofrt_form_of(void *prep);
uint32_t gofrt_block_leff(const anl_gofrt *g, uint32_t n);
uint32_t gofrt_nbin(const anl_gofrt *g);
uint32_t gofrt_skip(const anl_gofrt *g);
void *msd_prepare_blocks(anl_traj *tr, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, uint32_t skip,
                         uint32_t tmax, int cm_msd, int cm_self, uint32_t *leff_out);
void msd_enqueue_prepared(void *prep, double *out, hipStream_t s);
void msd_free_prepared(void *prep);
// One MSD calculation (reset(n) + calculate(primo)) with its block table,
// term counts and partial sums in caller-provided device memory; fails with
// ANL_ERR_ARG when they do not fit in scratch_bytes.  Allocates nothing.
void msd_calculate_in_scratch(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n,
                              uint64_t primo, double *out, unsigned char *scratch, size_t scratch_bytes,
                              hipStream_t s);
// Scratch of at least bytes in s->work (256-byte aligned sub-buffers are
// carved by the caller); growing synchronises the device first.
void *series_workspace(anl_series *s, size_t bytes);
// FP64 Green-Kubo epilogue (mean subtraction, running integrals, Schur
// complement) over d[leff][narr] whose first ncur^2 columns hold C(l); the
// mean and the integrals as the reference's serial scans, or as parallel
// sums (different rounding) when serial_scans is false.
void gk_epilogue(double *d, uint32_t narr, uint32_t ncur, uint32_t leff, int subtract_mean, uint32_t start_mean,
                 hipStream_t s, uint32_t nblocks, uint64_t stride, bool serial_scans);
// Lag products of every block (split-BF16 MFMA) into d + b * leff * narr:
// C(l) in the first ncur^2 columns, zeros elsewhere; scratch from the series.
void gk_mfma_enqueue_blocks(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t leff,
                            uint32_t n_seg, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, double *d,
                            hipStream_t st);
size_t gk_mfma_scratch_bytes(const anl_series *s, uint32_t ncur, uint32_t leff, uint32_t n);
bool gkm_shape_ok(uint32_t q, uint32_t cs, uint32_t wg);
// MediaVar over blocks (calcoliblocchi.h:21-65) of x_b = counts_b * incr_b or
// values_b, element-parallel, Welford steps in block order.
void mediavar_blocks(const unsigned long long *counts, const double *values, const double *incr, uint32_t nblocks,
                     uint64_t len, double *mean, double *var, hipStream_t s);
}  // namespace anl
