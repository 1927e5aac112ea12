/*
 * analisi_hip.h -- C ABI of the MI355X (gfx950) time-lagged pair-correlation
 * path: g(r,t) histograms, mean square displacement, Green-Kubo lag
 * correlations, spherical-harmonic density correlations and block statistics.
 *
 * Plain C types only (pointers, sizes, status codes).  Every call returns an
 * ANL_* status; on failure anl_last_error() returns the message, which the
 * C++ operator classes rethrow as std::runtime_error with the reference's
 * wording (gofrt.cpp:81-83, msd.cpp:54-60, greenkuboNcomponentionicfluid.cpp:
 * 168-169, :67).  Calls are synchronous unless named *_async.  A handle is
 * not re-entrant: one calculation at a time per handle, as in the reference
 * (th_data is a member, calculatemultithread.h:106-162).
 *
 * Each entry point names the reference interface it replaces (file:line of
 * rikigigi/analisi v0.5.2).
 */
#ifndef ANALISI_HIP_H
#define ANALISI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ANL_OK 0
#define ANL_ERR_ARG 1          /* invalid argument (shape, null pointer, sizes) */
#define ANL_ERR_TOO_SHORT 2    /* trajectory / series too short for (primo, n, lags) */
#define ANL_ERR_NOT_LOADED 3   /* MSD: not enough loaded timesteps (msd.cpp:56-59) */
#define ANL_ERR_HIP 4          /* HIP runtime error */
#define ANL_ERR_NO_DEVICE 5    /* no gfx950 device visible */
#define ANL_ERR_HEADER 6       /* GK header not found (greenkubo...cpp:62-69) */

/* g(r,t) arithmetic forms */
#define ANL_GOFRT_FP64 0       /* plain FP64 evaluation of every ordered pair */
#define ANL_GOFRT_SCREENED 1   /* 16-bit fixed-point screen of every pair, exact FP64
                                  evaluation of every candidate (same counts) */
#define ANL_GOFRT_AUTO 2       /* SCREENED when its preconditions hold, else FP64 */

/* Message of the last failed call on this thread ("" if none). */
const char *anl_last_error(void);

/* Select and initialise the HIP device (default 0).  Fails with
 * ANL_ERR_NO_DEVICE when no gfx950 device is present. */
int anl_init(int device);

/* Build/version string (replaces pyanalisi.info(), config.h.in:49-64). */
const char *anl_info(void);

/* ---- trajectory mirror ------------------------------------------------- */
/* Device-resident copy of what BaseTrajectory serves to the calculations:
 * positions<false>(t,i) (trajectory_numpy.h:18-19), box<false>(t) in the
 * internal format (basetrajectory.h:21-43, :94-105), get_type(i) ids
 * (basetrajectory.h:71-77, basetrajectory.cpp:51-89) and positions_cm(t,type)
 * (trajectory_numpy.cpp:201-225).  The host arrays are read once; the mirror
 * is immutable afterwards (Trajectory_numpy data never changes after construction).
 *   pos      [nframes][natoms][3] FP64
 *   box      [nframes][box_stride] FP64, internal format, box_stride 6 or 9
 *   type_id  [natoms] int32, dense ids 0..ntypes-1
 *   pos_cm   [nframes][ntypes][3] FP64 (may be NULL when MSD is not used) */
typedef struct anl_traj anl_traj;
int anl_traj_create(const double *pos, const double *box, uint32_t box_stride,
                    const int32_t *type_id, uint32_t ntypes, const double *pos_cm,
                    uint64_t nframes, uint64_t natoms, anl_traj **out);
int anl_traj_destroy(anl_traj *tr);

/* ---- g(r,t) ------------------------------------------------------------ */
/* Replaces Gofrt<double,T> (gofrt.h:37-45, gofrt.cpp:22-155) driven by
 * CalculateMultiThread::calculate (calculatemultithread.h:106-162).
 * Counts are exact integers: the reference's vdata[k] is counts[k]*incr with
 * incr = 1/int(n/skip) (gofrt.cpp:91-92).  Output layout
 * [leff][ntypes*(ntypes+1)][nbin] (gofrt.cpp:65-71). */
typedef struct anl_gofrt anl_gofrt;
int anl_gofrt_create(anl_traj *tr, double rmin, double rmax, uint32_t nbin, uint32_t tmax,
                     uint32_t skip, uint32_t every, int form, anl_gofrt **out);
int anl_gofrt_destroy(anl_gofrt *g);
/* Gofrt::reset(n) (gofrt.cpp:41-63): leff = min(n, tmax) (tmax 0 = n). */
int anl_gofrt_reset(anl_gofrt *g, uint32_t n, uint32_t *leff_out);
/* Gofrt::calculate(primo): whole (lags x t0 x ordered pairs) nest in one
 * launch; counts_host receives [leff][nch][nbin] uint64. */
int anl_gofrt_calculate(anl_gofrt *g, uint64_t primo, uint64_t *counts_host);
/* Same, device output (counts_dev zeroed and filled), enqueued on stream
 * (a hipStream_t, NULL = default stream), no host synchronisation. */
int anl_gofrt_calculate_async(anl_gofrt *g, uint64_t primo, uint64_t *counts_dev, void *stream);
/* Gofrt::calc_single_th(t, imedia, a0, a1, primo, ith) (gofrt.cpp:96-122):
 * one lag t, one t0 offset imedia, atoms a0 <= i < a1 against all j;
 * ADDS into counts_host [leff][nch][nbin]. */
int anl_gofrt_calc_single_th(anl_gofrt *g, int t, int imedia, int a0, int a1, uint64_t primo,
                             uint64_t *counts_host);
/* Uneven blocks in one launch (replaces the per-block loop of
 * BlockAverageG::calcola_custom, blockaverage.h:128-144, for Gofrt): block b
 * covers t0 in [0, n_b) from frame primo_b with leff_b = min(n_b, tmax) lags;
 * counts_dev is [nblocks][leff_max][nch][nbin], leff_max = max leff_b. */
int anl_gofrt_calculate_blocks_async(anl_gofrt *g, const uint64_t *primo, const uint32_t *n,
                                     uint32_t nblocks, uint64_t *counts_dev, void *stream);
/* Form actually used by the last calculate (ANL_GOFRT_FP64/SCREENED), the
 * number of candidate pairs the screen passed to the FP32 evaluation, and how
 * many of those were re-evaluated in FP64 (near a bin edge, rmin or rmax).
 * Diagnostics of this implementation; no counterpart in the reference. */
int anl_gofrt_last_stats(anl_gofrt *g, int *form_used, uint64_t *candidates, uint64_t *rechecks);
/* Of the last screened calculate: the candidates whose FP32 distance fell in
 * the window around a bin edge, rmin or rmax (window, the rechecks of
 * anl_gofrt_last_stats), re-evaluated in FP64 from their 32-bit fractions;
 * and how many of those were still within that evaluation's error bound of
 * an edge (about 1e-7 A) and took the reference's FP64 sequence from the
 * positions (exact).  Diagnostics of this implementation; no counterpart in
 * the reference. */
int anl_gofrt_recheck_stats(anl_gofrt *g, uint64_t *window, uint64_t *exact);
/* Of the last screened calculate: how many (wavefront, t0) units used the
 * cluster (MFMA) screen, out of how many (0 of 0 when it was off). */
int anl_gofrt_cluster_stats(anl_gofrt *g, uint64_t *clustered, uint64_t *units);
/* Of the last calculate: ordered pairs whose distance some test looked at
 * (evaluated = tile + screen16 + fp64: FP16 MFMA tiles of the cluster
 * screen, the 16-bit screen, the plain FP64 form), and all ordered pairs of
 * the calculation (the reference evaluates every one, calculatemultithread.h:
 * 114-139 around gofrt.cpp:101-104).  all_pairs - evaluated were pruned: the
 * cluster screen skips a 32-column half whose atoms are all farther than
 * rmax from the wavefront's bounding sphere.  Diagnostics of this
 * implementation; no counterpart in the reference. */
int anl_gofrt_pair_stats(anl_gofrt *g, uint64_t *evaluated, uint64_t *all_pairs, uint64_t *tile, uint64_t *screen16,
                         uint64_t *fp64);
/* Screen parameters for a constant cell, on the host (no device needed):
 * box9 is one frame's internal-format box [lo(3), half lengths(3), xy, xz,
 * yz].  valid: the 16-bit screen applies (rmax below half of every cell
 * width); threshold: its bound on |q|^2.  rho_cap: the largest wavefront
 * radius the cluster (FP16 MFMA) screen accepts, 0 when it is off; dq: the
 * 16-bit quantisation bound of a pair vector; q2_max and threshold_max: the
 * largest |q|^2 of a pair the cluster screen must pass and its largest
 * distance threshold, which stay far below both the FP16 maximum (65504) and
 * the sentinel 60000 of inactive tile rows and columns.  No counterpart in
 * the reference. */
int anl_gofrt_screen_info(const double *box9, double rmax, int *valid, uint32_t *threshold, double *rho_cap,
                          double *dq, double *q2_max, double *threshold_max);
/* Launch shape of the pair kernels: unroll of the 16-bit screen loop (4 or
 * 8), XCD-aware workgroup order (0/1), pair evaluations per workgroup (sets
 * the t0 chunk of one workgroup), and the cluster screen (0: off; 1: on,
 * wavefronts whose 64 atoms are compact screen their pairs with FP16 MFMA
 * distance tiles, skipping groups of 32 atoms that a per-frame table of
 * bounding spheres shows far; 2: as 1, with each step's 64 column atoms
 * cached in LDS as FP32 offsets from the wavefront's centre for the
 * candidates' evaluation; 3: on, visiting every j in order without the
 * table; 4: as 3, with the column cache of 2).  Shapes change speed only,
 * never results.  anl_gofrt_autotune
 * times each candidate (cand[4*c] = unroll, cand[4*c+1] = xcd, cand[4*c+2] =
 * log2 pairs per workgroup, cand[4*c+3] = cluster 0..4; NULL = a built-in
 * set of 9) on the first lags of the calculation set by reset() at primo (at least
 * nlag_probe, raised until the probe launch has >= 4096 workgroups), keeps
 * the fastest and reports ms_out[c] (may be NULL) and its index.  The reference's equivalent knob is nthreads
 * (calculatemultithread.h:80-104); no counterpart for the rest. */
int anl_gofrt_set_tuning(anl_gofrt *g, int unroll, int xcd, uint64_t pairs_per_wg, int cluster);
int anl_gofrt_get_tuning(anl_gofrt *g, int *unroll, int *xcd, uint64_t *pairs_per_wg, int *cluster);
int anl_gofrt_autotune(anl_gofrt *g, uint64_t primo, uint32_t nlag_probe, const int32_t *cand, uint32_t ncand,
                       float *ms_out, uint32_t *best_index);

/* ---- block engine ------------------------------------------------------ */
/* Replaces the per-block loop of BlockAverageG::calcola_custom with MediaVar
 * (blockaverage.h:107-203, calcoliblocchi.h:21-65) for Gofrt and MSD, with
 * blocks of uneven lengths: block b starts at frame primo[b] and averages
 * n[b] t0 values (reset(n[b]); every block must give the same number of
 * lags).  g(r,t) of all blocks runs in one launch of the pair kernel with the
 * parameters and form of `gofrt` (NULL: no g(r,t)); MSD (msd != 0) of all
 * blocks in one launch; then mean and variance of the mean over blocks on
 * the device, with the reference's Welford order.  use_graph != 0 captures
 * the whole sequence once into a HIP graph, replayed by every anl_blocks_run.
 * No counterpart in the reference beyond the loop it replaces. */
typedef struct anl_blocks anl_blocks;
int anl_blocks_create(anl_traj *tr, const uint64_t *primo, const uint32_t *n, uint32_t nblocks, anl_gofrt *gofrt,
                      int msd, uint32_t msd_skip, uint32_t msd_tmax, int cm_msd, int cm_self, int use_graph,
                      anl_blocks **out);
/* Enqueue one evaluation of every block on stream (NULL = default). */
int anl_blocks_run(anl_blocks *b, void *stream);
/* Synchronise and copy results: mean/var [leff][nt(nt+1)][nbin] FP64 (the
 * reference's vdata layout), counts [nblocks][leff][nt(nt+1)][nbin] uint64
 * (may be NULL); MSD mean/var [leff][f_cm][nt], per_block [nblocks][...]. */
int anl_blocks_gofrt_results(anl_blocks *b, double *mean, double *var, uint64_t *counts, uint32_t *leff);
int anl_blocks_msd_results(anl_blocks *b, double *mean, double *var, double *per_block, uint32_t *leff);
int anl_blocks_destroy(anl_blocks *b);

/* ---- spherical-harmonic density correlations -------------------------- */
/* Replaces SphericalCorrelations<lmax,double,T> (sphericalcorrelations.h:9-40,
 * sphericalcorrelations.cpp:10-270) and SphericalBase::calc
 * (sphericalbase.cpp:19-66) without neighbour lists.  rminmax holds
 * ntypes*ntypes (rmin, rmax) pairs indexed ntypes*type_i + type_j
 * (sphericalbase.cpp:7-16); lmax <= ANL_SPH_LMAX_MAX. */
#define ANL_SPH_LMAX_MAX 10
typedef struct anl_sph anl_sph;
int anl_sph_create(anl_traj *tr, uint32_t lmax, const double *rminmax, uint32_t nbin, anl_sph **out);
int anl_sph_destroy(anl_sph *h);
/* reset(n) + calculate(primo) (sphericalcorrelations.cpp:40-236): out_host
 * [leff][ntypes][ntypes][nbin][lmax+1] FP64, leff = min(n, tmax) (tmax = 0:
 * n); index k of the last axis holds l = lmax - k, as in the reference.
 * ANL_ERR_TOO_SHORT when leff + n + primo exceeds the frames (:63-65). */
int anl_sph_calculate(anl_sph *h, uint64_t primo, uint32_t n, uint32_t tmax, uint32_t skip, double *out_host);
/* SphericalBase::calc(frame) (sphericalbase.cpp:19-66): out_host
 * [natoms][ntypes][nbin][(lmax+1)^2] in original atom order and the
 * reference's (l, m) order (l = lmax first; m = -1..-l, then 0..l). */
int anl_sph_snapshot(anl_sph *h, uint64_t frame, double *out_host);
/* Steinhardt<lmax> (steinhardt.h:17-120, steinhardt.cpp:7-236) without
 * neighbour lists, not averaged, over the handle's ranges and bins: frames
 * primo + i, i = 0, skip, .. < n - n % skip; q_l = sqrt(4 pi/(2l+1) sum_m
 * |S_lm|^2) / (neighbour count).  do_histogram: out_host [nbin][ntypes]
 * [ntypes][nbin_st]^nl (one axis per entry of l_list, the first slowest),
 * count * incr with incr = 1/int(n/skip); else out_host [nbin][n/skip]
 * [natoms][ntypes][lmax] holding q_1..q_lmax (0 without neighbours).
 * averaged (neighbour lists only, :128-173): q_l of qbar_lm = S_lm/n of the
 * atom (for its own type) plus S_lm/n of each of its SANN neighbours.
 * ANL_ERR_TOO_SHORT when n + primo exceeds the frames (:80-83). */
int anl_steinhardt_calculate(anl_sph *h, uint64_t primo, uint32_t n, uint32_t skip, const uint32_t *l_list,
                             uint32_t nl, uint32_t nbin_st, int do_histogram, int averaged, double *out_host);
/* Switches the handle to neighbour lists (Neighbours, neighbour.h:10-104,
 * neighbour.cpp:8-101, the SANN variant of sphericalbase.cpp:67-93): for
 * every atom and type t, the atoms of type t within sqrt(cutoff2[t]), at most
 * nneigh[t] of them (else "Too many neighbours in shell!"), sorted by
 * distance; the harmonics of the first SANN-count ones go to bin 0 and the
 * radial ranges are ignored.  Sum of nneigh at most 1024, at most 16 types. */
int anl_sph_set_neighbours(anl_sph *h, const uint32_t *nneigh, const double *cutoff2);
/* Bytes of per-frame densities held on the device at once (default: half the
 * free device memory at creation, at most 64 GiB);
 * a smaller budget processes the t0 values in more chunks, same results up
 * to the order of the chunk sums.  No counterpart in the reference (its
 * CalcBuffer size, buffer_size, plays the same role on the host). */
int anl_sph_set_budget(anl_sph *h, uint64_t bytes);

/* ---- mean square displacement ----------------------------------------- */
/* Replaces MSD<T> (msd.h:30-47, msd.cpp:40-146): out_host [leff][f_cm][ntypes]
 * FP64, f_cm = 2 when cm_msd, per-type mean over (t0 in 0,skip,..<n, atoms);
 * cm_self subtracts the per-type centre-of-mass displacement. */
int anl_msd_calculate(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self,
                      uint32_t n, uint64_t primo, double *out_host);
int anl_msd_calculate_async(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self,
                            uint32_t n, uint64_t primo, double *out_dev, void *stream);
/* Allocation-free variant of anl_msd_calculate_async, with the same
 * arguments and results (bit for bit: the same kernels).  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its scratch (block table, term counts,
 * partial sums: about 8 * leff * ceil(natoms/64) * ntypes bytes at t0
 * stride 1, ceil(natoms/256) at other strides) lives in a fixed 16 MiB
 * device array per GPU, held until the process ends; larger calculations
 * fail with ANL_ERR_ARG before anything is launched.  The call synchronises
 * stream before returning.  Calls that overlap on one GPU share the array
 * and race. */
int anl_msd_calculate_noalloc(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n,
                              uint64_t primo, double *out_dev, void *stream);
/* Launch shape of the MSD kernel for this trajectory: the lag tile R of the
 * t0-stride-1 kernel, which holds R frames of an atom in registers and loads
 * two frames per t0 for R updates, and its wavefronts per workgroup W; a
 * workgroup covers R * W lags of 64 atoms, streaming their frames once for
 * all of them.  (R, W) is one of (8,4) (12,4) (16,4) (24,4) (16,8) (24,8)
 * (12,12) (16,12) (8,16); default (16,4).  Lag tile 0 selects the
 * eight-lags-per-thread kernel that every other stride uses (W ignored).
 * All shapes with R > 0 give identical results (each (atom, lag) sum runs
 * over t0 in order with the same operations, and the workgroup sums its 64
 * atoms in slot order); tile 0 agrees with them to rounding (both within
 * 1e-10 of the reference).  anl_msd_autotune times each candidate shape
 * (cand[2c], cand[2c+1]; NULL = the nine shapes above and tile 0) on the
 * calculation reset(n) at primo and keeps the fastest; ms_out[c] (may be
 * NULL) receives the times.  The reference's equivalent knob is nthreads
 * (msd.h:30-47, calculatemultithread.h:80-104); no counterpart for the
 * shape. */
int anl_msd_set_tuning(anl_traj *tr, uint32_t lag_tile, uint32_t waves);
int anl_msd_get_tuning(anl_traj *tr, uint32_t *lag_tile, uint32_t *waves);
int anl_msd_autotune(anl_traj *tr, uint32_t skip, uint32_t tmax, int cm_msd, int cm_self, uint32_t n, uint64_t primo,
                     const uint32_t *cand, uint32_t ncand, float *ms_out, uint32_t *best_index);

/* ---- time series and Green-Kubo --------------------------------------- */
/* Device mirror of ReadLog_numpy (readlog_numpy.h:6-26): row-major
 * [nrows][ncols] FP64. */
typedef struct anl_series anl_series;
int anl_series_create(const double *data, uint64_t nrows, uint32_t ncols, anl_series **out);
int anl_series_destroy(anl_series *s);
/* Replaces GreenKuboNComponentIonicFluid::calculate (greenkubo...cpp:156-380)
 * with reset(n) semantics (:122-131): idx_j[a] is the first column of current
 * a's (x,y,z) triplet; out_host [leff][3*ncur^2+2] FP64 in the reference
 * column order (C, (I,E) pairs, k_gk, k_einst). */
int anl_gk_calculate(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip,
                     uint32_t tmax, uint32_t n, uint32_t n_seg, int subtract_mean,
                     uint32_t start_mean, uint64_t primo, double *out_host);
/* Same interface and output as anl_gk_calculate, with the lag products
 * J_a(t0).J_b(t0+l) formed on the matrix cores as three split-BF16 MFMA
 * products (hi*hi + hi*lo + lo*hi), FP32 accumulation within chunks of 8192
 * t0 and FP64 across chunks; the integrals and the Schur complement are the
 * FP64 ones.  Not bit-exact: C(l) is within ~1e-6 of max|C| of the FP64
 * form.  Replaces the same loop (greenkubo...cpp:156-249). */
int anl_gk_calculate_mfma(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip,
                          uint32_t tmax, uint32_t n, uint32_t n_seg, int subtract_mean,
                          uint32_t start_mean, uint64_t primo, double *out_host);
/* Green-Kubo of nblocks blocks in one call (the block loop of
 * BlockAverageG::calcola_custom, blockaverage.h:107-190, with
 * MediaVarCovar, calcoliblocchi.h:70-176, and no covariance pairs): block b
 * is reset(n[b]) + calculate(primo[b]); all blocks must have the same
 * leff = min(n[b], tmax).  out_dev [nblocks][leff][3*ncur^2+2]; mean_dev and
 * var_dev ([leff][3*ncur^2+2], both or neither, nblocks >= 2) receive the
 * mean over blocks and the variance of that mean.  form: ANL_GK_FP64 (the
 * reference's values) or ANL_GK_MFMA (split-BF16 lag products).  The FP64
 * lag loop of all blocks is one launch, and so is the epilogue of all
 * blocks. */
#define ANL_GK_FP64 0
#define ANL_GK_MFMA 1
int anl_gk_calculate_blocks_async(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip,
                                  uint32_t tmax, uint32_t n_seg, int subtract_mean, uint32_t start_mean,
                                  const uint64_t *primo, const uint32_t *n, uint32_t nblocks, int form,
                                  double *out_dev, double *mean_dev, double *var_dev, void *stream);
/* Stream-ordered variants writing [leff][3*ncur^2+2] FP64 to device memory.
 * Scratch lives in the series (grow-only; growing synchronises the device),
 * so the calls on one series must be ordered on one stream. */
int anl_gk_calculate_async(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip,
                           uint32_t tmax, uint32_t n, uint32_t n_seg, int subtract_mean,
                           uint32_t start_mean, uint64_t primo, double *out_dev, void *stream);
int anl_gk_calculate_mfma_async(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip,
                                uint32_t tmax, uint32_t n, uint32_t n_seg, int subtract_mean,
                                uint32_t start_mean, uint64_t primo, double *out_dev, void *stream);

/* Launch shape of the split-BF16 MFMA lag products of this series: tiles of
 * 32 x 32 (t0 x lag) per wavefront (4 or 8; default 8), MFMA steps of 512 t0
 * per FP32 chunk before the sums go to FP64 (8 or 16; default 16), and
 * wavefronts per workgroup sharing one X operand block (2 or 4; default 4).
 * Tiles per wavefront and wavefronts per workgroup never change results; the
 * chunk length changes the FP32 rounding within the MFMA form's error bound.
 * anl_gk_autotune times each candidate (cand[3c..3c+2]; NULL = a built-in
 * set of 6) on the lag products of the given blocks (reset(n[b]) at primo[b],
 * as anl_gk_calculate_blocks_async: one untimed call, then the fastest of
 * three timed ones) and keeps the fastest; ms_out[c] (may be NULL) receives
 * the times.  The reference's only knob is nthreads
 * (greenkuboNcomponentionicfluid.h:26-51); no counterpart for the shape. */
int anl_gk_set_tuning(anl_series *s, uint32_t tiles_per_wave, uint32_t chunk_steps, uint32_t waves_per_group);
int anl_gk_get_tuning(anl_series *s, uint32_t *tiles_per_wave, uint32_t *chunk_steps, uint32_t *waves_per_group);
int anl_gk_autotune(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t tmax, uint32_t n_seg,
                    const uint64_t *primo, const uint32_t *n, uint32_t nblocks, const uint32_t *cand, uint32_t ncand,
                    float *ms_out, uint32_t *best_index);

/* ---- block statistics --------------------------------------------------- */
/* MediaVar::calculate (calcoliblocchi.h:35-58) on device vectors of length
 * len: one Welford step with x = block iblock (0-based); mean/var updated in
 * place.  anl_mediavar_end divides var by (n_b-1)*n_b (:59-61). */
int anl_mediavar_update_async(double *mean_dev, double *var_dev, const double *x_dev, uint64_t len,
                              uint32_t iblock, void *stream);
int anl_mediavar_end_async(double *var_dev, uint64_t len, uint32_t n_b, void *stream);
/* Converts counts_dev (uint64) to the reference FP64 layout: out = count*incr. */
int anl_counts_to_double_async(const uint64_t *counts_dev, double incr, uint64_t len, double *out_dev,
                               void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ANALISI_HIP_H */
