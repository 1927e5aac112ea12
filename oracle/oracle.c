/*
 * oracle.c -- CPU restatement of analisi's time-lagged pair-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (analisi_amd/, include/)
 * links, loads or calls this file.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker and by bench.py's cpu_baseline leg
 * (timed on the host cores, never as the measured GPU value).
 *
 * Every function restates the reference algorithm from its source, cited as
 * file:line of rikigigi/analisi v0.5.2.  The arithmetic is kept in the same
 * order as the reference (x86-64, SSE2 doubles, no FMA contraction: build
 * with -ffp-contract=off) so that integer histogram counts are bit-exact and
 * FP64 outputs match to the last bit where the reference order is known.
 *
 * Parity pinning (what checks each function; see tests/):
 *   orc_gk_calculate       pinned: reference goldens gk/gk_2comp, test_gk.csv,
 *                          CLI GK_heat_single/double (with orc_mvc_*).
 *                          Schur complement (schur00): pinned against Eigen
 *                          3.3.9's evaluation of the reference's expression
 *                          (ref_schur.cpp, the reference's vendored headers):
 *                          bit-identical up to 5 currents, 1e-13 beyond.
 *   orc_mvc_* / orc_mv_*   pinned: CLI GK_heat_* text outputs; closed form.
 *   orc_min_image          pinned: brute-force lattice search, |x_k|<=l_k/2
 *                          invariant, hand-derived iteration cases.
 *   orc_gofrt_*            pinned: brute force on tiny inputs, exact-rational
 *                          bin edges, lag-0 self channel, pair symmetry,
 *                          structural facts of test_gofr.csv.
 *   orc_msd_calculate      pinned: closed forms (linear motion, uniform
 *                          translation with cm_self, cm channel), thread
 *                          invariance.
 *   orc_type_ids, orc_cm   pinned: closed forms.
 *   orc_sph_ylm            pinned: scipy.special.sph_harm_y (complex Y_l^m,
 *                          Condon-Shortley phase) turned into the real form.
 *   orc_sph_density        pinned: an independent numpy brute force (lattice
 *                          minimum image, scipy harmonics) on small frames.
 *   orc_sph_correlations   pinned: the addition theorem, sum_m Y_lm(a)Y_lm(b)
 *                          = (2l+1)/(4 pi) P_l(a.b), evaluated by brute force
 *                          over neighbour pairs with scipy's Legendre P_l.
 *                          The reference's t1 visiting order depends on its
 *                          thread split; this restatement visits t1 in
 *                          ascending order (agreement to rounding).
 *   orc_steinhardt         pinned: sc / bcc / fcc crystals (every atom's q_l
 *                          equals scipy's from its neighbour vectors; q4, q6
 *                          equal the tabulated values), q_l = 1 for a lone
 *                          neighbour, the histogram equal to the binned q_l.
 *   orc_*_sann             pinned: a numpy brute force of the published SANN
 *                          rule (van Meel et al., J. Chem. Phys. 136, 234107,
 *                          2012) with scipy harmonics; fcc keeps its 12
 *                          nearest neighbours; addition theorem over the
 *                          selected neighbours.  List capacity: each list is
 *                          checked against its own count (the reference checks
 *                          both atoms' lists for a pair within either cutoff).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Trajectory view: what BaseTrajectory serves to the calculations.          */
/* ------------------------------------------------------------------------ */
typedef struct {
    const double *pos;     /* [nframes][natoms][3], as positions<false>() serves them */
    const double *box;     /* [nframes][box_stride], internal format (basetrajectory.h:21-43) */
    const int *type_id;    /* [natoms], dense ids 0..ntypes-1 */
    const double *cm;      /* [nframes][ntypes][3], positions_cm<false>() */
    uint64_t nframes;
    uint64_t natoms;
    uint64_t ntypes;
    uint64_t box_stride;   /* 6 or 9 */
    int32_t triclinic;
    int32_t pad_;
    uint64_t loaded;       /* get_nloaded_timesteps() */
} orc_traj;

/* basetrajectory.cpp:51-89 -- type ids are the rank of each distinct type
 * value in the sorted list of distinct values. Returns ntypes. */
uint64_t orc_type_ids(const int *types, uint64_t natoms, int *type_id)
{
    if (natoms == 0) return 0;
    int *distinct = (int *)malloc(natoms * sizeof(int));
    uint64_t nd = 0;
    for (uint64_t i = 0; i < natoms; i++) {
        int seen = 0;
        for (uint64_t k = 0; k < nd; k++)
            if (distinct[k] == types[i]) { seen = 1; break; }
        if (!seen) distinct[nd++] = types[i];
    }
    /* insertion sort of the distinct values */
    for (uint64_t a = 1; a < nd; a++) {
        int v = distinct[a];
        uint64_t b = a;
        while (b > 0 && distinct[b - 1] > v) { distinct[b] = distinct[b - 1]; b--; }
        distinct[b] = v;
    }
    for (uint64_t i = 0; i < natoms; i++)
        for (uint64_t k = 0; k < nd; k++)
            if (distinct[k] == types[i]) { type_id[i] = (int)k; break; }
    free(distinct);
    return nd;
}

/* basetrajectory.h:94-105 -- lammps box (xlo,xhi,ylo,yhi,zlo,zhi[,xy,xz,yz])
 * to internal (xlo,ylo,zlo,lx/2,ly/2,lz/2[,xy,xz,yz]). */
void orc_lammps_to_internal(double *c)
{
    double t = c[4];
    c[5] = (c[5] - c[4]) / 2;
    c[4] = (c[3] - c[2]) / 2;
    c[3] = (c[1] - c[0]) / 2;
    c[1] = c[2];
    c[2] = t;
}

/* trajectory_numpy.cpp:88-101 -- copy and convert every frame's box. */
void orc_boxes_from_lammps(const double *in, uint64_t nframes, int triclinic, double *out)
{
    uint64_t stride = triclinic ? 9 : 6;
    for (uint64_t f = 0; f < nframes; f++) {
        memcpy(out + f * stride, in + f * stride, stride * sizeof(double));
        orc_lammps_to_internal(out + f * stride);
    }
}

/* trajectory_numpy.cpp:201-225 -- per-frame, per-type running mean of the raw
 * input coordinates (unwrapped, unrotated). */
void orc_cm(const double *a, const int *type_id, uint64_t nframes, uint64_t natoms,
            uint64_t ntypes, double *cm)
{
    int *cont = (int *)malloc(ntypes * sizeof(int));
    for (uint64_t i = 0; i < nframes; ++i) {
        for (uint64_t it = 0; it < ntypes; ++it) {
            cont[it] = 0;
            for (int k = 0; k < 3; k++) cm[(i * ntypes + it) * 3 + k] = 0.0;
        }
        for (uint64_t ia = 0; ia < natoms; ++ia) {
            int it = type_id[ia];
            cont[it]++;
            for (int k = 0; k < 3; k++)
                cm[(i * ntypes + it) * 3 + k] +=
                    (a[(i * natoms + ia) * 3 + k] - cm[(i * ntypes + it) * 3 + k]) / (double)cont[it];
        }
    }
    free(cont);
}

/* basetrajectory.h:224-268 -- iterative minimum image, z then y then x; in
 * a triclinic cell a z step also shifts y by yz and x by xz, a y step shifts
 * x by xy. */
void orc_min_image(double *delta, const double *l_half, const double *xy_xz_yz, int triclinic)
{
    double xy = 0, xz = 0, yz = 0;
    if (triclinic) { xy = xy_xz_yz[0]; xz = xy_xz_yz[1]; yz = xy_xz_yz[2]; }
    while (fabs(delta[2]) > l_half[2]) {
        if (delta[2] < 0.0) {
            delta[2] += l_half[2] * 2;
            if (triclinic) { delta[1] += yz; delta[0] += xz; }
        } else {
            delta[2] -= l_half[2] * 2;
            if (triclinic) { delta[1] -= yz; delta[0] -= xz; }
        }
    }
    while (fabs(delta[1]) > l_half[1]) {
        if (delta[1] < 0.0) {
            delta[1] += l_half[1] * 2;
            if (triclinic) delta[0] += xy;
        } else {
            delta[1] -= l_half[1] * 2;
            if (triclinic) delta[0] -= xy;
        }
    }
    while (fabs(delta[0]) > l_half[0]) {
        if (delta[0] < 0.0) delta[0] += l_half[0] * 2;
        else delta[0] -= l_half[0] * 2;
    }
}

/* basetrajectory.h:200-219 -- x = xi - xj, min image, d2 = 0 + x0^2 + x1^2 + x2^2 */
double orc_d2_min_image(const double *xi, const double *xj, const double *box, int triclinic, double *x)
{
    double d2 = 0.0;
    for (int k = 0; k < 3; k++) x[k] = xi[k] - xj[k];
    orc_min_image(x, box + 3, box + 6, triclinic);
    for (int k = 0; k < 3; k++) d2 += x[k] * x[k];
    return d2;
}

/* basetrajectory.h:145-161 -- wrap every atom of every frame around the cell
 * centre: x -= l/2, min image, x += l/2. */
void orc_pbc_wrap(double *pos, const double *boxes, uint64_t stride, uint64_t nframes,
                  uint64_t natoms, int triclinic)
{
    for (uint64_t f = 0; f < nframes; f++) {
        const double *c = boxes + stride * f;
        double x0[3] = {c[3], c[4], c[5]};
        for (uint64_t ia = 0; ia < natoms; ++ia) {
            double *xa = pos + f * natoms * 3 + ia * 3;
            for (int k = 0; k < 3; ++k) xa[k] = xa[k] - x0[k];
            orc_min_image(xa, c + 3, c + 6, triclinic);
            for (int k = 0; k < 3; ++k) xa[k] = xa[k] + x0[k];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Gofrt                                                                      */
/* ------------------------------------------------------------------------ */

/* gofrt.h:86-104 -- channel of an unordered type pair, reverse upper-
 * triangular order. */
static unsigned get_itype(unsigned ntypes, unsigned t1, unsigned t2)
{
    if (t2 < t1) { unsigned tmp = t2; t2 = t1; t1 = tmp; }
    return ntypes * (ntypes + 1) / 2 - (t2 + 1) * (t2 + 2) / 2 + t1;
}

unsigned orc_gofrt_channel(unsigned ntypes, unsigned t1, unsigned t2, int same_atom)
{
    unsigned c = get_itype(ntypes, t1, t2);
    if (same_atom) c += ntypes * (ntypes + 1) / 2;
    return c;
}

/* gofrt.cpp:114-115 -- the bin of a squared distance that passed the gate:
 * sqrt in double, (d-rmin)/dr in double, rounded to float, floorf, to int. */
int orc_gofrt_bin(double d2, double rmin, double dr)
{
    double d = sqrt(d2);
    return (int)floorf((float)((d - rmin) / dr));
}

typedef struct {
    const orc_traj *tr;
    double rmin, rmax, rmax2, rmin2, dr, incr;
    unsigned nbin, leff;
    uint64_t data_length;
} gofrt_cfg;

/* gofrt.cpp:96-122 -- one (lag, t0) over iatom in [a0,a1), all jatom. */
static void gofrt_single(const gofrt_cfg *g, int t, int imedia, uint64_t a0, uint64_t a1,
                         uint64_t primo, uint64_t *cnt, double *th)
{
    const orc_traj *tr = g->tr;
    unsigned nt = (unsigned)tr->ntypes;
    uint64_t N = tr->natoms;
    uint64_t fi = primo + (uint64_t)imedia, fj = primo + (uint64_t)imedia + (uint64_t)t;
    const double *box = tr->box + fi * tr->box_stride;
    uint64_t nch = (uint64_t)nt * (nt + 1);
    for (uint64_t ia = a0; ia < a1; ia++) {
        const double *xi = tr->pos + (fi * N + ia) * 3;
        for (uint64_t ja = 0; ja < N; ja++) {
            double x[3];
            double d = orc_d2_min_image(xi, tr->pos + (fj * N + ja) * 3, box, tr->triclinic, x);
            if (d > g->rmax2 || d < g->rmin2) continue;
            unsigned it = orc_gofrt_channel(nt, (unsigned)tr->type_id[ia], (unsigned)tr->type_id[ja], ia == ja);
            d = sqrt(d);
            int idx = (int)floorf((float)((d - g->rmin) / g->dr));
            if ((unsigned)idx < g->nbin && idx >= 0) {
                /* gofrt.h:71-76 gofr_idx */
                uint64_t k = (uint64_t)t * nch * g->nbin + (uint64_t)g->nbin * it + (uint64_t)idx;
                if (cnt) cnt[k] += 1;
                if (th) th[k] += g->incr;
            }
        }
    }
}

typedef struct {
    const gofrt_cfg *g;
    int t, imedia;
    uint64_t a0, a1, primo;
    uint64_t *cnt;
    double *th;
} gofrt_job;

static void *gofrt_worker(void *p)
{
    gofrt_job *j = (gofrt_job *)p;
    gofrt_single(j->g, j->t, j->imedia, j->a0, j->a1, j->primo, j->cnt, j->th);
    return NULL;
}

/* gofrt.cpp:41-63 reset, :74-93 calc_init, calculatemultithread.h:106-162
 * calculate with PARALLEL_SPLIT_ATOM|SERIAL_LOOP_TIME|SERIAL_LOOP_AVERAGE,
 * :50-78 splitter, gofrt.cpp:127-135 merge of the per-thread copies.
 *
 * counts: [leff][nch][nbin] exact integer hits (may be NULL).
 * vdata : [leff][nch][nbin] the reference's FP64 values, accumulated with
 *         the same per-thread copies and merge order (may be NULL).
 * Returns 0, or -1 when the trajectory is too short (gofrt.cpp:81-83). */
int orc_gofrt_calculate(const orc_traj *tr, double rmin, double rmax, unsigned nbin,
                        unsigned lmax, unsigned nthreads, unsigned skip, unsigned every,
                        unsigned n, uint64_t primo, uint64_t *counts, double *vdata)
{
    if (nthreads == 0) nthreads = 1;          /* calculatemultithread.h:43-45 */
    if (skip == 0) skip = 1;
    if (every == 0) every = 1;
    gofrt_cfg g;
    g.tr = tr;
    g.rmin = rmin; g.rmax = rmax;
    g.dr = (rmax - rmin) / (double)nbin;      /* gofrt.cpp:27-29 */
    g.rmax2 = rmax * rmax;
    g.rmin2 = rmin * rmin;
    g.nbin = nbin;
    g.leff = (n < lmax || lmax == 0) ? n : lmax;
    unsigned nt = (unsigned)tr->ntypes;
    g.data_length = (uint64_t)g.leff * nt * (nt + 1) * nbin;
    if ((uint64_t)g.leff + n + primo > tr->nframes + 1) return -1;
    g.incr = (n / skip > 0) ? 1.0 / (double)(int)(n / skip) : 1.0;

    if (counts) memset(counts, 0, g.data_length * sizeof(uint64_t));
    if (vdata) memset(vdata, 0, g.data_length * sizeof(double));
    double *th = NULL;
    uint64_t *thc = NULL;
    if (nthreads > 1) {
        if (vdata) th = (double *)calloc(g.data_length * (nthreads - 1), sizeof(double));
        if (counts) thc = (uint64_t *)calloc(g.data_length * (nthreads - 1), sizeof(uint64_t));
    }
    uint64_t N = tr->natoms;
    uint64_t npass = N / nthreads;
    pthread_t *tid = (pthread_t *)malloc(nthreads * sizeof(pthread_t));
    gofrt_job *jobs = (gofrt_job *)malloc(nthreads * sizeof(gofrt_job));
    for (unsigned t = 0; t < g.leff; t += every) {
        for (unsigned i = 0; i < n; i += skip) {
            for (unsigned ith = 0; ith < nthreads; ith++) {
                gofrt_job *j = &jobs[ith];
                j->g = &g; j->t = (int)t; j->imedia = (int)i; j->primo = primo;
                j->a0 = ith * npass;
                j->a1 = (ith == nthreads - 1) ? N : (ith + 1) * npass;
                j->cnt = (ith == 0) ? counts : (thc ? thc + (uint64_t)(ith - 1) * g.data_length : NULL);
                j->th = (ith == 0) ? vdata : (th ? th + (uint64_t)(ith - 1) * g.data_length : NULL);
                if (nthreads > 1) pthread_create(&tid[ith], NULL, gofrt_worker, j);
                else gofrt_worker(j);
            }
            if (nthreads > 1)
                for (unsigned ith = 0; ith < nthreads; ith++) pthread_join(tid[ith], NULL);
        }
    }
    for (unsigned ith = 0; ith + 1 < nthreads; ith++) {
        for (uint64_t k = 0; k < g.data_length; ++k) {
            if (vdata) vdata[k] += th[ith * g.data_length + k];
            if (counts) counts[k] += thc[ith * g.data_length + k];
        }
    }
    free(th); free(thc); free(tid); free(jobs);
    return 0;
}

/* gofrt.cpp:96-122 as a single call: (lag t, t0 index imedia, atoms [a0,a1)),
 * adding exact hit counts into counts [leff][nch][nbin]. */
void orc_gofrt_single(const orc_traj *tr, double rmin, double rmax, unsigned nbin, unsigned leff,
                      int t, int imedia, uint64_t a0, uint64_t a1, uint64_t primo, uint64_t *counts)
{
    gofrt_cfg g;
    g.tr = tr; g.rmin = rmin; g.rmax = rmax;
    g.dr = (rmax - rmin) / (double)nbin;
    g.rmax2 = rmax * rmax; g.rmin2 = rmin * rmin;
    g.nbin = nbin; g.leff = leff; g.incr = 1.0;
    unsigned nt = (unsigned)tr->ntypes;
    g.data_length = (uint64_t)leff * nt * (nt + 1) * nbin;
    gofrt_single(&g, t, imedia, a0, a1, primo, counts, NULL);
}

/* gofrt.cpp:37-39 (msd.cpp:36-38 is identical) */
unsigned orc_n_extra_timesteps(uint64_t nframes, unsigned n_b, unsigned lmax)
{
    unsigned v = (unsigned)(nframes / (n_b + 1) + 1);
    return (v < lmax || lmax == 0) ? v : lmax;
}

/* ------------------------------------------------------------------------ */
/* MSD                                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_traj *tr;
    unsigned skip, ntimesteps;
    uint64_t f_cm, ntypes;
    int cm_msd, cm_self;
    uint64_t primo;
    double *vdata;
    uint64_t begin, end;
} msd_job;

/* msd.cpp:64-125 */
static void *msd_worker(void *p)
{
    msd_job *j = (msd_job *)p;
    const orc_traj *tr = j->tr;
    uint64_t nt = j->ntypes, fcm = j->f_cm, N = tr->natoms;
    uint64_t *cont = (uint64_t *)malloc(nt * fcm * sizeof(uint64_t));
    double *v = j->vdata;
    for (uint64_t t = j->begin; t < j->end; t++) {
        for (uint64_t i = 0; i < nt * fcm; i++) { v[nt * t * fcm + i] = 0.0; cont[i] = 0; }
        for (uint64_t im = 0; im < j->ntimesteps; im += j->skip) {
            uint64_t fa = j->primo + im, fb = j->primo + im + t;
            for (uint64_t ia = 0; ia < N; ia++) {
                uint64_t it = (uint64_t)tr->type_id[ia];
                const double *xa = tr->pos + (fa * N + ia) * 3;
                const double *xb = tr->pos + (fb * N + ia) * 3;
                double s;
                if (j->cm_self) {
                    const double *ca = tr->cm + (fa * nt + it) * 3;
                    const double *cb = tr->cm + (fb * nt + it) * 3;
                    double e0 = xa[0] - xb[0] - (ca[0] - cb[0]);
                    double e1 = xa[1] - xb[1] - (ca[1] - cb[1]);
                    double e2 = xa[2] - xb[2] - (ca[2] - cb[2]);
                    s = e0 * e0 + e1 * e1 + e2 * e2;
                } else {
                    double e0 = xa[0] - xb[0], e1 = xa[1] - xb[1], e2 = xa[2] - xb[2];
                    s = e0 * e0 + e1 * e1 + e2 * e2;
                }
                double delta = s - v[nt * t * fcm + it];
                v[nt * t * fcm + it] += delta / (double)(++cont[it]);
            }
            if (j->cm_msd) {
                for (uint64_t it = 0; it < nt; it++) {
                    const double *ca = tr->cm + (fa * nt + it) * 3;
                    const double *cb = tr->cm + (fb * nt + it) * 3;
                    double e0 = ca[0] - cb[0], e1 = ca[1] - cb[1], e2 = ca[2] - cb[2];
                    double delta = (e0 * e0 + e1 * e1 + e2 * e2) - v[nt * t * fcm + nt + it];
                    v[nt * t * fcm + nt + it] += delta / (double)(++cont[nt + it]);
                }
            }
        }
    }
    free(cont);
    return NULL;
}

/* msd.cpp:40-49 reset, :52-61 calc_init, calculatemultithread.h:80-104 with
 * PARALLEL_SPLIT_TIME (one round; contiguous lag ranges per thread).
 * vdata: [leff][f_cm][ntypes].  Returns 0, -1 (too short), -2 (not loaded). */
int orc_msd_calculate(const orc_traj *tr, unsigned skip, unsigned lmax, unsigned nthreads,
                      int cm_msd, int cm_self, unsigned n, uint64_t primo, double *vdata)
{
    if (nthreads == 0) nthreads = 1;
    if (skip == 0) skip = 1;
    unsigned leff = (n < lmax || lmax == 0) ? n : lmax;
    if ((uint64_t)leff + n + primo > tr->nframes) return -1;
    if ((uint64_t)leff + n > tr->loaded) return -2;
    uint64_t npass = leff / nthreads;
    pthread_t *tid = (pthread_t *)malloc(nthreads * sizeof(pthread_t));
    msd_job *jobs = (msd_job *)malloc(nthreads * sizeof(msd_job));
    for (unsigned ith = 0; ith < nthreads; ith++) {
        msd_job *j = &jobs[ith];
        j->tr = tr; j->skip = skip; j->ntimesteps = n;
        j->f_cm = cm_msd ? 2 : 1; j->ntypes = tr->ntypes;
        j->cm_msd = cm_msd; j->cm_self = cm_self; j->primo = primo; j->vdata = vdata;
        j->begin = ith * npass;
        j->end = (ith == nthreads - 1) ? leff : (ith + 1) * npass;
        if (nthreads > 1) pthread_create(&tid[ith], NULL, msd_worker, j);
        else msd_worker(j);
    }
    if (nthreads > 1)
        for (unsigned ith = 0; ith < nthreads; ith++) pthread_join(tid[ith], NULL);
    free(tid); free(jobs);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Green-Kubo N components                                                    */
/* ------------------------------------------------------------------------ */

/* In-place inverse of an m x m column-major matrix by LU with partial
 * pivoting (the algorithm Eigen's MatrixBase::inverse() uses for dynamic
 * sizes).  For m == 1 this is exactly 1/a. */
static void inverse_pplu(double *a, int m, double *inv)
{
    int *perm = (int *)malloc(m * sizeof(int));
    for (int i = 0; i < m; i++) perm[i] = i;
    for (int k = 0; k < m; k++) {
        int p = k;
        double best = fabs(a[k * m + k]);
        for (int r = k + 1; r < m; r++)
            if (fabs(a[k * m + r]) > best) { best = fabs(a[k * m + r]); p = r; }
        if (p != k) {
            for (int c = 0; c < m; c++) { double t = a[c * m + k]; a[c * m + k] = a[c * m + p]; a[c * m + p] = t; }
            int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
        }
        for (int r = k + 1; r < m; r++) a[k * m + r] /= a[k * m + k];
        for (int c = k + 1; c < m; c++)
            for (int r = k + 1; r < m; r++) a[c * m + r] -= a[k * m + r] * a[c * m + k];
    }
    for (int col = 0; col < m; col++) {
        double *x = inv + col * m;
        for (int r = 0; r < m; r++) x[r] = (perm[r] == col) ? 1.0 : 0.0;
        for (int r = 0; r < m; r++)
            for (int k = 0; k < r; k++) x[r] -= a[k * m + r] * x[k];
        /* back substitution column by column from the last (Eigen's small-
         * panel upper solve): each row's terms are subtracted in decreasing
         * column order, then the row is scaled by the reciprocal pivot */
        for (int r = m - 1; r >= 0; r--) {
            for (int k = m - 1; k > r; k--) x[r] -= a[k * m + r] * x[k];
            x[r] = x[r] * (1.0 / a[r * m + r]);
        }
    }
    free(perm);
}

/* greenkuboNcomponentionicfluid.cpp:290-346 -- Schur complement of the
 * (0,0) entry: k = c00 - c(0,1:) inv(c(1:,1:)) c(1:,0); c column-major n x n. */
static double schur00(const double *c, int n)
{
    if (n <= 1) return c[0];
    int m = n - 1;
    double *sub = (double *)malloc(m * m * sizeof(double));
    double *inv = (double *)malloc(m * m * sizeof(double));
    double *tmp = (double *)malloc(m * sizeof(double));
    for (int cc = 0; cc < m; cc++)
        for (int r = 0; r < m; r++) sub[cc * m + r] = c[(cc + 1) * n + (r + 1)];
    inverse_pplu(sub, m, inv);
    for (int cc = 0; cc < m; cc++) {
        double s = c[(0 + 1) * n + 0] * inv[cc * m + 0];
        for (int k = 1; k < m; k++) s += c[(k + 1) * n + 0] * inv[cc * m + k];
        tmp[cc] = s;
    }
    double s = tmp[0] * c[0 * n + 1];
    for (int cc = 1; cc < m; cc++) s += tmp[cc] * c[0 * n + (cc + 1)];
    double k = c[0] - s;
    free(sub); free(inv); free(tmp);
    return k;
}

/* schur00 exported for its own test (against Eigen's evaluation of the same
 * expression, built by oracle/Makefile's ref target into oracle/_ref). */
double orc_schur00(const double *c, int n)
{
    return schur00(c, n);
}

/* greenkuboNcomponentionicfluid.cpp:122-131 reset, :156-356 calculate.
 * data: row-major [nrows][ncols]; idx_j[a] is the first column of current
 * a's (x,y,z) triplet.  vdata: [leff][narr], narr = 3*ncur^2 + 2.
 * Returns 0 or -1 when the series is too short (:168-169). */
int orc_gk_calculate(const double *data, uint64_t nrows, uint64_t ncols, const unsigned *idx_j,
                     unsigned ncur, unsigned skip, unsigned lmax, int subtract_mean,
                     unsigned start_mean, unsigned n_seg, unsigned n, unsigned primo, double *vdata)
{
    if (n_seg < 1) n_seg = 1;
    if (skip == 0) skip = 1;
    unsigned leff = (n < lmax || lmax == 0) ? n : lmax;
    unsigned N_corr = ncur * ncur;
    unsigned narr = 3 * N_corr + 2;
    for (uint64_t i = 0; i < (uint64_t)leff * narr; i++) vdata[i] = 0.0;
    if ((uint64_t)leff + primo + n > nrows) return -1;

    double *JJ = (double *)malloc(N_corr * sizeof(double));
    double *matr = (double *)malloc(N_corr * sizeof(double));
    double *JJm_T = (double *)calloc(N_corr, sizeof(double));
    unsigned cont_JJ = 0;
    unsigned seg = n / n_seg;
    for (unsigned iseg = 0; iseg < n_seg; iseg++) {
        ++cont_JJ;
        for (unsigned it = 0; it < leff; it++) {
            for (unsigned j = 0; j < N_corr; j++) JJ[j] = 0.0;
            unsigned cont = 0;
            unsigned allinea = (seg * iseg) + (seg * iseg) % skip;
            for (unsigned jm = allinea; jm < seg * (iseg + 1); jm += skip) {
                cont++;
                unsigned idx = 0;
                const double *la = data + (uint64_t)(primo + jm) * ncols;
                const double *lb = data + (uint64_t)(primo + jm + it) * ncols;
                for (unsigned j1 = 0; j1 < ncur; j1++)
                    for (unsigned j2 = 0; j2 < ncur; j2++) {
                        const double *a = la + idx_j[j1], *b = lb + idx_j[j2];
                        double delta = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / 3.0 - JJ[idx];
                        JJ[idx] += delta / (double)cont;
                        idx++;
                    }
            }
            for (unsigned j1 = 0; j1 < N_corr; j1++) {
                double d = JJ[j1] - vdata[(uint64_t)it * narr + j1];
                vdata[(uint64_t)it * narr + j1] += d / (double)cont_JJ;
            }
        }
    }
    if (subtract_mean) {
        unsigned c = 0;
        for (unsigned it = start_mean; it < leff; it++) {
            c++;
            for (unsigned j1 = 0; j1 < N_corr; j1++) {
                double d = vdata[(uint64_t)it * narr + j1] - JJm_T[j1];
                JJm_T[j1] += d / (double)c;
            }
        }
        for (unsigned it = 0; it < leff; it++)
            for (unsigned j = 0; j < N_corr; j++) vdata[(uint64_t)it * narr + j] -= JJm_T[j];
    }
    for (unsigned it = 1; it < leff; it++) {
        double *cur = vdata + (uint64_t)it * narr, *prev = vdata + (uint64_t)(it - 1) * narr;
        for (unsigned j = 0; j < N_corr; j++) {
            cur[N_corr + 2 * j] = prev[N_corr + 2 * j] + prev[j] / 2.0 + cur[j] / 2.0;
            cur[N_corr + 2 * j + 1] = prev[N_corr + 2 * j + 1] + prev[j] * (double)it / 2.0 + cur[j] * (double)it / 2.0;
        }
        for (unsigned j = 0; j < N_corr; j++) matr[j] = 0.0;
        unsigned idx = 0;
        for (unsigned j1 = 0; j1 < ncur; j1++)
            for (unsigned j2 = 0; j2 < ncur; j2++) {
                matr[j2 * ncur + j1] += cur[N_corr + 2 * idx] / 2.0;
                matr[j1 * ncur + j2] += cur[N_corr + 2 * idx] / 2.0;
                idx++;
            }
        cur[3 * N_corr + 0] = schur00(matr, (int)ncur);
        for (unsigned j = 0; j < N_corr; j++) matr[j] = 0.0;
        idx = 0;
        for (unsigned j1 = 0; j1 < ncur; j1++)
            for (unsigned j2 = 0; j2 < ncur; j2++) {
                double v = cur[N_corr + 2 * idx] - cur[N_corr + 2 * idx + 1] / (double)it;
                matr[j2 * ncur + j1] += v / 2.0;
                matr[j1 * ncur + j2] += v / 2.0;
                idx++;
            }
        cur[3 * N_corr + 1] = schur00(matr, (int)ncur);
    }
    for (unsigned it = 1; it < leff; it++)
        for (unsigned j = 0; j < N_corr; j++) vdata[(uint64_t)it * narr + N_corr + 2 * j + 1] /= (double)it;
    free(JJ); free(matr); free(JJm_T);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Block statistics                                                           */
/* ------------------------------------------------------------------------ */

/* calcoliblocchi.h:35-58 MediaVar::calculate on whole vectors (iblock = number
 * of blocks already accumulated). */
void orc_mv_update(double *mean, double *var, const double *x, uint64_t len, unsigned iblock)
{
    for (uint64_t i = 0; i < len; i++) {
        double delta = x[i] - mean[i];
        double tmp = delta / (double)(iblock + 1);
        mean[i] += tmp;
        tmp = x[i] - mean[i];
        tmp *= delta;
        var[i] += tmp;
    }
}

/* calcoliblocchi.h:59-61 */
void orc_mv_end(double *var, uint64_t len, unsigned n_b)
{
    double d = (double)((n_b - 1) * n_b);
    for (uint64_t i = 0; i < len; i++) var[i] /= d;
}

/* calcoliblocchi.h:136-162 MediaVarCovar::calculate; x is one block's
 * [size_block][n_columns] vector; mean/var are [n_columns][size_block];
 * cov is [ncv][size_block]; iblock is the count *after* this block. */
void orc_mvc_update(double *mean, double *var, double *cov, const double *x, uint64_t size_block,
                    unsigned n_columns, const unsigned *cvar_pairs, unsigned ncv, unsigned iblock,
                    double *dx, double *dx2)
{
    for (uint64_t ib = 0; ib < size_block; ib++) {
        for (unsigned ic = 0; ic < n_columns; ic++) {
            double v = x[ib * n_columns + ic];
            dx[ic] = v - mean[(uint64_t)ic * size_block + ib];
            mean[(uint64_t)ic * size_block + ib] += dx[ic] / (double)iblock;
            dx2[ic] = v - mean[(uint64_t)ic * size_block + ib];
            var[(uint64_t)ic * size_block + ib] += dx[ic] * dx2[ic];
        }
        for (unsigned k = 0; k < ncv; k++)
            cov[(uint64_t)k * size_block + ib] += dx[cvar_pairs[2 * k]] * dx2[cvar_pairs[2 * k + 1]];
    }
}

/* calcoliblocchi.h:163-179 */
void orc_mvc_end(double *var, double *cov, uint64_t size_block, unsigned n_columns, unsigned ncv,
                 unsigned iblock)
{
    double d = (double)(iblock * (iblock - 1));
    for (uint64_t ib = 0; ib < size_block; ib++) {
        for (unsigned ic = 0; ic < n_columns; ic++) var[(uint64_t)ic * size_block + ib] /= d;
        for (unsigned k = 0; k < ncv; k++) cov[(uint64_t)k * size_block + ib] /= d;
    }
}

/* ------------------------------------------------------------------------ */
/* Spherical-harmonic density correlations (SphericalCorrelations)          */
/* ------------------------------------------------------------------------ */

#define ORC_PI 3.14159265358979323846264338327950288419716939937510
#define ORC_SH_LMAX 10

/* specialfunctions.h:33-49 -- coefficient of the real harmonic (l, +-m):
 * sqrt((2l+1)/(4 pi)) for m = 0, else sqrt((2l+1)/(4 pi) * 2 / v) (-1)^m with
 * v = (l-m+1)(l-m+2)...(l+m) multiplied up in that order from 1. */
static double sh_coeff(int l, int m)
{
    if (m < 0) m = -m;
    if (m == 0) return sqrt((double)(2 * l + 1) / (4 * ORC_PI));
    double value = 1.0;
    for (int v = l - m + 1; v <= l + m; v++) value = value * v;
    return sqrt((double)(2 * l + 1) / (4 * ORC_PI) * 2 / value) * pow(-1.0, m);
}

/* specialfunctions.h:279-360 (SphericalHarmonics<lmax,double,true,true>) --
 * real spherical harmonics of the direction of (x, y, z):
 *   cos(theta) = z/r, cos(phi) = x/rxy, sin(phi) = y/rxy;
 *   P_l^l = -(2l-1) sqrt(1-c^2) P_{l-1}^{l-1}, P_{m+1}^m = P_m^m c (2m+1),
 *   P_l^m = (P_{l-1}^m c (2l-1) - P_{l-2}^m (l+m-1)) / (l-m)   (:262-318);
 *   cos/sin(m phi) by the Chebyshev recursion (:11-22);
 *   Y_lm = P_l^|m| coeff(l,m) cos(m phi) for m >= 0, sin(|m| phi) for m < 0.
 * Output in the reference's memory order (MultiValDynamic::init, :149-154):
 * for l = lmax down to 0: m = -1, -2, ..., -l, then m = 0, 1, ..., l. */
void orc_sph_ylm(double x, double y, double z, int lmax, double *out)
{
    double P[ORC_SH_LMAX + 1][ORC_SH_LMAX + 1];
    double cs[2 * (ORC_SH_LMAX + 1) + 2];
    double rxy = sqrt(x * x + y * y);
    double r = sqrt(x * x + y * y + z * z);
    double ct = z / r, sp = y / rxy, cp = x / rxy;
    P[0][0] = 1.0;
    for (int l = 1; l <= lmax; l++) P[l][l] = (double)(-(2 * l - 1)) * sqrt(1 - ct * ct) * P[l - 1][l - 1];
    for (int m = 0; m < lmax; m++) {
        P[m + 1][m] = P[m][m] * ct * (2 * m + 1);
        for (int l = m + 2; l <= lmax; l++)
            P[l][m] = (P[l - 1][m] * ct * (2 * l - 1) - P[l - 2][m] * (l + m - 1)) / (l - m);
    }
    cs[0] = 1.0;
    cs[1] = 0.0;
    cs[2] = cp;
    cs[3] = sp;
    for (int i = 2; i <= lmax; i++) {
        cs[2 * i] = 2 * cp * cs[2 * (i - 1)] - cs[2 * (i - 2)];
        cs[2 * i + 1] = 2 * cp * cs[2 * (i - 1) + 1] - cs[2 * (i - 2) + 1];
    }
    int off = 0;
    for (int l = lmax; l >= 0; l--) {
        for (int m = 1; m <= l; m++) out[off + m - 1] = P[l][m] * sh_coeff(l, m) * cs[2 * m + 1];
        for (int m = 0; m <= l; m++) out[off + l + m] = P[l][m] * sh_coeff(l, m) * cs[2 * m];
        off += 2 * l + 1;
    }
}

/* sphericalbase.cpp:19-66 (no neighbour list) -- result[i][type_j][idx][:] is
 * the sum, over the other atoms j in index order, of the harmonics of the
 * minimum image x_i - x_j (frame ts for both) whose distance d falls in bin
 * idx = (int)floorf((d - rmin_p) / dr_p), 0 <= idx < nbin, of the range
 * p = ntypes * type_i + type_j (dr_p = (rmax_p - rmin_p) / nbin, :5-14). */
void orc_sph_density(const orc_traj *tr, uint64_t ts, const double *rminmax, unsigned nbin, int lmax,
                     double *res)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1);
    double y[(ORC_SH_LMAX + 1) * (ORC_SH_LMAX + 1)];
    double *dr = malloc(nt * nt * sizeof(double));
    for (uint64_t p = 0; p < nt * nt; p++) dr[p] = (rminmax[2 * p + 1] - rminmax[2 * p]) / (double)nbin;
    memset(res, 0, N * nt * nbin * L2 * sizeof(double));
    const double *box = tr->box + ts * tr->box_stride;
    for (uint64_t i = 0; i < N; i++) {
        const unsigned ti = (unsigned)tr->type_id[i];
        for (uint64_t j = 0; j < N; j++) {
            const unsigned tj = (unsigned)tr->type_id[j];
            const uint64_t p = nt * ti + tj;
            if (i == j) continue;
            double x[3];
            double d = sqrt(orc_d2_min_image(tr->pos + (ts * N + i) * 3, tr->pos + (ts * N + j) * 3, box,
                                             tr->triclinic, x));
            int idx = (int)floorf((float)((d - rminmax[2 * p]) / dr[p]));
            if (idx < (int)nbin && idx >= 0) {
                orc_sph_ylm(x[0], x[1], x[2], lmax, y);
                double *dst = res + L2 * (nbin * (nt * i + tj) + (uint64_t)idx);
                for (size_t k = 0; k < L2; k++) dst[k] += y[k];
            }
        }
    }
    free(dr);
}

/* sphericalcorrelations.cpp:62-236 and corr_sh_calc :238-270, one thread --
 * for every lag dt < leff and t1 = 0, skip, ... < n (ascending): the product
 * of the densities of frames primo+t1 and primo+t1+dt, element by element;
 * per (atom, type_j, bin) and l the sum over m (the m = 0 term, then
 * m = -1..-l, then m = 1..l; MultiVal::sum_in_m_zero); the running mean over
 * the central atoms of each type (in index order); then the running mean over
 * t1 per lag.  out[dt][type_i][type_j][bin][k] holds l = lmax - k
 * (running_average_m_zero_vector walks l from lmax down).  leff = min(n, tmax)
 * (tmax = 0: n).  Returns -1 when leff + n + primo > nframes (:63-65). */
/* The t1/lag loops of orc_sph_correlations over densities already computed:
 * dens[f] is the density of frame primo + f, f < n + leff - 1. */
void orc_sph_corr_from_densities(const orc_traj *tr, const double *dens, unsigned nbin, int lmax, unsigned leff,
                                 unsigned skip, unsigned n, double *out)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1), L1 = (size_t)lmax + 1;
    if (skip == 0) skip = 1;
    const uint64_t snap = N * nt * nbin * L2;
    const uint64_t fin = nt * nt * nbin * L1;
    double *ave = malloc(fin * sizeof(double));
    double *prod = malloc(L2 * sizeof(double));
    unsigned *cnt = calloc(leff, sizeof(unsigned));
    int *acnt = malloc(nt * sizeof(int));
    memset(out, 0, (size_t)leff * fin * sizeof(double));
    for (unsigned t1 = 0; t1 < n; t1 += skip) {
        for (unsigned dt = 0; dt < leff; dt++) {
            const double *s1 = dens + (uint64_t)t1 * snap, *s2 = dens + (uint64_t)(t1 + dt) * snap;
            memset(ave, 0, fin * sizeof(double));
            for (uint64_t t = 0; t < nt; t++) acnt[t] = 0;
            for (uint64_t i = 0; i < N; i++) {
                const unsigned ti = (unsigned)tr->type_id[i];
                acnt[ti]++;
                for (uint64_t tj = 0; tj < nt; tj++)
                    for (unsigned b = 0; b < nbin; b++) {
                        const uint64_t base = L2 * (nbin * (nt * i + tj) + b);
                        for (size_t k = 0; k < L2; k++) prod[k] = s1[base + k] * s2[base + k];
                        double *av = ave + L1 * (nbin * (nt * ti + tj) + b);
                        size_t off = 0;
                        for (int l = lmax, k = 0; l >= 0; l--, k++) {
                            double v = prod[off + l];
                            for (int m = 0; m < l; m++) v += prod[off + m];
                            for (int m = 0; m < l; m++) v += prod[off + l + 1 + m];
                            av[k] += (v - av[k]) / (double)acnt[ti];
                            off += 2 * (size_t)l + 1;
                        }
                    }
            }
            cnt[dt]++;
            double *o = out + (uint64_t)dt * fin;
            for (uint64_t k = 0; k < fin; k++) o[k] += (ave[k] - o[k]) / (double)cnt[dt];
        }
    }
    free(ave);
    free(prod);
    free(cnt);
    free(acnt);
}

/* sphericalcorrelations.cpp:62-236 and corr_sh_calc :238-270, one thread --
 * for every lag dt < leff and t1 = 0, skip, ... < n (ascending): the product
 * of the densities of frames primo+t1 and primo+t1+dt, element by element;
 * per (atom, type_j, bin) and l the sum over m (the m = 0 term, then
 * m = -1..-l, then m = 1..l; MultiVal::sum_in_m_zero); the running mean over
 * the central atoms of each type (in index order); then the running mean over
 * t1 per lag.  out[dt][type_i][type_j][bin][k] holds l = lmax - k
 * (running_average_m_zero_vector walks l from lmax down).  leff = min(n, tmax)
 * (tmax = 0: n).  Returns -1 when leff + n + primo > nframes (:63-65). */
int orc_sph_correlations(const orc_traj *tr, const double *rminmax, unsigned nbin, int lmax, unsigned tmax,
                         unsigned skip, unsigned n, uint64_t primo, double *out)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1);
    const unsigned leff = (n < tmax || tmax == 0) ? n : tmax;
    if ((uint64_t)leff + n + primo > tr->nframes) return -1;
    const uint64_t nfr = (uint64_t)n + leff - 1, snap = N * nt * nbin * L2;
    double *dens = malloc(nfr * snap * sizeof(double));
    for (uint64_t f = 0; f < nfr; f++) orc_sph_density(tr, primo + f, rminmax, nbin, lmax, dens + f * snap);
    orc_sph_corr_from_densities(tr, dens, nbin, lmax, leff, skip, n, out);
    free(dens);
    return 0;
}

/* steinhardt.cpp:90-236 (no neighbour list, not averaged), one thread -- for
 * frames primo + i, i = 0, skip, ... < n - n % skip (the split of
 * calculatemultithread.h:50-79): the densities (with the neighbour count n_a
 * of every (atom, type_j, bin)), q_l = sqrt(sum_m S_lm^2 / n_a / n_a * 4 * pi
 * / (2l+1)) with the m sum in the order m = 0, -1..-l, 1..l.
 *   do_hist: for n_a > 0, bin index floorf(q_l * nbin_st) (clamped to
 *   nbin_st - 1) of every l in l_list, row-major with the first l slowest;
 *   out[bin][type_i][type_j][hist] += incr, incr = 1/int(n/skip) (1 if 0).
 *   else: out[bin][i/skip][atom][type_j][l-1] = q_l for l = 1..lmax (0 if
 *   n_a = 0).
 * Returns -1 when n + primo > nframes (:80-83). */
int orc_steinhardt(const orc_traj *tr, const double *rminmax, unsigned nbin, int lmax, unsigned nbin_st,
                   const unsigned *l_list, unsigned nl, unsigned skip, unsigned n, uint64_t primo, int do_hist,
                   double *out)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1);
    if ((uint64_t)n + primo > tr->nframes) return -1;
    if (skip == 0) skip = 1;
    const double incr = (n / skip > 0) ? 1.0 / (int)(n / skip) : 1.0;
    size_t hs = 1;
    for (unsigned k = 0; k < nl; k++) hs *= nbin_st;
    const uint64_t nfs = n / skip;
    const size_t len = do_hist ? nbin * nt * nt * hs : nbin * nfs * N * nt * (size_t)lmax;
    memset(out, 0, len * sizeof(double));
    double *dens = malloc(N * nt * nbin * L2 * sizeof(double));
    int *cnt = malloc(N * nt * nbin * sizeof(int));
    double *dr = malloc(nt * nt * sizeof(double));
    for (uint64_t p = 0; p < nt * nt; p++) dr[p] = (rminmax[2 * p + 1] - rminmax[2 * p]) / (double)nbin;
    for (unsigned i = 0; i < n - n % skip; i += skip) {
        const uint64_t ts = primo + i;
        orc_sph_density(tr, ts, rminmax, nbin, lmax, dens);
        /* sphericalbase.cpp:33-60: the counter of the same loop */
        memset(cnt, 0, N * nt * nbin * sizeof(int));
        const double *box = tr->box + ts * tr->box_stride;
        for (uint64_t a = 0; a < N; a++)
            for (uint64_t b = 0; b < N; b++) {
                if (a == b) continue;
                const uint64_t p = nt * (uint64_t)tr->type_id[a] + (uint64_t)tr->type_id[b];
                double x[3];
                double d = sqrt(orc_d2_min_image(tr->pos + (ts * N + a) * 3, tr->pos + (ts * N + b) * 3, box,
                                                 tr->triclinic, x));
                int idx = (int)floorf((float)((d - rminmax[2 * p]) / dr[p]));
                if (idx < (int)nbin && idx >= 0) cnt[nbin * (nt * a + (uint64_t)tr->type_id[b]) + (uint64_t)idx]++;
            }
        for (uint64_t a = 0; a < N; a++) {
            const unsigned ti = (unsigned)tr->type_id[a];
            for (uint64_t tj = 0; tj < nt; tj++)
                for (unsigned b = 0; b < nbin; b++) {
                    const double *s = dens + L2 * (nbin * (nt * a + tj) + b);
                    const double na = (double)cnt[nbin * (nt * a + tj) + b];
                    double q[ORC_SH_LMAX + 1];
                    size_t off = 0;
                    for (int l = lmax; l >= 0; l--) {
                        double v = s[off + l] * s[off + l];
                        for (int m = 0; m < l; m++) v += s[off + m] * s[off + m];
                        for (int m = 0; m < l; m++) v += s[off + l + 1 + m] * s[off + l + 1 + m];
                        q[l] = na > 0 ? sqrt(v / na / na * 4 * ORC_PI / (2 * l + 1)) : 0.0;
                        off += 2 * (size_t)l + 1;
                    }
                    if (do_hist) {
                        if (!(na > 0)) continue;
                        size_t hidx = 0, stride = hs;
                        for (unsigned k = 0; k < nl; k++) {
                            stride /= nbin_st;
                            const float x = floorf((float)(q[l_list[k]] * nbin_st));
                            /* beyond the last cell, and NaN, go to the last cell: what the
                             * size_t conversion and clamp of steinhardt.cpp:199-202 do on x86-64 */
                            const size_t jh = x < (float)nbin_st ? (size_t)x : nbin_st - 1;
                            hidx += jh * stride;
                        }
                        out[((b * nt + ti) * nt + tj) * hs + hidx] += incr;
                    } else {
                        for (int l = 1; l <= lmax; l++)
                            out[(((b * nfs + i / skip) * N + a) * nt + tj) * (size_t)lmax + (size_t)(l - 1)] = q[l];
                    }
                }
        }
    }
    free(dens);
    free(cnt);
    free(dr);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SANN neighbour lists (Neighbours, neighbour.h:10-104, neighbour.cpp)      */
/* ------------------------------------------------------------------------ */

typedef struct {
    double r, x[3];
    uint64_t j;
} orc_neigh;

static int cmp_neigh(const void *a, const void *b)
{
    const orc_neigh *p = a, *q = b;
    if (p->r < q->r) return -1;
    if (p->r > q->r) return 1;
    return p->j < q->j ? -1 : (p->j > q->j);
}

/* neighbour.cpp:85-101 -- the SANN count of a list sorted by distance: 0 for
 * fewer than 3 neighbours, else the smallest n >= 3 with
 * r_0 + .. + r_{n-1} <= r_n (n - 2), or the whole list. */
static unsigned sann_count(const orc_neigh *l, unsigned n)
{
    if (n < 3) return 0;
    unsigned m = 3;
    double s = 0.0;
    for (unsigned i = 0; i < m; i++) s += l[i].r;
    while (m < n) {
        if (s <= l[m].r * (m - 2)) break;
        s += l[m].r;
        m++;
    }
    return m;
}

/* neighbour.cpp:8-82 and sphericalbase.cpp:67-93 -- for every atom i and type
 * t, the atoms j != i of type t with d2(i, j) <= cut2[t] (x = x_i - x_j, the
 * minimum image of frame ts), sorted by distance (ties by index), at most
 * nneigh[t] of them (-1 beyond: "Too many neighbours in shell!"); the
 * harmonics of the first SANN-count ones are summed into bin 0 of
 * res[i][t][nbin][(lmax+1)^2], and the count goes to cnt[i][t] (may be NULL).
 * The capacity check is the reference's own pair loop (neighbour.cpp:15-47):
 * pairs i < j in index order; a pair within either atom type's cutoff throws
 * when i's list of j's type or j's list of i's type is already full; a list
 * counts a partner only within the partner type's cutoff. */
static int sph_density_sann(const orc_traj *tr, uint64_t ts, const unsigned *nneigh, const double *cut2,
                            unsigned nbin, int lmax, double *res, unsigned *cnt, uint64_t *nbr, unsigned capmax)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1);
    double y[(ORC_SH_LMAX + 1) * (ORC_SH_LMAX + 1)];
    unsigned cap = 0;
    for (uint64_t t = 0; t < nt; t++) cap = nneigh[t] > cap ? nneigh[t] : cap;
    orc_neigh *l = malloc((cap + 1) * sizeof(orc_neigh));
    memset(res, 0, N * nt * nbin * L2 * sizeof(double));
    const double *box = tr->box + ts * tr->box_stride;
    int rc = 0;
    unsigned *size = calloc(N * nt, sizeof(unsigned)); /* [atom][type] list lengths */
    for (uint64_t i = 0; i < N && rc == 0; i++)
        for (uint64_t j = i + 1; j < N; j++) {
            const int ti = tr->type_id[i], tj = tr->type_id[j];
            double x[3];
            const double d2 = orc_d2_min_image(tr->pos + (ts * N + i) * 3, tr->pos + (ts * N + j) * 3, box,
                                               tr->triclinic, x);
            if (!(d2 <= cut2[tj] || d2 <= cut2[ti])) continue;
            if (!(size[i * nt + tj] < nneigh[tj] && size[j * nt + ti] < nneigh[ti])) { rc = -1; break; }
            if (d2 <= cut2[tj]) size[i * nt + tj]++;
            if (d2 <= cut2[ti]) size[j * nt + ti]++;
        }
    free(size);
    for (uint64_t i = 0; i < N && rc == 0; i++)
        for (uint64_t t = 0; t < nt && rc == 0; t++) {
            unsigned n = 0;
            for (uint64_t j = 0; j < N; j++) {
                if (j == i || (uint64_t)tr->type_id[j] != t) continue;
                double x[3];
                const double d2 = orc_d2_min_image(tr->pos + (ts * N + i) * 3, tr->pos + (ts * N + j) * 3, box,
                                                   tr->triclinic, x);
                if (!(d2 <= cut2[t])) continue;
                if (n >= nneigh[t]) { rc = -1; break; }
                l[n].r = sqrt(d2);
                l[n].x[0] = x[0];
                l[n].x[1] = x[1];
                l[n].x[2] = x[2];
                l[n].j = j;
                n++;
            }
            if (rc) break;
            qsort(l, n, sizeof(orc_neigh), cmp_neigh);
            const unsigned m = sann_count(l, n);
            double *dst = res + L2 * (nbin * (nt * i + t));
            for (unsigned k = 0; k < m; k++) {
                orc_sph_ylm(l[k].x[0], l[k].x[1], l[k].x[2], lmax, y);
                for (size_t e = 0; e < L2; e++) dst[e] += y[e];
            }
            if (cnt) cnt[nt * i + t] = m;
            if (nbr)
                for (unsigned k = 0; k < m; k++) nbr[(nt * i + t) * capmax + k] = l[k].j;
        }
    free(l);
    return rc;
}

int orc_sph_density_sann(const orc_traj *tr, uint64_t ts, const unsigned *nneigh, const double *cut2,
                         unsigned nbin, int lmax, double *res, unsigned *cnt)
{
    return sph_density_sann(tr, ts, nneigh, cut2, nbin, lmax, res, cnt, NULL, 0);
}

/* orc_sph_correlations with the SANN densities (sphericalcorrelations.cpp
 * with a neighbour list, :115-117): returns -1 too short, -2 too many
 * neighbours. */
int orc_sph_correlations_sann(const orc_traj *tr, const unsigned *nneigh, const double *cut2, unsigned nbin,
                              int lmax, unsigned tmax, unsigned skip, unsigned n, uint64_t primo, double *out)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1);
    const unsigned leff = (n < tmax || tmax == 0) ? n : tmax;
    if ((uint64_t)leff + n + primo > tr->nframes) return -1;
    const uint64_t nfr = (uint64_t)n + leff - 1, snap = N * nt * nbin * L2;
    double *dens = malloc(nfr * snap * sizeof(double));
    for (uint64_t f = 0; f < nfr; f++)
        if (orc_sph_density_sann(tr, primo + f, nneigh, cut2, nbin, lmax, dens + f * snap, NULL)) {
            free(dens);
            return -2;
        }
    orc_sph_corr_from_densities(tr, dens, nbin, lmax, leff, skip, n, out);
    free(dens);
    return 0;
}

/* orc_steinhardt with the SANN densities and counts (steinhardt.cpp with a
 * neighbour list; nbin is 1 there).  averaged (:128-173): per atom and type
 * t, qbar_lm = S_lm(i)/n(i) if type(i) = t, plus S_lm(k)/n(k) of every SANN
 * neighbour k of type t, in list order; then q_l from qbar with n(i) as
 * before.  Returns -1 too short, -2 too many neighbours. */
int orc_steinhardt_sann(const orc_traj *tr, const unsigned *nneigh, const double *cut2, int lmax, unsigned nbin_st,
                        const unsigned *l_list, unsigned nl, unsigned skip, unsigned n, uint64_t primo, int do_hist,
                        int averaged, double *out)
{
    const uint64_t N = tr->natoms, nt = tr->ntypes;
    const size_t L2 = (size_t)(lmax + 1) * (lmax + 1);
    if ((uint64_t)n + primo > tr->nframes) return -1;
    if (skip == 0) skip = 1;
    const double incr = (n / skip > 0) ? 1.0 / (int)(n / skip) : 1.0;
    size_t hs = 1;
    for (unsigned k = 0; k < nl; k++) hs *= nbin_st;
    const uint64_t nfs = n / skip;
    const size_t len = do_hist ? nt * nt * hs : nfs * N * nt * (size_t)lmax;
    memset(out, 0, len * sizeof(double));
    double *dens = malloc(N * nt * L2 * sizeof(double));
    double *avg = malloc(L2 * sizeof(double));
    unsigned *cnt = malloc(N * nt * sizeof(unsigned));
    unsigned capmax = 1;
    for (uint64_t t = 0; t < nt; t++) capmax = nneigh[t] > capmax ? nneigh[t] : capmax;
    uint64_t *nbr = malloc(N * nt * capmax * sizeof(uint64_t));
    int rc = 0;
    for (unsigned i = 0; i < n - n % skip && rc == 0; i += skip) {
        if (sph_density_sann(tr, primo + i, nneigh, cut2, 1, lmax, dens, cnt, nbr, capmax)) { rc = -2; break; }
        for (uint64_t a = 0; a < N; a++) {
            const unsigned ti = (unsigned)tr->type_id[a];
            for (uint64_t tj = 0; tj < nt; tj++) {
                const double *s = dens + L2 * (nt * a + tj);
                if (averaged) {
                    const unsigned na_ = cnt[nt * a + tj];
                    for (size_t e = 0; e < L2; e++) avg[e] = ti == tj ? s[e] / na_ : 0.0;
                    for (unsigned k = 0; k < na_; k++) {
                        const uint64_t b = nbr[(nt * a + tj) * capmax + k];
                        const double nb = (double)cnt[nt * b + tj];
                        const double *sb = dens + L2 * (nt * b + tj);
                        for (size_t e = 0; e < L2; e++) avg[e] += sb[e] / nb;
                    }
                    s = avg;
                }
                const double na = (double)cnt[nt * a + tj];
                double q[ORC_SH_LMAX + 1];
                size_t off = 0;
                for (int l = lmax; l >= 0; l--) {
                    double v = s[off + l] * s[off + l];
                    for (int m = 0; m < l; m++) v += s[off + m] * s[off + m];
                    for (int m = 0; m < l; m++) v += s[off + l + 1 + m] * s[off + l + 1 + m];
                    q[l] = na > 0 ? sqrt(v / na / na * 4 * ORC_PI / (2 * l + 1)) : 0.0;
                    off += 2 * (size_t)l + 1;
                }
                if (do_hist) {
                    if (!(na > 0)) continue;
                    size_t hidx = 0, stride = hs;
                    for (unsigned k = 0; k < nl; k++) {
                        stride /= nbin_st;
                        const float x = floorf((float)(q[l_list[k]] * nbin_st));
                        const size_t jh = x < (float)nbin_st ? (size_t)x : nbin_st - 1;  /* NaN: last cell */
                        hidx += jh * stride;
                    }
                    out[(ti * nt + tj) * hs + hidx] += incr;
                } else {
                    for (int l = 1; l <= lmax; l++)
                        out[(((uint64_t)(i / skip) * N + a) * nt + tj) * (size_t)lmax + (size_t)(l - 1)] = q[l];
                }
            }
        }
    }
    free(dens);
    free(avg);
    free(cnt);
    free(nbr);
    return rc;
}
