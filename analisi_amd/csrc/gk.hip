// Green-Kubo N-component lag correlations (greenkuboNcomponentionicfluid.cpp:
// 156-356) on the device, in the reference's FP64 operation order so that the
// correlation functions, their integrals and the transport coefficients match
// the reference's values:
//  * gk_lag_kernel: one thread per lag l holds the ncur^2 running means
//      JJ[ab] += ((Ja(t0).Jb(t0+l))/3 - JJ[ab]) / cont      (:200-228)
//    over t0 = align, align+skip, .. < seg*(iseg+1) of each of the n_seg
//    segments, and the running mean of the segment means (:230-237).
//    Rows t0 are shared by all threads of a workgroup (broadcast loads), rows
//    t0+l are consecutive across threads.
//  * gk_mean_kernel / gk_integral_kernel / gk_schur_kernel: the serial
//    epilogue (:259-356) as column-parallel prefix scans and a lag-parallel
//    Schur complement of the symmetrised integral matrix.
#include <algorithm>

#include "anl_internal.h"

#pragma clang fp contract(off)

namespace anl {

constexpr int GK_MAXC = 4;  // up to 4 vector currents (16 correlation pairs)

struct GkArgs {
    const double *data;
    uint32_t ncols;
    uint32_t ncur;
    uint32_t idx[GK_MAXC];
    uint64_t primo;
    uint32_t n, n_seg, skip, leff, narr;
    double *out;  // [leff][narr]
};

template <int NC>
__global__ __launch_bounds__(256) void gk_lag_kernel(const GkArgs a) {
    const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= a.leff) return;
    constexpr int NP = NC * NC;
    double v[NP], JJ[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) v[k] = 0.0;
    const uint32_t seg = a.n / a.n_seg;
    uint32_t cont_JJ = 0;
    for (uint32_t iseg = 0; iseg < a.n_seg; iseg++) {
        ++cont_JJ;
#pragma unroll
        for (int k = 0; k < NP; k++) JJ[k] = 0.0;
        uint32_t cont = 0;
        const uint32_t allinea = (seg * iseg) + (seg * iseg) % a.skip;
        for (uint32_t jm = allinea; jm < seg * (iseg + 1); jm += a.skip) {
            cont++;
            const double *la = a.data + (a.primo + jm) * a.ncols;
            const double *lb = a.data + (a.primo + jm + it) * a.ncols;
            double A[NC][3], B[NC][3];
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    A[c][x] = la[a.idx[c] + x];
                    B[c][x] = lb[a.idx[c] + x];
                }
            const double dc = (double)cont;
#pragma unroll
            for (int j1 = 0; j1 < NC; j1++)
#pragma unroll
                for (int j2 = 0; j2 < NC; j2++) {
                    const int k = j1 * NC + j2;
                    double p = A[j1][0] * B[j2][0];
                    p = p + A[j1][1] * B[j2][1];
                    p = p + A[j1][2] * B[j2][2];
                    const double delta = p / 3.0 - JJ[k];
                    JJ[k] = JJ[k] + delta / dc;
                }
        }
        const double dcj = (double)cont_JJ;
#pragma unroll
        for (int k = 0; k < NP; k++) v[k] = v[k] + (JJ[k] - v[k]) / dcj;
    }
    double *o = a.out + (uint64_t)it * a.narr;
#pragma unroll
    for (int k = 0; k < NP; k++) o[k] = v[k];
    for (uint32_t k = NP; k < a.narr; k++) o[k] = 0.0;
}

// :259-277 subtract the mean of C over lags [start_mean, leff)
__global__ void gk_mean_kernel(double *__restrict__ out, uint32_t narr, uint32_t ncorr, uint32_t leff,
                               uint32_t start_mean) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncorr) return;
    double m = 0.0;
    uint32_t c = 0;
    for (uint32_t it = start_mean; it < leff; it++) {
        c++;
        const double d = out[(uint64_t)it * narr + j] - m;
        m = m + d / (double)c;
    }
    for (uint32_t it = 0; it < leff; it++) out[(uint64_t)it * narr + j] = out[(uint64_t)it * narr + j] - m;
}

// :279-288 trapezoid I and Einstein E running integrals, per correlation column
__global__ void gk_integral_kernel(double *__restrict__ out, uint32_t narr, uint32_t ncorr, uint32_t leff) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncorr) return;
    for (uint32_t it = 1; it < leff; it++) {
        double *cur = out + (uint64_t)it * narr;
        const double *prev = out + (uint64_t)(it - 1) * narr;
        const double dit = (double)it;
        double I = prev[ncorr + 2 * j] + prev[j] / 2.0;
        cur[ncorr + 2 * j] = I + cur[j] / 2.0;
        double E = prev[ncorr + 2 * j + 1] + prev[j] * dit / 2.0;
        cur[ncorr + 2 * j + 1] = E + cur[j] * dit / 2.0;
    }
}

// k = c00 - c(0,1:) inv(c(1:,1:)) c(1:,0), c column-major n x n; the inverse
// by LU with partial pivoting (Eigen's dynamic-size inverse); for n == 2 this
// is c00 - (c01 * (1/c11)) * c10.
__device__ double schur00(const double *c, int n) {
    if (n <= 1) return c[0];
    const int m = n - 1;
    double lu[(GK_MAXC - 1) * (GK_MAXC - 1)], inv[(GK_MAXC - 1) * (GK_MAXC - 1)], tmp[GK_MAXC - 1];
    int perm[GK_MAXC - 1];
    for (int cc = 0; cc < m; cc++)
        for (int r = 0; r < m; r++) lu[cc * m + r] = c[(cc + 1) * n + (r + 1)];
    for (int i = 0; i < m; i++) perm[i] = i;
    for (int k = 0; k < m; k++) {
        int p = k;
        double best = fabs(lu[k * m + k]);
        for (int r = k + 1; r < m; r++)
            if (fabs(lu[k * m + r]) > best) {
                best = fabs(lu[k * m + r]);
                p = r;
            }
        if (p != k) {
            for (int cc = 0; cc < m; cc++) {
                double t = lu[cc * m + k];
                lu[cc * m + k] = lu[cc * m + p];
                lu[cc * m + p] = t;
            }
            int t = perm[k];
            perm[k] = perm[p];
            perm[p] = t;
        }
        for (int r = k + 1; r < m; r++) lu[k * m + r] /= lu[k * m + k];
        for (int cc = k + 1; cc < m; cc++)
            for (int r = k + 1; r < m; r++) lu[cc * m + r] -= lu[k * m + r] * lu[cc * m + k];
    }
    for (int col = 0; col < m; col++) {
        double *x = inv + col * m;
        for (int r = 0; r < m; r++) x[r] = (perm[r] == col) ? 1.0 : 0.0;
        for (int r = 0; r < m; r++)
            for (int k = 0; k < r; k++) x[r] -= lu[k * m + r] * x[k];
        for (int r = m - 1; r >= 0; r--) {
            for (int k = r + 1; k < m; k++) x[r] -= lu[k * m + r] * x[k];
            x[r] = x[r] * (1.0 / lu[r * m + r]);
        }
    }
    for (int cc = 0; cc < m; cc++) {
        double s = c[1 * n + 0] * inv[cc * m + 0];
        for (int k = 1; k < m; k++) s += c[(k + 1) * n + 0] * inv[cc * m + k];
        tmp[cc] = s;
    }
    double s = tmp[0] * c[0 * n + 1];
    for (int cc = 1; cc < m; cc++) s += tmp[cc] * c[0 * n + (cc + 1)];
    return c[0] - s;
}

// :289-346 k_gk and k_einst per lag, then :352-356 E /= l
__global__ void gk_schur_kernel(double *__restrict__ out, uint32_t narr, uint32_t ncur, uint32_t leff) {
    const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it == 0 || it >= leff) return;
    const uint32_t ncorr = ncur * ncur;
    double *cur = out + (uint64_t)it * narr;
    double matr[GK_MAXC * GK_MAXC];
    for (uint32_t j = 0; j < ncorr; j++) matr[j] = 0.0;
    uint32_t idx = 0;
    for (uint32_t j1 = 0; j1 < ncur; j1++)
        for (uint32_t j2 = 0; j2 < ncur; j2++) {
            matr[j2 * ncur + j1] += cur[ncorr + 2 * idx] / 2.0;
            matr[j1 * ncur + j2] += cur[ncorr + 2 * idx] / 2.0;
            idx++;
        }
    const double kgk = schur00(matr, (int)ncur);
    for (uint32_t j = 0; j < ncorr; j++) matr[j] = 0.0;
    idx = 0;
    const double dit = (double)it;
    for (uint32_t j1 = 0; j1 < ncur; j1++)
        for (uint32_t j2 = 0; j2 < ncur; j2++) {
            const double v = cur[ncorr + 2 * idx] - cur[ncorr + 2 * idx + 1] / dit;
            matr[j2 * ncur + j1] += v / 2.0;
            matr[j1 * ncur + j2] += v / 2.0;
            idx++;
        }
    const double kein = schur00(matr, (int)ncur);
    cur[3 * ncorr + 0] = kgk;
    cur[3 * ncorr + 1] = kein;
    for (uint32_t j = 0; j < ncorr; j++) cur[ncorr + 2 * j + 1] = cur[ncorr + 2 * j + 1] / dit;
}

void gk_epilogue(double *d, uint32_t narr, uint32_t ncur, uint32_t leff, int subtract_mean, uint32_t start_mean,
                 hipStream_t s) {
    const uint32_t ncorr = ncur * ncur;
    if (subtract_mean) {
        hipLaunchKernelGGL(gk_mean_kernel, dim3(1), dim3(64), 0, s, d, narr, ncorr, leff, start_mean);
        ANL_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(gk_integral_kernel, dim3(1), dim3(64), 0, s, d, narr, ncorr, leff);
    ANL_HIP(hipGetLastError());
    hipLaunchKernelGGL(gk_schur_kernel, dim3((leff + 255) / 256), dim3(256), 0, s, d, narr, ncur, leff);
    ANL_HIP(hipGetLastError());
}

}  // namespace anl

using namespace anl;

namespace {

uint32_t gk_fp64_checks(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t tmax, uint32_t n,
                        uint64_t primo) {
    if (!s || !idx_j) fail(ANL_ERR_ARG, "null pointer passed to anl_gk_calculate");
    if (ncur == 0 || ncur > GK_MAXC) fail(ANL_ERR_ARG, "between 1 and 4 currents are supported");
    for (uint32_t c = 0; c < ncur; c++)
        if (idx_j[c] + 3 > s->ncols) fail(ANL_ERR_HEADER, "current column triplet out of range");
    const uint32_t leff = (n < tmax || tmax == 0) ? n : tmax;  // :123
    if ((uint64_t)leff + primo + n > s->nrows)                 // :168-169
        fail(ANL_ERR_TOO_SHORT,
             "trajectory is too short for this kind of calculation. Select a different starting timestep or "
             "lower the size of the average or the lenght of the correlation function");
    return leff;
}

void gk_fp64_enqueue(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t leff, uint32_t n,
                     uint32_t n_seg, int subtract_mean, uint32_t start_mean, uint64_t primo, double *d,
                     hipStream_t st) {
    const uint32_t narr = 3 * ncur * ncur + 2;
    GkArgs a{};
    a.data = s->data;
    a.ncols = s->ncols;
    a.ncur = ncur;
    for (uint32_t c = 0; c < ncur; c++) a.idx[c] = idx_j[c];
    a.primo = primo;
    a.n = n;
    a.n_seg = n_seg;
    a.skip = skip;
    a.leff = leff;
    a.narr = narr;
    a.out = d;
    dim3 grid((leff + 255) / 256);
    switch (ncur) {
        case 1: hipLaunchKernelGGL(gk_lag_kernel<1>, grid, dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL(gk_lag_kernel<2>, grid, dim3(256), 0, st, a); break;
        case 3: hipLaunchKernelGGL(gk_lag_kernel<3>, grid, dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL(gk_lag_kernel<4>, grid, dim3(256), 0, st, a); break;
    }
    ANL_HIP(hipGetLastError());
    gk_epilogue(d, narr, ncur, leff, subtract_mean, start_mean, st);
}

}  // namespace

extern "C" {

int anl_gk_calculate_async(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t tmax,
                           uint32_t n, uint32_t n_seg, int subtract_mean, uint32_t start_mean, uint64_t primo,
                           double *out_dev, void *stream) {
    return guard([&] {
        const uint32_t leff = gk_fp64_checks(s, idx_j, ncur, tmax, n, primo);
        if (!out_dev) fail(ANL_ERR_ARG, "null output passed to anl_gk_calculate_async");
        if (n_seg < 1) n_seg = 1;  // greenkubo...cpp:50-53
        if (skip == 0) skip = 1;
        if (leff == 0) return;
        ensure_device();
        gk_fp64_enqueue(s, idx_j, ncur, skip, leff, n, n_seg, subtract_mean, start_mean, primo, out_dev,
                        (hipStream_t)stream);
    });
}

int anl_gk_calculate(anl_series *s, const uint32_t *idx_j, uint32_t ncur, uint32_t skip, uint32_t tmax, uint32_t n,
                     uint32_t n_seg, int subtract_mean, uint32_t start_mean, uint64_t primo, double *out_host) {
    return guard([&] {
        const uint32_t leff = gk_fp64_checks(s, idx_j, ncur, tmax, n, primo);
        if (!out_host) fail(ANL_ERR_ARG, "null output passed to anl_gk_calculate");
        if (n_seg < 1) n_seg = 1;
        if (skip == 0) skip = 1;
        if (leff == 0) return;
        ensure_device();
        const size_t bytes = (size_t)leff * (3 * ncur * ncur + 2) * sizeof(double);
        double *d = (double *)series_workspace(s, bytes);
        gk_fp64_enqueue(s, idx_j, ncur, skip, leff, n, n_seg, subtract_mean, start_mean, primo, d, 0);
        ANL_HIP(hipMemcpy(out_host, d, bytes, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
