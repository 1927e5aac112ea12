// Runtime of libanalisi_hip: errors, device selection, trajectory and time
// series mirrors, block statistics kernels.
#include <algorithm>
#include <array>
#include <thread>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "anl_internal.h"

namespace anl {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(ANL_ERR_HIP, std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
}

static int g_device = -1;

void ensure_device() {
    if (g_device >= 0) {
        ANL_HIP(hipSetDevice(g_device));
        return;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail(ANL_ERR_NO_DEVICE, "no HIP device visible: the analisi_amd kernels need an MI355X (gfx950)");
    hipDeviceProp_t p;
    ANL_HIP(hipGetDeviceProperties(&p, 0));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        fail(ANL_ERR_NO_DEVICE, std::string("device 0 is ") + p.gcnArchName + ", the kernels are built for gfx950");
    g_device = 0;
    ANL_HIP(hipSetDevice(0));
}

Cell cell_of(const double *b) {
    Cell c;
    for (int k = 0; k < 3; k++) {
        c.lo[k] = b[k];
        c.lh[k] = b[3 + k];
        c.L[k] = b[3 + k] * 2;
    }
    c.xy = b[6];
    c.xz = b[7];
    c.yz = b[8];
    return c;
}

// Fractional coordinates of x in the cell (x = lo + A s with columns
// a=(Lx,0,0), b=(xy,Ly,0), c=(xz,yz,Lz)).
__host__ __device__ static inline void to_fractional(const Cell &c, double x, double y, double z, double *s) {
    double sc = (z - c.lo[2]) / c.L[2];
    double sb = (y - c.lo[1] - c.yz * sc) / c.L[1];
    double sa = (x - c.lo[0] - c.xy * sb - c.xz * sc) / c.L[0];
    s[0] = sa;
    s[1] = sb;
    s[2] = sc;
}

// 32-bit fixed point of the fractional part: round(frac(s) * 2^32) mod 2^32
// (s - floor(s) and the power-of-two scaling are exact in FP64).
__device__ static inline uint32_t fix32(double s) {
    double f = s - floor(s);
    double u = rint(f * 4294967296.0);
    return (uint32_t)(uint64_t)u;
}

// Per screen slot k = (frame, s): the atom at slot order[k] of that frame
// (slot s itself without an order) as {a, b, c} 32-bit fixed-point fractions
// (frac32) and the same rounded to 16 bits, packed {a | b << 16, c} (frac16).
// out32.w = type | slot << 8, so that one 16-byte load gives a candidate's
// position, type and identity.  outpos holds the FP64 positions in the same
// order, for the exact re-evaluation of candidates near a bin edge (reads of
// neighbouring screen slots then share cache lines).
__global__ void frac_kernel(const double4 *__restrict__ pos, const int32_t *__restrict__ type_slot,
                            const int32_t *__restrict__ order, uint64_t natoms, uint2 *__restrict__ out16,
                            uint4 *__restrict__ out32, double4 *__restrict__ outpos, uint64_t total, Cell c) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const uint64_t f = k / natoms;
    const uint32_t sl = order ? (uint32_t)order[k] : (uint32_t)(k - f * natoms);
    double4 x = pos[f * natoms + sl];
    double s[3];
    to_fractional(c, x.x, x.y, x.z, s);
    uint32_t a = fix32(s[0]), b = fix32(s[1]), cc = fix32(s[2]);
    uint32_t a16 = ((a + 0x8000u) >> 16) & 0xFFFFu;
    uint32_t b16 = ((b + 0x8000u) >> 16) & 0xFFFFu;
    uint32_t c16 = ((cc + 0x8000u) >> 16) & 0xFFFFu;
    out16[k] = make_uint2(a16 | (b16 << 16), c16);
    out32[k] = make_uint4(a, b, cc, (uint32_t)type_slot[sl] | (sl << 8));
    outpos[k] = x;
}

static __device__ __forceinline__ int wrap16(uint32_t x) { return (int)(short)(x & 0xFFFFu); }

// Bounding sphere of each group of 32 consecutive screen slots of a frame
// (the last group may be shorter): a 16-bit fractional centre c_g (the
// first member plus the rounded mean wrapped offset of the members) and the
// radius R_g = max |A wrap16(f_j - c_g)| 2^-16 over the members, in FP64,
// rounded up to float.  Groups are compact because each frame's screen slots
// follow that frame's k-d order.
__global__ void jgroup_kernel(const uint2 *__restrict__ frac16, uint64_t natoms, uint32_t ngroups, uint64_t total,
                              Cell c, uint4 *__restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const uint64_t f = k / ngroups;
    const uint32_t g = (uint32_t)(k - f * ngroups);
    const uint2 *p = frac16 + f * natoms + 32ull * g;
    const uint32_t n = (uint32_t)min<uint64_t>(32, natoms - 32ull * g);
    const uint2 r = p[0];
    int sa = 0, sb = 0, sc = 0;
    for (uint32_t m = 0; m < n; m++) {
        sa += wrap16(p[m].x - r.x);
        sb += wrap16((p[m].x >> 16) - (r.x >> 16));
        sc += wrap16(p[m].y - r.y);
    }
    const uint32_t ca = (r.x + (uint32_t)__double2int_rn((double)sa / n)) & 0xFFFFu;
    const uint32_t cb = ((r.x >> 16) + (uint32_t)__double2int_rn((double)sb / n)) & 0xFFFFu;
    const uint32_t cc = (r.y + (uint32_t)__double2int_rn((double)sc / n)) & 0xFFFFu;
    double R2 = 0.0;
    for (uint32_t m = 0; m < n; m++) {
        const double da = wrap16(p[m].x - ca), db = wrap16((p[m].x >> 16) - cb), dc = wrap16(p[m].y - cc);
        const double x = (c.L[0] * da + c.xy * db + c.xz * dc) * 0x1p-16;
        const double y = (c.L[1] * db + c.yz * dc) * 0x1p-16;
        const double z = (c.L[2] * dc) * 0x1p-16;
        R2 = fmax(R2, x * x + y * y + z * z);
    }
    const float R = __double2float_ru(sqrt(R2) * (1.0 + 1e-9));
    out[k] = make_uint4(ca | (cb << 16), cc, __float_as_uint(R), 0u);
}

void build_frac16(anl_traj *tr, hipStream_t s) {
    if (tr->frac16) return;
    uint64_t total = tr->F * tr->N;
    ANL_HIP(hipMalloc(&tr->frac16, total * sizeof(uint2)));
    ANL_HIP(hipMalloc(&tr->frac32, total * sizeof(uint4)));
    ANL_HIP(hipMalloc(&tr->pos_scr, total * sizeof(double4)));
    const uint32_t ngroups = (uint32_t)((tr->N + 31) / 32);
    const uint64_t gtotal = tr->F * ngroups;
    ANL_HIP(hipMalloc(&tr->jgroups, gtotal * sizeof(uint4)));
    unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(frac_kernel, dim3(blocks), dim3(256), 0, s, tr->pos, tr->type_slot, tr->frac_order, tr->N,
                       tr->frac16, tr->frac32, tr->pos_scr, total, tr->cell0);
    ANL_HIP(hipGetLastError());
    hipLaunchKernelGGL(jgroup_kernel, dim3((unsigned)((gtotal + 255) / 256)), dim3(256), 0, s, tr->frac16, tr->N,
                       ngroups, gtotal, tr->cell0, tr->jgroups);
    ANL_HIP(hipGetLastError());
    ANL_HIP(hipStreamSynchronize(s));
    (void)hipFree(tr->frac_order);  // the fractions hold the order from now on
    tr->frac_order = nullptr;
}

// Recursive median bisection of order[lo, hi) (see the slot order below).
static void kd_order(int32_t *order, const std::array<float, 3> *fr, size_t lo, size_t hi, const double len[3]) {
    const size_t n = hi - lo;
    if (n <= 32) return;
    float mn[3] = {2.0f, 2.0f, 2.0f}, mx[3] = {-1.0f, -1.0f, -1.0f};
    for (size_t i = lo; i < hi; i++)
        for (int k = 0; k < 3; k++) {
            mn[k] = std::min(mn[k], fr[order[i]][k]);
            mx[k] = std::max(mx[k], fr[order[i]][k]);
        }
    int ax = 0;
    for (int k = 1; k < 3; k++)
        if ((mx[k] - mn[k]) * len[k] > (mx[ax] - mn[ax]) * len[ax]) ax = k;
    const size_t half = (n + 63) / 64 * 32;  // a multiple of 32, < n
    std::nth_element(order + lo, order + lo + half, order + hi, [&](int32_t a, int32_t b) {
        return fr[a][ax] < fr[b][ax] || (fr[a][ax] == fr[b][ax] && a < b);
    });
    kd_order(order, fr, lo, lo + half, len);
    kd_order(order, fr, lo + half, hi, len);
}


__global__ void mediavar_update_kernel(double *__restrict__ mean, double *__restrict__ var,
                                       const double *__restrict__ x, uint64_t len, double bp1) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    // calcoliblocchi.h:35-58, element-wise, same operation order
    double delta = x[i] - mean[i];
    double tmp = delta / bp1;
    double m = mean[i] + tmp;
    mean[i] = m;
    tmp = x[i] - m;
    tmp = tmp * delta;
    var[i] = var[i] + tmp;
}

__global__ void div_kernel(double *__restrict__ v, uint64_t len, double d) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) v[i] = v[i] / d;
}

__global__ void counts_to_double_kernel(const unsigned long long *__restrict__ c, double incr, uint64_t len,
                                        double *__restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) out[i] = (double)c[i] * incr;
}

}  // namespace anl

using namespace anl;

namespace anl {
void *series_workspace(anl_series *s, size_t bytes) {
    if (bytes <= s->work_cap) return s->work;
    ANL_HIP(hipDeviceSynchronize());  // earlier async calls may still read the old buffer
    (void)hipFree(s->work);
    s->work = nullptr;
    s->work_cap = 0;
    ANL_HIP(hipMalloc(&s->work, bytes));
    s->work_cap = bytes;
    return s->work;
}
}  // namespace anl

extern "C" {

const char *anl_last_error(void) { return g_last_error.c_str(); }

int anl_init(int device) {
    return guard([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
            fail(ANL_ERR_NO_DEVICE, "no HIP device visible: the analisi_amd kernels need an MI355X (gfx950)");
        if (device < 0 || device >= n) fail(ANL_ERR_ARG, "device index out of range");
        hipDeviceProp_t p;
        ANL_HIP(hipGetDeviceProperties(&p, device));
        if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
            fail(ANL_ERR_NO_DEVICE, std::string("device is ") + p.gcnArchName + ", the kernels are built for gfx950");
        g_device = device;
        ANL_HIP(hipSetDevice(device));
    });
}

const char *anl_info(void) {
    return "analisi_amd: MI355X (gfx950) g(r,t) / MSD / Green-Kubo lag-correlation path; "
           "restates rikigigi/analisi v0.5.2 behaviour (Gofrt, MSD, GreenKuboNComponentIonicFluid, MediaVar)";
}

}  // extern "C"

namespace anl {

// Host-to-device copies of anl_traj_create: a non-blocking stream of its own,
// so that they do not wait for kernels on the default stream (a trajectory
// window prepared on a host thread while the current window's kernels run),
// and two pinned staging buffers filled and copied alternately.
struct Uploader {
    hipStream_t s = nullptr;
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    size_t cap;
    int k = 0;
    explicit Uploader(size_t bytes) : cap(std::max<size_t>(bytes, 1 << 16)) {
        ANL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (int i = 0; i < 2; i++) {
            ANL_HIP(hipHostMalloc(&stage[i], cap, hipHostMallocDefault));
            ANL_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
        }
    }
    ~Uploader() {
        if (s) (void)hipStreamSynchronize(s);
        for (int i = 0; i < 2; i++) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (stage[i]) (void)hipHostFree(stage[i]);
        }
        if (s) (void)hipStreamDestroy(s);
    }
    Uploader(const Uploader &) = delete;
    Uploader &operator=(const Uploader &) = delete;
    // the staging buffer of the next copy, once its previous copy has ended
    void *next() {
        ANL_HIP(hipEventSynchronize(done[k]));
        return stage[k];
    }
    // copies the first bytes of that buffer to dst
    void send(void *dst, size_t bytes) {
        ANL_HIP(hipMemcpyAsync(dst, stage[k], bytes, hipMemcpyHostToDevice, s));
        ANL_HIP(hipEventRecord(done[k], s));
        k ^= 1;
    }
    void put(void *dst, const void *src, size_t bytes) {
        for (size_t off = 0; off < bytes; off += cap) {
            const size_t n = std::min(cap, bytes - off);
            std::memcpy(next(), (const char *)src + off, n);
            send((char *)dst + off, n);
        }
    }
    void finish() { ANL_HIP(hipStreamSynchronize(s)); }
};

}  // namespace anl

extern "C" {

int anl_traj_create(const double *pos, const double *box, uint32_t box_stride, const int32_t *type_id,
                    uint32_t ntypes, const double *pos_cm, uint64_t nframes, uint64_t natoms, anl_traj **out) {
    return guard([&] {
        if (!pos || !box || !type_id || !out) fail(ANL_ERR_ARG, "null pointer passed to anl_traj_create");
        if (box_stride != 6 && box_stride != 9) fail(ANL_ERR_ARG, "box stride must be 6 or 9");
        if (nframes == 0 || natoms == 0 || ntypes == 0) fail(ANL_ERR_ARG, "empty trajectory");
        if (natoms >= (1u << 24)) fail(ANL_ERR_ARG, "at most 2^24-1 atoms are supported");
        for (uint64_t i = 0; i < natoms; i++)
            if (type_id[i] < 0 || (uint32_t)type_id[i] >= ntypes) fail(ANL_ERR_ARG, "type id out of range");
        // The device min-image loops step by one cell length per iteration;
        // bounded, finite inputs keep every loop short and finite.
        double lmin = HUGE_VAL;
        for (uint64_t f = 0; f < nframes; f++)
            for (int k = 3; k < 6; k++) {
                double h = box[f * box_stride + k];
                if (!std::isfinite(h) || !(h > 0)) fail(ANL_ERR_ARG, "cell half-lengths must be finite and positive");
                lmin = std::min(lmin, h);
            }
        for (uint64_t f = 0; f < nframes; f++)
            for (uint32_t k = 6; k < box_stride; k++)
                if (!std::isfinite(box[f * box_stride + k])) fail(ANL_ERR_ARG, "cell tilts must be finite");
        const double bound = 1e4 * 2 * lmin;
        double pos_max = 0.0;
        for (uint64_t i = 0; i < nframes * natoms * 3; i++) {
            const double v = std::fabs(pos[i]);
            if (!(v <= bound))
                fail(ANL_ERR_ARG, "positions must be finite and within 10^4 (shortest) cell lengths of the origin");
            pos_max = std::max(pos_max, v);
        }
        ensure_device();
        // on any failure below (a device allocation, an upload) everything
        // allocated so far is released with the trajectory
        std::unique_ptr<anl_traj, void (*)(anl_traj *)> owner(new anl_traj(), [](anl_traj *t) {
            (void)hipDeviceSynchronize();
            (void)anl_traj_destroy(t);
        });
        anl_traj *tr = owner.get();
        tr->F = nframes;
        tr->N = natoms;
        tr->nt = ntypes;
        tr->stride = box_stride;
        tr->triclinic = box_stride == 9;
        tr->pos_max = pos_max;
        tr->box_host.assign(nframes * 9, 0.0);
        for (uint64_t f = 0; f < nframes; f++)
            for (uint32_t k = 0; k < box_stride; k++) tr->box_host[f * 9 + k] = box[f * box_stride + k];
        tr->box_const = true;
        for (uint64_t f = 1; f < nframes && tr->box_const; f++)
            tr->box_const = std::memcmp(&tr->box_host[f * 9], &tr->box_host[0], 9 * sizeof(double)) == 0;
        tr->cell0 = cell_of(&tr->box_host[0]);

        // slot order: balanced k-d bisection of the frame-0 fractional
        // positions (median cut along the widest cell-scaled extent, every
        // cut at a multiple of 32 atoms), so that each run of 32 or 64
        // consecutive slots is a compact box
        std::vector<std::array<float, 3>> fr(natoms);
        for (uint64_t i = 0; i < natoms; i++) {
            double s3[3];
            to_fractional(tr->cell0, pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2], s3);
            for (int k = 0; k < 3; k++) fr[i][k] = (float)(s3[k] - std::floor(s3[k]));
        }
        std::vector<int32_t> order(natoms);
        std::iota(order.begin(), order.end(), 0);
        const double len[3] = {tr->cell0.L[0], tr->cell0.L[1], tr->cell0.L[2]};
        kd_order(order.data(), fr.data(), 0, natoms, len);
        tr->type_of_slot.resize(natoms);
        tr->type_count.assign(ntypes, 0);
        for (uint64_t s = 0; s < natoms; s++) {
            tr->type_of_slot[s] = type_id[order[s]];
            tr->type_count[type_id[order[s]]]++;
        }
        // permuted, padded positions, staged frame by frame in pinned memory
        Uploader up(natoms * sizeof(double4));
        ANL_HIP(hipMalloc(&tr->pos, nframes * natoms * sizeof(double4)));
        for (uint64_t f = 0; f < nframes; f++) {
            const double *p = pos + f * natoms * 3;
            double4 *buf = (double4 *)up.next();
            for (uint64_t s = 0; s < natoms; s++) {
                const double *x = p + (uint64_t)order[s] * 3;
                buf[s] = make_double4(x[0], x[1], x[2], 0.0);
            }
            up.send(tr->pos + f * natoms, natoms * sizeof(double4));
        }
        ANL_HIP(hipMalloc(&tr->box, nframes * 9 * sizeof(double)));
        up.put(tr->box, tr->box_host.data(), nframes * 9 * sizeof(double));
        ANL_HIP(hipMalloc(&tr->type_slot, natoms * sizeof(int32_t)));
        up.put(tr->type_slot, tr->type_of_slot.data(), natoms * sizeof(int32_t));
        ANL_HIP(hipMalloc(&tr->orig_slot, natoms * sizeof(int32_t)));
        up.put(tr->orig_slot, order.data(), natoms * sizeof(int32_t));
        // screen order, per frame (the screened g(r,t) needs a constant
        // cell): the same bisection of that frame's positions, so that the
        // atoms of one wavefront stay compact in every frame while they
        // diffuse.  frac_order[f][s] = slot of the atom at screen slot s.
        if (tr->box_const) {
            std::vector<int32_t> ford(nframes * natoms);
            const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (unsigned k = 0; k < nth; k++)
                pool.emplace_back([&, k] {
                    std::vector<std::array<float, 3>> fs(natoms);
                    for (uint64_t f = k; f < nframes; f += nth) {
                        const double *p = pos + f * natoms * 3;
                        for (uint64_t sl = 0; sl < natoms; sl++) {
                            const double *x = p + (uint64_t)order[sl] * 3;
                            double s3[3];
                            to_fractional(tr->cell0, x[0], x[1], x[2], s3);
                            for (int c = 0; c < 3; c++) fs[sl][c] = (float)(s3[c] - std::floor(s3[c]));
                        }
                        int32_t *o = ford.data() + f * natoms;
                        std::iota(o, o + natoms, 0);
                        kd_order(o, fs.data(), 0, natoms, len);
                    }
                });
            for (auto &t : pool) t.join();
            ANL_HIP(hipMalloc(&tr->frac_order, nframes * natoms * sizeof(int32_t)));
            up.put(tr->frac_order, ford.data(), nframes * natoms * sizeof(int32_t));
        }
        if (pos_cm) {
            ANL_HIP(hipMalloc(&tr->cm, nframes * ntypes * 3 * sizeof(double)));
            up.put(tr->cm, pos_cm, nframes * ntypes * 3 * sizeof(double));
        }
        up.finish();
        *out = owner.release();
    });
}

int anl_traj_destroy(anl_traj *tr) {
    return guard([&] {
        if (!tr) return;
        (void)hipFree(tr->pos);
        (void)hipFree(tr->box);
        (void)hipFree(tr->type_slot);
        (void)hipFree(tr->orig_slot);
        (void)hipFree(tr->frac_order);
        (void)hipFree(tr->cm);
        (void)hipFree(tr->frac16);
        (void)hipFree(tr->frac32);
        (void)hipFree(tr->jgroups);
        (void)hipFree(tr->pos_scr);
        delete tr;
    });
}


int anl_series_create(const double *data, uint64_t nrows, uint32_t ncols, anl_series **out) {
    return guard([&] {
        if (!data || !out) fail(ANL_ERR_ARG, "null pointer passed to anl_series_create");
        if (nrows == 0 || ncols == 0) fail(ANL_ERR_ARG, "empty time series");
        ensure_device();
        auto *s = new anl_series();
        s->nrows = nrows;
        s->ncols = ncols;
        ANL_HIP(hipMalloc(&s->data, nrows * ncols * sizeof(double)));
        ANL_HIP(hipMemcpy(s->data, data, nrows * ncols * sizeof(double), hipMemcpyHostToDevice));
        *out = s;
    });
}

int anl_series_destroy(anl_series *s) {
    return guard([&] {
        if (!s) return;
        (void)hipDeviceSynchronize();
        (void)hipFree(s->work);
        (void)hipFree(s->data);
        delete s;
    });
}

int anl_mediavar_update_async(double *mean_dev, double *var_dev, const double *x_dev, uint64_t len, uint32_t iblock,
                              void *stream) {
    return guard([&] {
        if (!mean_dev || !var_dev || !x_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_mediavar_update_async");
        if (len == 0) return;
        ensure_device();
        hipLaunchKernelGGL(mediavar_update_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, mean_dev, var_dev, x_dev, len, (double)(iblock + 1));
        ANL_HIP(hipGetLastError());
    });
}

int anl_mediavar_end_async(double *var_dev, uint64_t len, uint32_t n_b, void *stream) {
    return guard([&] {
        if (!var_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_mediavar_end_async");
        if (len == 0) return;
        ensure_device();
        double d = (double)((n_b - 1) * n_b);  // unsigned product, then to double (calcoliblocchi.h:60)
        hipLaunchKernelGGL(div_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, (hipStream_t)stream, var_dev,
                           len, d);
        ANL_HIP(hipGetLastError());
    });
}

int anl_counts_to_double_async(const uint64_t *counts_dev, double incr, uint64_t len, double *out_dev, void *stream) {
    return guard([&] {
        if (!counts_dev || !out_dev) fail(ANL_ERR_ARG, "null pointer passed to anl_counts_to_double_async");
        if (len == 0) return;
        ensure_device();
        hipLaunchKernelGGL(counts_to_double_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, (const unsigned long long *)counts_dev, incr, len, out_dev);
        ANL_HIP(hipGetLastError());
    });
}

}  // extern "C"
