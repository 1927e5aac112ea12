// FP64 minimum image of the reference (basetrajectory.h:200-268), shared by
// the pair kernels.  Include inside namespace anl, in a translation unit
// compiled with "#pragma clang fp contract(off)".
#pragma once

struct DevCell {
    double lh0, lh1, lh2, L0, L1, L2, xy, xz, yz;
};

// Internal-format box [xlo, ylo, zlo, lx/2, ly/2, lz/2, xy, xz, yz] of a frame
// (b: a pointer to double in any address space).
template <class P>
__device__ __forceinline__ DevCell load_cell(P b) {
    DevCell c;
    c.lh0 = b[3];
    c.lh1 = b[4];
    c.lh2 = b[5];
    c.L0 = b[3] * 2;
    c.L1 = b[4] * 2;
    c.L2 = b[5] * 2;
    c.xy = b[6];
    c.xz = b[7];
    c.yz = b[8];
    return c;
}

// basetrajectory.h:224-268 in the reference's operation order: z, then y,
// then x; a z step also shifts y by yz and x by xz, a y step shifts x by xy.
// A step "if (d<0) d += 2l else d -= 2l" is written d -= s*2l with s = +-1
// (exact products; x - (-y) == x + y in IEEE arithmetic).
// The first iteration of each loop is written as straight-line code whose
// result each lane keeps only if it needs the step (the loop condition),
// so a wavefront whose lanes disagree does not branch; the loop then runs
// the further steps, which a component within 1.5 cell lengths never needs.
// Every lane performs the reference's operations in the reference's order
// either way.
template <bool TRI>
__device__ __forceinline__ void min_image_ref(double &x0, double &x1, double &x2, const DevCell &c) {
    {
        const bool t = fabs(x2) > c.lh2;
        const double s = x2 < 0.0 ? -1.0 : 1.0;
        const double n2 = x2 - s * c.L2;
        x2 = t ? n2 : x2;
        if (TRI) {
            const double n1 = x1 - s * c.yz, n0 = x0 - s * c.xz;
            x1 = t ? n1 : x1;
            x0 = t ? n0 : x0;
        }
    }
    // Inputs are validated on the host (finite, within 10^4 cell lengths), so
    // each loop ends after a few steps; the counters only guard the device.
    int guard = 0;
    while (fabs(x2) > c.lh2 && ++guard < (1 << 20)) {
        double s = x2 < 0.0 ? -1.0 : 1.0;
        x2 = x2 - s * c.L2;
        if (TRI) {
            x1 = x1 - s * c.yz;
            x0 = x0 - s * c.xz;
        }
    }
    {
        const bool t = fabs(x1) > c.lh1;
        const double s = x1 < 0.0 ? -1.0 : 1.0;
        const double n1 = x1 - s * c.L1;
        x1 = t ? n1 : x1;
        if (TRI) {
            const double n0 = x0 - s * c.xy;
            x0 = t ? n0 : x0;
        }
    }
    while (fabs(x1) > c.lh1 && ++guard < (2 << 20)) {
        double s = x1 < 0.0 ? -1.0 : 1.0;
        x1 = x1 - s * c.L1;
        if (TRI) x0 = x0 - s * c.xy;
    }
    {
        const bool t = fabs(x0) > c.lh0;
        const double s = x0 < 0.0 ? -1.0 : 1.0;
        const double n0 = x0 - s * c.L0;
        x0 = t ? n0 : x0;
    }
    while (fabs(x0) > c.lh0 && ++guard < (3 << 20)) {
        double s = x0 < 0.0 ? -1.0 : 1.0;
        x0 = x0 - s * c.L0;
    }
}

// basetrajectory.h:200-219: x = xi - xj, minimum image, d2 = 0 + x0^2 + x1^2
// + x2^2 with separate products and sums.
__device__ __forceinline__ double d2_of(double x0, double x1, double x2) {
    double d2 = __dmul_rn(x0, x0);
    d2 = __dadd_rn(d2, __dmul_rn(x1, x1));
    d2 = __dadd_rn(d2, __dmul_rn(x2, x2));
    return d2;
}

template <bool TRI>
__device__ __forceinline__ double d2_ref(double xi0, double xi1, double xi2, double xj0, double xj1, double xj2,
                                         const DevCell &c) {
    double x0 = xi0 - xj0;
    double x1 = xi1 - xj1;
    double x2 = xi2 - xj2;
    min_image_ref<TRI>(x0, x1, x2, c);
    return d2_of(x0, x1, x2);
}
