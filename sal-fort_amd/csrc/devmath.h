// Device (and host-device) math shared by the kernels: cubed-sphere maps, barycentric Lagrange
// basis, and the SAL-gradient kernel in its reference and cancellation-free forms.
//
// The library is compiled with -ffp-contract=off: every fused multiply-add below is explicit.
#pragma once

#include <hip/hip_runtime.h>

namespace sal {

constexpr double KCONS = -2.343256857908601e-9;  // fast_sum_module.f90:14
constexpr double KP1 = 1.0 - 6.21196;             // fast_sum_module.f90:23
constexpr double KP2 = 2.7 + 6.0;                 // fast_sum_module.f90:24
constexpr int MAX_DEG = 16;

// xyz_from_xieta, cube_sphere_module.f90:64-95 (same expression order)
__host__ __device__ inline void xyz_from_xieta(double xi, double eta, int face, double& x, double& y, double& z) {
    double ax = tan(xi), ay = tan(eta);
    double r = sqrt(1.0 + ax * ax + ay * ay);
    switch (face) {
        case 1: x = 1.0 / r; y = ax * x; z = ay * x; break;
        case 2: y = 1.0 / r; x = -ax * y; z = ay * y; break;
        case 3: x = -1.0 / r; y = ax * x; z = -ay * x; break;
        case 4: y = -1.0 / r; x = -ax * y; z = -ay * y; break;
        case 5: z = 1.0 / r; x = -ay * z; y = ax * z; break;
        default: z = -1.0 / r; x = -ay * z; y = -ax * z; break;
    }
}

// face_from_xyz, cube_sphere_module.f90:7-35 (exact comparisons)
__host__ __device__ inline int face_from_xyz(double x, double y, double z) {
    double ax = fabs(x), ay = fabs(y), az = fabs(z);
    if (ax >= ay && ax >= az) return x >= 0 ? 1 : 3;
    if (ay >= ax && ay >= az) return y >= 0 ? 2 : 4;
    return z >= 0 ? 5 : 6;
}

// xieta_from_xyz, cube_sphere_module.f90:38-61
__host__ __device__ inline void xieta_from_xyz(double x, double y, double z, int face, double& xi, double& eta) {
    switch (face) {
        case 1: xi = atan(y / x); eta = atan(z / x); break;
        case 2: xi = atan(-x / y); eta = atan(z / y); break;
        case 3: xi = atan(y / x); eta = atan(-z / x); break;
        case 4: xi = atan(-x / y); eta = atan(-z / y); break;
        case 5: xi = atan(y / z); eta = atan(-x / z); break;
        default: xi = atan(-y / z); eta = atan(-x / z); break;
    }
}

// cos(pi i / degree): node i of the degree-d grid on [-1, 1]
__host__ __device__ inline double bli_cheb(int degree, int i) {
    const double pi = 4.0 * 0.78539816339744830962;  // 4*atan(1), exact in double
    return cos(pi / (double)degree * (double)i);
}

// node i of [a, b] from its [-1, 1] value cs = bli_cheb(degree, i) (degree >= 1)
__host__ __device__ inline double bli_node_cs(double a, double b, double cs) {
    double range = 0.5 * (b - a), shift = 0.5 * (b + a);
    return cs * range + shift;
}

// Chebyshev-2 nodes on [a, b], descending (bli_interp_points_shift, bli_module.f90:28-48)
__host__ __device__ inline double bli_node(double a, double b, int degree, int i) {
    if (degree == 0) return 0.5 * (b + a);
    return bli_node_cs(a, b, bli_cheb(degree, i));
}

// barycentric weights (bli_coeffs, bli_module.f90:6-26)
__host__ __device__ inline double bli_coeff(int degree, int i) {
    if (degree == 0) return 1.0;
    double s = (i % 2 == 0) ? 1.0 : -1.0;
    return (i == 0 || i == degree) ? 0.5 * s : s;
}

// One dimension of interp_vals_bli (bli_module.f90:71-105): f[i] = l_i(x) for nodes[0..degree].
__host__ __device__ inline void bli_basis_1d(double x, const double* nodes, int degree, double* f) {
    bool found = false;
    for (int i = 0; i <= degree; ++i) {
        f[i] = 0.0;
        if (fabs(x - nodes[i]) < 1.0e-16) found = true;
    }
    if (found) {
        for (int i = 0; i <= degree; ++i) f[i] = (fabs(x - nodes[i]) < 1.0e-16) ? 1.0 : 0.0;
        return;
    }
    double den = 0.0;
    for (int i = 0; i <= degree; ++i) {
        double v = bli_coeff(degree, i) / (x - nodes[i]);
        f[i] = v;
        den = den + v;
    }
    double inv = 1.0 / den;
    for (int i = 0; i <= degree; ++i) f[i] = f[i] * inv;
}

// sal_grad_gfunc (fast_sum_module.f90:9-31), exactly as written: used where bit-identical values
// matter (the kernel test entry point).
__host__ __device__ inline void gfunc_ref_exact(double tx, double ty, double tz, double sx, double sy, double sz,
                                                double& gx, double& gy) {
    gx = 0.0;
    gy = 0.0;
    if (fabs(tz - 1.0) > 1e-15 && fabs(tz + 1.0) > 1e-15) {
        double g = tx * sx + ty * sy + tz * sz;
        g = g < 1.0 ? g : 1.0;
        g = g > -1.0 ? g : -1.0;
        double mp = 2.0 - 2.0 * g;
        double sqp = sqrt(mp);
        double p1 = KP1 / (sqp * mp + 1e-16);
        double p2 = KP2 * (2.0 * g + sqp) / (2.0 * (g * g - 1.0) + 1e-16);
        double val = (p1 + p2) * KCONS;
        double x32 = tz * tz;
        double mp2 = sqrt(1.0 - x32);
        gy = (sz * (1.0 - x32) - tz * (tx * sx + ty * sy)) / mp2 * val;
        gx = (tx * sy - ty * sx) / mp2 * val;
    }
}

// Per-target constants of the kernel: pole flag, 1/sqrt(1 - tz^2), and that times the kernel's constant.
struct Tgt {
    double x, y, z, a /* 1 - tz^2 */, inv_c /* 1/sqrt(a), 0 at the poles */, kc /* KCONS inv_c */;
};
__host__ __device__ inline Tgt make_tgt(double x, double y, double z) {
    Tgt t;
    t.x = x; t.y = y; t.z = z;
    double x32 = z * z;
    t.a = 1.0 - x32;
    bool pole = !(fabs(z - 1.0) > 1e-15 && fabs(z + 1.0) > 1e-15);
    t.inv_c = pole ? 0.0 : 1.0 / sqrt(t.a);
    t.kc = KCONS * t.inv_c;
    return t;
}

// 1/x for a finite, normal, non-zero x: the hardware reciprocal estimate r (relative error below
// 2^-24 on gfx950, measured over [2^-40, 4]) refined by one Newton step r <- r + r (1 - x r), which
// squares that error: within about 2^-48 of 1/x, at a fraction of the cost of an IEEE division.  The
// kernel's denominators are products of terms bounded away from zero by its 1e-16 regularisers, so
// they always qualify.
__device__ inline double rcp_nr(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// sqrt(x) for a normal x > 0: the hardware reciprocal square root y (relative error below 2^-24)
// and one Goldschmidt step g <- g + g (1/2 - g y / 2) from g = x y, which leaves 3/2 of the squared
// error: within about 2^-47 of sqrt(x).
__device__ inline double sqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y;
    return fma(g, fma(-g, 0.5 * y, 0.5), g);
}

// Reference expression of the kernel with the products that cancel (dot product, g*g - 1, the
// north and east numerators) rounded as the reference rounds them, and the two divisions p1, p2
// merged into one and KCONS / sqrt(1 - tz^2) hoisted per target: each value agrees with
// sal_grad_gfunc to about 2^-46 relative (the one-step root and reciprocal above), the self pair's
// rounding noise included.  Steps that involve 2 g
// are fused exactly (2 g is exact, so fma(-2, g, 2) is the rounded 2 - 2 g, fma(2, x, 1e-16) the
// rounded 2 x + 1e-16, fma(2, g, c) the rounded 2 g + c); the products that do not cancel (d1, num)
// are fused, a rounding of at most an ulp.  The chord's root is taken of mp + 1e-300: mp is 0 or at
// least 2^-52, so that sum is mp itself for every mp > 0, and for mp = 0 the root 1e-150 only meets
// terms of order 1 and 1e-16, which it leaves unchanged (the self pair gets the same bits as with
// a root of 0).  gfunc_ref_u leaves the per-target factor KCONS / sqrt(1 - tz^2) out: it returns
// the scalar factor and the two projections, whose products summed over sources and then scaled
// once by t.kc give the target's sums.
__device__ inline double gfunc_ref_u(const Tgt& t, double sx, double sy, double sz, double& pe, double& pn) {
    const double dxy = t.x * sx + t.y * sy;
    double g = dxy + t.z * sz;
    g = g < 1.0 ? g : 1.0;
    g = g > -1.0 ? g : -1.0;
    const double mp = fma(-2.0, g, 2.0);
    const double sqp = sqrt_pos(mp + 1e-300);
    const double d1 = fma(sqp, mp, 1e-16);
    const double d2 = fma(2.0, g * g - 1.0, 1e-16);
    const double num = fma(KP1, d2, KP2 * fma(2.0, g, sqp) * d1);
    pn = sz * t.a - t.z * dxy;
    pe = t.x * sy - t.y * sx;
    return num * rcp_nr(d1 * d2);
}
// gfunc_ref_u for a well-separated pair (the source is a cluster's proxy or the target a cluster's
// proxy: PC, CP and CC entries, radius / dist < theta), where g = t.s stays below 1 by far: the
// upper clamp and the root's 1e-300 guard are no-ops there and are left out (the same bits); the
// lower clamp stays (antipodal pairs).  The two projections are fused: their terms are bounded away
// from zero at these distances, so the rounding moves each by about an ulp of the term (below 1e-15
// relative at the closest cluster distances of a leaf of a few dozen points on a million-point grid).
// The constant KP1 is taken out of the numerator (KP1 d2 + KP2 c d1 = KP1 (d2 + (KP2 / KP1) c d1)):
// the value returned is the factor over KP1, and callers put KP1 into the source weights or the
// final scaling.  2 (g^2 - 1) is formed as -(2 - 2g)(1 + g) from the 2 - 2g the root needs anyway:
// one operation fewer, and no rounding of g^2 where g is close to 1.  g = t.s itself must be rounded
// as the reference rounds it, (t_x s_x + t_y s_y) + t_z s_z: 1 - g of a close pair (a point and a
// small cluster at the acceptance distance) can be 1e-5, where another rounding of g moves the
// kernel value by 1e-11.
constexpr double KP21 = KP2 / KP1;
// the factor over KP1 from g = t.s alone
__device__ inline double far_val(double g) {
    g = g > -1.0 ? g : -1.0;
    const double mp = fma(-2.0, g, 2.0);
    const double sqp = sqrt_pos(mp);
    const double d1 = fma(sqp, mp, 1e-16);
    const double d2 = fma(-mp, g + 1.0, 1e-16);  // 2 (g^2 - 1) = -(2 - 2g)(1 + g)
    const double num = fma(KP21, fma(2.0, g, sqp) * d1, d2);
    return num * rcp_nr(d1 * d2);
}
__device__ inline double gfunc_ref_far_u(const Tgt& t, double sx, double sy, double sz, double& pe, double& pn) {
    const double dxy = t.x * sx + t.y * sy;
    pn = fma(sz, t.a, -(t.z * dxy));
    pe = fma(t.x, sy, -(t.y * sx));
    return far_val(dxy + t.z * sz);  // t.s rounded as the reference rounds it
}
__device__ inline void gfunc_ref(const Tgt& t, double sx, double sy, double sz, double& gx, double& gy) {
    double pe, pn;
    const double f = gfunc_ref_u(t, sx, sy, sz, pe, pn) * t.kc;
    gy = pn * f;
    gx = pe * f;
}

// Scalar part val(c) / KCONS of the kernel from the chord squared c2 = |s - t|^2 (cancellation-free
// form: 2g + c = 2 - c^2 + c, 2(g^2 - 1) = -c^2 (4 - c^2)/2), regularisers kept; c2 = 0 gives the
// regularised value of the self pair, whose projections are 0.
__device__ inline double val_stable_nc(double c2) {
    const double c = sqrt_pos(c2 + 1e-300);
    const double d1 = fma(c, c2, 1e-16);
    const double d2 = fma(-0.5 * c2, 4.0 - c2, 1e-16);
    const double num = fma(KP1, d2, KP2 * ((2.0 - c2) + c) * d1);
    return num * rcp_nr(d1 * d2);
}
__device__ inline double val_stable(double c2) { return val_stable_nc(c2) * KCONS; }

// Cancellation-free kernel: (east, north) for target t and source s, from d = s - t; gfunc_stable_u
// leaves out the per-target factor t.kc as gfunc_ref_u does.
__device__ inline double gfunc_stable_u(const Tgt& t, double sx, double sy, double sz, double& pe, double& pn) {
    const double dx = sx - t.x, dy = sy - t.y, dz = sz - t.z;
    const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
    pn = dz * t.a - t.z * fma(t.x, dx, t.y * dy);
    pe = t.x * dy - t.y * dx;
    return val_stable_nc(c2);
}
__device__ inline void gfunc_stable(const Tgt& t, double sx, double sy, double sz, double& gx, double& gy) {
    double pe, pn;
    const double f = gfunc_stable_u(t, sx, sy, sz, pe, pn) * t.kc;
    gy = pn * f;
    gx = pe * f;
}

__host__ __device__ inline void project_en(const Tgt& t, double vx, double vy, double vz, double& ge, double& gn) {
    ge = (t.x * vy - t.y * vx) * t.inv_c;
    gn = (vz * t.a - t.z * (t.x * vx + t.y * vy)) * t.inv_c;
}

// Workgroups go to the 8 XCDs round robin by id (XCD = id % 8).  With nx > 0 items, chunks of
// XCD_CHUNK consecutive items are dealt to the XCDs in turn: XCD x runs chunks x, x + 8, x + 16, ...
// in order, so the items running at once on one XCD are few chunks of neighbouring items (leaves,
// targets in spatial order), which read mostly the same sources from its L2, while every XCD's
// share spreads over the whole launch and all finish together.  (Contiguous per-XCD ranges balanced
// by stored bytes shared rows as well but left XCDs finishing apart: 9.24 vs 8.88 ms for the mixed
// near field at leaves <= 64.)  The grid is nx rounded up to whole rounds of 8 chunks; ids past nx
// return -1.  nx <= 0: one item per id.
constexpr int XCD_CHUNK = 16;
__device__ __forceinline__ int xcd_item(int nx) {
    if (nx <= 0) return blockIdx.x;
    const int x = blockIdx.x & 7, j = blockIdx.x >> 3;
    const int item = ((j / XCD_CHUNK) * 8 + x) * XCD_CHUNK + j % XCD_CHUNK;
    return item < nx ? item : -1;
}
static inline unsigned xcd_grid(int nx) {
    return (unsigned)((nx + 8 * XCD_CHUNK - 1) / (8 * XCD_CHUNK) * (8 * XCD_CHUNK));
}

}  // namespace sal
