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

// Per-target constants of the kernel: pole flag and 1/sqrt(1 - tz^2).
struct Tgt {
    double x, y, z, a /* 1 - tz^2 */, inv_c /* 1/sqrt(a), 0 at the poles */;
};
__host__ __device__ inline Tgt make_tgt(double x, double y, double z) {
    Tgt t;
    t.x = x; t.y = y; t.z = z;
    double x32 = z * z;
    t.a = 1.0 - x32;
    bool pole = !(fabs(z - 1.0) > 1e-15 && fabs(z + 1.0) > 1e-15);
    t.inv_c = pole ? 0.0 : 1.0 / sqrt(t.a);
    return t;
}

// 1/x for a finite, normal, non-zero x: the hardware reciprocal estimate refined by two Newton steps
// r <- r + r (1 - x r) (each squares the relative error), within an ulp or so of the rounded
// quotient; a fraction of the cost of an IEEE division.  The kernel's denominators are products of
// terms bounded away from zero by its 1e-16 regularisers, so they always qualify.
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// sqrt(x) for x = 0 or a normal x <= 4 (the kernel's chord squared 2 - 2g or |s - t|^2): the
// hardware reciprocal square root y, then one Goldschmidt step on (g, h) = (x y, y / 2) and one
// Newton correction g <- g + (x - g^2) h, each squaring the relative error; within an ulp or so of
// the rounded root.  The library sqrt adds a second correction and the scaling and special-value
// handling this range does not need.  x = 0 gives 0.
__device__ inline double sqrt_chord(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, x), h, g);
    return x > 0.0 ? g : 0.0;
}

// Reference expression of the kernel with the products that cancel (dot product, g*g - 1, the
// north numerator) evaluated in the reference's order, and the two divisions p1, p2 merged into
// one and 1/sqrt(1 - tz^2) hoisted per target: each value agrees with sal_grad_gfunc to a few ulp,
// the self pair's rounding noise included.
__device__ inline void gfunc_ref(const Tgt& t, double sx, double sy, double sz, double& gx, double& gy) {
    double dxy = t.x * sx + t.y * sy;
    double g = dxy + t.z * sz;
    g = g < 1.0 ? g : 1.0;
    g = g > -1.0 ? g : -1.0;
    double mp = 2.0 - 2.0 * g;
    double sqp = sqrt_chord(mp);
    double d1 = sqp * mp + 1e-16;
    double d2 = 2.0 * (g * g - 1.0) + 1e-16;
    double num = KP1 * d2 + KP2 * (2.0 * g + sqp) * d1;
    double f = (num * rcp_nr(d1 * d2)) * KCONS * t.inv_c;
    gy = (sz * t.a - t.z * dxy) * f;
    gx = (t.x * sy - t.y * sx) * f;
}

// Scalar part val(c) of the kernel from the chord squared c2 = |s - t|^2 (cancellation-free form:
// 2g + c = 2 - c^2 + c, 2(g^2 - 1) = -c^2 (4 - c^2)/2), regularisers kept.
__device__ inline double val_stable(double c2) {
    double c = sqrt_chord(c2);
    double d1 = c * c2 + 1e-16;
    double d2 = -0.5 * c2 * (4.0 - c2) + 1e-16;
    double num = KP1 * d2 + KP2 * ((2.0 - c2) + c) * d1;
    return (num * rcp_nr(d1 * d2)) * KCONS;
}

// Cancellation-free kernel: (east, north) for target t and source s, from d = s - t.
__device__ inline void gfunc_stable(const Tgt& t, double sx, double sy, double sz, double& gx, double& gy) {
    double dx = sx - t.x, dy = sy - t.y, dz = sz - t.z;
    double c2 = dx * dx + dy * dy + dz * dz;
    double f = val_stable(c2) * t.inv_c;
    gy = (dz * t.a - t.z * (t.x * dx + t.y * dy)) * f;
    gx = (t.x * dy - t.y * dx) * f;
}

// project a Cartesian tangential vector onto the target's east/north basis
__host__ __device__ inline void project_en(const Tgt& t, double vx, double vy, double vz, double& ge, double& gn) {
    ge = (t.x * vy - t.y * vx) * t.inv_c;
    gn = (vz * t.a - t.z * (t.x * vx + t.y * vy)) * t.inv_c;
}

}  // namespace sal
