// One interaction at a time, as the reference's per-entry routines compute it (FP64, the reference's
// kernel expression and layouts):
//
//   sal_interact          pp_interact / pc_interact / cp_interact / cc_interact of the MPAS evaluator
//                         (fast_sum_module.f90:33-288): one DTT entry between two panels of the tree
//   sal_mom6_interact     pc_interact / pp_interact of the MOM6 evaluator (mom6_fast_sum_module.f90:
//                         76-150): one target point against one source panel
//   sal_mom6_pp_points    pp_interact with the sources given as points (the rank's own points and
//                         halo copies a caller at several ranks holds)
//
// Each call costs what the entry itself touches, as in the reference: the source panel's points (or
// proxies) and the target panel's points (or proxies), never the whole grid.  Source fields go to
// the device as the source's own values (gathered on the host through the handle's copy of the tree
// permutation), results come back for the target's points only, and the device scratch is the
// handle's, grown when an entry needs more.  A handle is not re-entrant (one call at a time).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// P2M of one source panel (pc_interact :99-111, cc_interact :229-241): W[k], k = a + b nd (a the xi
// node, b the eta node), = sum over the panel's points s of l_a(xi_s) l_b(eta_s) q_s; the points are
// the panel's DFS range [s0, s0 + ns), q_src their area * ssh in that order.  One block of 256.
__global__ __launch_bounds__(256) void k_one_weights(int nd, int64_t s0, int ns, const double* __restrict__ bounds,
                                                     const int* __restrict__ perm, const double* __restrict__ xi,
                                                     const double* __restrict__ eta,
                                                     const double* __restrict__ q_src, double* __restrict__ W) {
    __shared__ double fb[2][128][MAX_DEG + 1], fq[128], nodes[2][MAX_DEG + 1];
    const int m = nd * nd, deg = nd - 1, tid = threadIdx.x;
    if (tid < 2 * nd) nodes[tid / nd][tid % nd] = bli_node(bounds[3 * (tid / nd)], bounds[3 * (tid / nd) + 2], deg, tid % nd);
    __syncthreads();
    double acc[2] = {0.0, 0.0};  // weights tid, tid + 256 (m <= 289)
    for (int c0 = 0; c0 < ns; c0 += 128) {
        const int nc = min(128, ns - c0), d = tid >> 7, t = tid & 127;
        if (t < nc) {
            const int p = perm[s0 + c0 + t];
            bli_basis_1d(d ? eta[p] : xi[p], nodes[d], deg, fb[d][t]);
            if (d == 0) fq[t] = q_src[c0 + t];
        }
        __syncthreads();
        for (int r = 0; r < 2; ++r) {
            const int w = tid + 256 * r;
            if (w < m) {
                const int a = w % nd, b = w / nd;
                double v = acc[r];
                for (int k = 0; k < nc; ++k) v = fma(fb[0][k][a] * fb[1][k][b], fq[k], v);
                acc[r] = v;
            }
        }
        __syncthreads();
    }
    for (int r = 0; r < 2; ++r)
        if (tid + 256 * r < m) W[tid + 256 * r] = acc[r];
}

// Proxy point k = a + b nd of a panel (its Chebyshev grid mapped to the sphere, xyz_from_xieta)
__device__ inline void proxy_point(const double* b, int face, int nd, int k, double& x, double& y, double& z) {
    const int deg = nd - 1;
    xyz_from_xieta(bli_node(b[0], b[2], deg, k % nd), bli_node(b[3], b[5], deg, k / nd), face, x, y, z);
}

// One DTT entry of kind KIND (0 PP, 1 PC, 2 CP, 3 CC) between target panel T and source panel S, in
// the reference's semantics (kernel expression with its regularisers, self pair included, projected
// east/north components interpolated at target proxies, fast_sum_module.f90:133-288).  Sources:
// S's points [s0, s0 + ns) with weights q_src (PP, CP) or S's proxies with weights W (PC, CC).
// Targets: T's points [t0, t0 + nt) (PP, PC), or T's proxies whose potentials are then interpolated
// at T's points (CP, CC).  out[2 i], out[2 i + 1]: (east, north) of target point t0 + i.  One block.
template <int KIND>
__global__ __launch_bounds__(256) void k_one_eval(int nd, int64_t t0, int nt, int64_t s0, int ns,
                                                  const double* __restrict__ bt, int ft, const double* __restrict__ bs,
                                                  int fs, const int* __restrict__ perm, const double* __restrict__ px,
                                                  const double* __restrict__ py, const double* __restrict__ pz,
                                                  const double* __restrict__ xi, const double* __restrict__ eta,
                                                  const double* __restrict__ q_src, const double* __restrict__ W,
                                                  double* __restrict__ out) {
    constexpr bool SRC_PROXY = (KIND == 1 || KIND == 3), TGT_PROXY = (KIND >= 2);
    constexpr int MMAX = (MAX_DEG + 1) * (MAX_DEG + 1);
    __shared__ double sx[SRC_PROXY ? MMAX : 1], sy[SRC_PROXY ? MMAX : 1], sz[SRC_PROXY ? MMAX : 1];
    __shared__ double phi[TGT_PROXY ? 2 : 1][TGT_PROXY ? MMAX : 1], nodes[2][MAX_DEG + 1];
    const int m = nd * nd, deg = nd - 1, tid = threadIdx.x;
    const int nsrc = SRC_PROXY ? m : ns;
    if (SRC_PROXY)
        for (int k = tid; k < m; k += 256) proxy_point(bs, fs, nd, k, sx[k], sy[k], sz[k]);
    __syncthreads();
    auto source = [&](int k, double& x, double& y, double& z, double& w) {
        if (SRC_PROXY) { x = sx[k]; y = sy[k]; z = sz[k]; w = W[k]; }
        else { x = px[s0 + k]; y = py[s0 + k]; z = pz[s0 + k]; w = q_src[k]; }
    };
    if (!TGT_PROXY) {
        for (int i = tid; i < nt; i += 256) {
            const Tgt T = make_tgt(px[t0 + i], py[t0 + i], pz[t0 + i]);
            double ax = 0.0, ay = 0.0;
            for (int k = 0; k < nsrc; ++k) {
                double x, y, z, w, ge, gn;
                source(k, x, y, z, w);
                gfunc_ref(T, x, y, z, ge, gn);
                ax = fma(ge, w, ax);
                ay = fma(gn, w, ay);
            }
            out[2 * i] = ax;
            out[2 * i + 1] = ay;
        }
        return;
    }
    // target proxies: potentials (east, north) at T's grid, then L2P at T's points
    for (int j = tid; j < m; j += 256) {
        double zx, zy, zz;
        proxy_point(bt, ft, nd, j, zx, zy, zz);
        const Tgt Z = make_tgt(zx, zy, zz);
        double ax = 0.0, ay = 0.0;
        for (int k = 0; k < nsrc; ++k) {
            double x, y, z, w, ge, gn;
            source(k, x, y, z, w);
            gfunc_ref(Z, x, y, z, ge, gn);
            ax = fma(ge, w, ax);
            ay = fma(gn, w, ay);
        }
        phi[0][j] = ax;
        phi[1][j] = ay;
    }
    if (tid < 2 * nd) nodes[tid / nd][tid % nd] = bli_node(bt[3 * (tid / nd)], bt[3 * (tid / nd) + 2], deg, tid % nd);
    __syncthreads();
    for (int i = tid; i < nt; i += 256) {
        const int p = perm[t0 + i];
        double fx[MAX_DEG + 1], fy[MAX_DEG + 1];
        bli_basis_1d(xi[p], nodes[0], deg, fx);
        bli_basis_1d(eta[p], nodes[1], deg, fy);
        double ax = 0.0, ay = 0.0;
        for (int b = 0; b < nd; ++b)
            for (int a = 0; a < nd; ++a) {
                const double bv = fx[a] * fy[b];
                ax = fma(bv, phi[0][a + b * nd], ax);
                ay = fma(bv, phi[1][a + b * nd], ay);
            }
        out[2 * i] = ax;
        out[2 * i + 1] = ay;
    }
}

// One MOM6 target point (x, y, z) against one source (mom6_fast_sum_module.f90:76-150), one thread,
// in the reference's loop order: PROXY = true, source panel S's proxies (node a in xi inner, b in eta
// outer) with weights w (read at b nd + a when transposed, the reference's eta-major read of the
// xi-major proxy_source_compute layout, else at a nd + b); PROXY = false, ns points (sx, sy, sz) with
// q = ssh * area.  out[0..1] = (east, north).
template <bool PROXY>
__global__ void k_one_point(double x, double y, double z, int nd, const double* __restrict__ bs, int fs, int transposed,
                            int ns, const double* __restrict__ sx, const double* __restrict__ sy,
                            const double* __restrict__ sz, const double* __restrict__ w, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const Tgt T = make_tgt(x, y, z);
    double ax = 0.0, ay = 0.0;
    if (PROXY) {
        for (int b = 0; b < nd; ++b)
            for (int a = 0; a < nd; ++a) {
                double cx, cy, cz, ge, gn;
                proxy_point(bs, fs, nd, a + b * nd, cx, cy, cz);
                gfunc_ref(T, cx, cy, cz, ge, gn);
                const double wv = transposed ? w[b * nd + a] : w[a * nd + b];
                ax = fma(ge, wv, ax);
                ay = fma(gn, wv, ay);
            }
    } else {
        for (int k = 0; k < ns; ++k) {
            double ge, gn;
            gfunc_ref(T, sx[k], sy[k], sz[k], ge, gn);
            ax = fma(w[k], ge, ax);
            ay = fma(w[k], gn, ay);
        }
    }
    out[0] = ax;
    out[1] = ay;
}

// The handle's host copy of the tree permutation (DFS position -> caller index), made once.
static const std::vector<int>& host_perm(sal_geom g) {
    std::lock_guard<std::mutex> lk(g->mu);
    if ((int64_t)g->h_perm.size() != g->n) {
        std::vector<int> p(g->n);
        d2h(p.data(), g->perm.p, g->n, g->stream);
        SAL_HIP(hipStreamSynchronize(g->stream));
        g->h_perm.swap(p);
    }
    return g->h_perm;
}

// the handle's scratch for one interaction: at least `count` doubles (grown, never shrunk)
static double* one_scratch(sal_geom g, size_t count) {
    if (g->one_buf.n < count) g->one_buf.alloc(std::max(count, (size_t)4096));
    return g->one_buf.p;
}

}  // namespace sal

using namespace sal;

extern "C" {

int sal_interact(sal_geom g, int kind, int target, int source, int degree, const double* area, const double* ssh,
                 double* gx, double* gy, int accumulate) {
    return guard([&] {
        SAL_REQUIRE(g && area && ssh && gx && gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(kind >= 0 && kind <= 3 && target >= 0 && target < g->P && source >= 0 && source < g->P,
                    SAL_ERR_ARG, "bad interaction");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        hipStream_t s = g->stream;
        const std::vector<int>& perm = host_perm(g);
        const int nd = degree + 1, m = nd * nd;
        const int64_t t0 = g->start[target], s0 = g->start[source];
        const int nt = g->count[target], ns = g->count[source];
        if (!accumulate) {
            std::memset(gx, 0, sizeof(double) * g->n);
            std::memset(gy, 0, sizeof(double) * g->n);
        }
        if (nt == 0 || ns == 0) return;
        // device scratch: [source q: ns][weights: m][outputs: 2 nt]
        double* q = one_scratch(g, (size_t)ns + m + 2 * (size_t)nt);
        double* W = q + ns;
        double* out = W + m;
        std::vector<double> hq(ns), ho(2 * (size_t)nt);
        for (int k = 0; k < ns; ++k) {
            const int p = perm[s0 + k];
            hq[k] = area[p] * ssh[p];
        }
        h2d(q, hq.data(), ns, s);
        const double* bt = g->d_bounds.p + 6 * (int64_t)target;
        const double* bs = g->d_bounds.p + 6 * (int64_t)source;
        if (kind == 1 || kind == 3)
            k_one_weights<<<1, 256, 0, s>>>(nd, s0, ns, bs, g->perm.p, g->xi.p, g->eta.p, q, W);
#define SAL_ONE(K)                                                                                               \
    k_one_eval<K><<<1, 256, 0, s>>>(nd, t0, nt, s0, ns, bt, g->face[target], bs, g->face[source], g->perm.p,  \
                                    g->px.p, g->py.p, g->pz.p, g->xi.p, g->eta.p, q, W, out)
        switch (kind) {
            case 0: SAL_ONE(0); break;
            case 1: SAL_ONE(1); break;
            case 2: SAL_ONE(2); break;
            default: SAL_ONE(3); break;
        }
#undef SAL_ONE
        SAL_HIP(hipGetLastError());
        d2h(ho.data(), out, 2 * (size_t)nt, s);
        SAL_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nt; ++i) {
            const int p = perm[t0 + i];
            gx[p] += ho[2 * (size_t)i];
            gy[p] += ho[2 * (size_t)i + 1];
        }
    });
}

int sal_mom6_interact(sal_geom g, int kind, double x, double y, double z, int source, int degree, int transposed,
                      const double* w, const double* area, const double* ssh, double* gx, double* gy) {
    return guard([&] {
        SAL_REQUIRE(g && gx && gy && (kind == 0 || kind == 1), SAL_ERR_ARG, "bad arguments");
        SAL_REQUIRE(source >= 0 && source < g->P, SAL_ERR_ARG, "source panel out of range");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        hipStream_t s = g->stream;
        const int nd = degree + 1, m = nd * nd;
        const double* bs = g->d_bounds.p + 6 * (int64_t)source;
        double ho[2] = {0.0, 0.0};
        if (kind == 1) {
            SAL_REQUIRE(w, SAL_ERR_ARG, "null weights");
            double* dw = one_scratch(g, (size_t)m + 2);
            h2d(dw, w, m, s);
            k_one_point<true><<<1, 64, 0, s>>>(x, y, z, nd, bs, g->face[source], transposed, 0, nullptr, nullptr,
                                               nullptr, dw, dw + m);
            SAL_HIP(hipGetLastError());
            d2h(ho, dw + m, 2, s);
        } else {
            SAL_REQUIRE(area && ssh, SAL_ERR_ARG, "null area / ssh");
            const std::vector<int>& perm = host_perm(g);
            const int64_t s0 = g->start[source];
            const int ns = g->count[source];
            if (ns > 0) {
                double* dq = one_scratch(g, (size_t)ns + 2);
                std::vector<double> hq(ns);
                for (int k = 0; k < ns; ++k) {
                    const int p = perm[s0 + k];
                    hq[k] = ssh[p] * area[p];  // the reference's ssh * area (mom6_fast_sum_module.f90:146-147)
                }
                h2d(dq, hq.data(), ns, s);
                k_one_point<false><<<1, 64, 0, s>>>(x, y, z, 0, nullptr, 0, 0, ns, g->px.p + s0, g->py.p + s0,
                                                    g->pz.p + s0, dq, dq + ns);
                SAL_HIP(hipGetLastError());
                d2h(ho, dq + ns, 2, s);
            }
        }
        SAL_HIP(hipStreamSynchronize(s));
        *gx += ho[0];
        *gy += ho[1];
    });
}

int sal_mom6_pp_points(sal_geom g, double x, double y, double z, int ns, const double* sx, const double* sy,
                       const double* sz, const double* ssh, const double* area, double* gx, double* gy) {
    return guard([&] {
        SAL_REQUIRE(g && gx && gy && ns >= 0, SAL_ERR_ARG, "bad arguments");
        if (ns == 0) return;
        SAL_REQUIRE(sx && sy && sz && ssh && area, SAL_ERR_ARG, "null pointer");
        hipStream_t s = g->stream;
        double* d = one_scratch(g, 4 * (size_t)ns + 2);
        std::vector<double> hq(ns);
        for (int k = 0; k < ns; ++k) hq[k] = ssh[k] * area[k];
        h2d(d, sx, ns, s);
        h2d(d + ns, sy, ns, s);
        h2d(d + 2 * (size_t)ns, sz, ns, s);
        h2d(d + 3 * (size_t)ns, hq.data(), ns, s);
        k_one_point<false><<<1, 64, 0, s>>>(x, y, z, 0, nullptr, 0, 0, ns, d, d + ns, d + 2 * (size_t)ns,
                                            d + 3 * (size_t)ns, d + 4 * (size_t)ns);
        SAL_HIP(hipGetLastError());
        double ho[2];
        d2h(ho, d + 4 * (size_t)ns, 2, s);
        SAL_HIP(hipStreamSynchronize(s));
        *gx += ho[0];
        *gy += ho[1];
    });
}

}  // extern "C"
