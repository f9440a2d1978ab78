// MOM6-variant evaluator on the GPU.
//
// Single tree traversal (replaces SINGLE_TREE_TRAVERSAL, tree_module.f90:401-487): every target
// point walks the tree first-in first-out from the six roots; a panel is a cluster (PC) when
// radius / dist < theta, a leaf that is not is a direct (PP) source, otherwise its four children
// are queued.  One thread per target, two passes: count, exclusive scan, fill, so the lists come
// out target-major in the reference's order.  MAC decisions within a few ulp of theta mark the
// target for a host walk with the host's acos.
//
// proxy_source_compute (mom6_fast_sum_module.f90:33-74): the weights of every panel, summed over
// the panel's points (a point contributes to every panel on its root path), stored xi-major like
// the reference.  pc_compute / pp_compute (:153-192): one thread per target sums its PC entries
// (source proxies, weights read eta-major as the reference reads them unless asked otherwise) and
// its PP entries (source points, (ssh * area) * K as :146).
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
static double stt_tie_ulps() { return 16.0; }

template <bool FILL>
__global__ void k_stt(int64_t nt, const double* __restrict__ tx, const double* __restrict__ ty,
                      const double* __restrict__ tz, const int* __restrict__ count, const int* __restrict__ is_leaf,
                      const int* __restrict__ child1, const double* __restrict__ radius,
                      const double* __restrict__ centre, double theta, int cap, int* __restrict__ queue,
                      int* __restrict__ npc, int* __restrict__ npp, const int* __restrict__ pc_off,
                      const int* __restrict__ pp_off, int* __restrict__ pc_t, int* __restrict__ pc_s,
                      int* __restrict__ pp_t, int* __restrict__ pp_s, const unsigned char* __restrict__ skip,
                      unsigned char* __restrict__ tie, int* __restrict__ overflow, double ulps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    if (FILL && skip[i]) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int* q = queue + i;  // element k of this target's ring at q[(k % cap) * stride]
    int head = 0, tail = 0;
    for (int r = 0; r < 6; ++r) q[(int64_t)(tail++ % cap) * stride] = r;
    const double x1t = tx[i], x2t = ty[i], x3t = tz[i];
    int cpc = 0, cpp = 0;
    bool t_ = false;
    const int pco = FILL ? pc_off[i] : 0, ppo = FILL ? pp_off[i] : 0;
    while (head < tail) {
        int S = q[(int64_t)(head % cap) * stride];
        ++head;
        if (count[S] <= 0) continue;  // tree_module.f90:431
        double dot = x1t * centre[3 * S] + x2t * centre[3 * S + 1] + x3t * centre[3 * S + 2];
        dot = dot > -1.0 ? dot : -1.0;
        dot = dot < 1.0 ? dot : 1.0;
        double dist = acos(dot);
        double sep = radius[S] / dist;
        if (dist > 0 && fabs(sep - theta) <= ulps * 2.220446049250313e-16 * theta) t_ = true;
        if (dist > 0 && sep < theta) {  // :438-449
            if (FILL) { pc_t[pco + cpc] = (int)i; pc_s[pco + cpc] = S; }
            ++cpc;
        } else if (is_leaf[S]) {  // :451-457
            if (FILL) { pp_t[ppo + cpp] = (int)i; pp_s[ppo + cpp] = S; }
            ++cpp;
        } else {  // :459-465
            if (tail - head + 4 > cap) {
                atomicMax(overflow, tail - head + 4);
                return;
            }
            int c1 = child1[S];
            for (int k = 0; k < 4; ++k) q[(int64_t)(tail++ % cap) * stride] = c1 + k;
        }
    }
    if (!FILL) {
        npc[i] = cpc;
        npp[i] = cpp;
        if (t_) tie[i] = 1;
    }
}

// host walk of one target with the host libm (tree_module.f90:419-470)
static void host_stt(sal_geom g, double theta, double x1t, double x2t, double x3t, std::vector<int>& pc,
                     std::vector<int>& pp) {
    std::vector<int> q = {0, 1, 2, 3, 4, 5};
    for (size_t h = 0; h < q.size(); ++h) {
        int S = q[h];
        if (g->count[S] <= 0) continue;
        const double* c = &g->centre[3 * (size_t)S];
        double dot = x1t * c[0] + x2t * c[1] + x3t * c[2];
        dot = dot > -1.0 ? dot : -1.0;
        dot = dot < 1.0 ? dot : 1.0;
        double dist = std::acos(dot);
        double sep = g->radius[S] / dist;
        if (dist > 0 && sep < theta) pc.push_back(S);
        else if (g->is_leaf[S]) pp.push_back(S);
        else
            for (int k = 0; k < 4; ++k) q.push_back(g->child1[S] + k);
    }
}

struct SttBuf {
    DBuf<double> tx, ty, tz;
    int64_t nt = 0;
    DBuf<int> pc_off, pp_off;  // per target
    DBuf<int> order;           // the sums' thread -> target map: targets in spatial order
};

// Thread order of the per-target sums: targets sorted by cube face, then by the Morton code of their
// face angles on a 16384 x 16384 grid, so that the 64 targets of a wave lie close together and walk
// mostly the same source panels (their loads coalesce).  Each target's sum is still one thread's, in
// its own list order: the order changes no result, only the speed.
__global__ void k_stt_key(int64_t nt, const double* __restrict__ tx, const double* __restrict__ ty,
                          const double* __restrict__ tz, int* __restrict__ key, int* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const int f = face_from_xyz(tx[i], ty[i], tz[i]);
    double xi, eta;
    xieta_from_xyz(tx[i], ty[i], tz[i], f, xi, eta);
    const double q = 0.78539816339744830962;
    auto cell = [&](double a) {
        const double u = fmin(fmax((a + q) / (2 * q), 0.0), 1.0);
        return (unsigned)fmin(16383.0, u * 16384.0);
    };
    auto spread = [](unsigned v) {  // 14 bits to the even bit positions
        v = (v | (v << 8)) & 0x00FF00FFu;
        v = (v | (v << 4)) & 0x0F0F0F0Fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    key[i] = (int)(((unsigned)(f - 1) << 28) | (spread(cell(xi)) << 1) | spread(cell(eta)));
    idx[i] = (int)i;
}

static void stt_order(SttBuf* sb, hipStream_t s) {
    const int64_t nt = sb->nt;
    DBuf<int> key(nt), skey(nt), idx(nt);
    sb->order.alloc(std::max<int64_t>(nt, 1));
    k_stt_key<<<grid_for(nt, 256), 256, 0, s>>>(nt, sb->tx.p, sb->ty.p, sb->tz.p, key.p, idx.p);
    SAL_HIP(hipGetLastError());
    sort_pairs_i32(key.p, skey.p, idx.p, sb->order.p, nt, 31, s);
    SAL_HIP(hipStreamSynchronize(s));
}

static SttBuf* g_stt_buf(sal_geom g);

// MOM6 sums: one thread per target point.
//
// The proxies of a panel (pc_interact forms them per interaction, mom6_fast_sum_module.f90:86-98:
// node a in xi, node j in eta, xyz_from_xieta) are p_aj = ir_aj (c + X_a e1 + Y_j e2) on the panel's
// face, with X_a = tan xi_a, Y_j = tan eta_j, ir_aj = 1 / sqrt(1 + X_a^2 + Y_j^2) and (c, e1, e2) the
// face's signed axes: componentwise the same products xyz_from_xieta rounds (1/r, X (1/r), Y (1/r)).
// Per panel the sums read a header {c, e1, e2, X_0 .. X_{nd-1}, Y_0 .. Y_{nd-1}} and one 16-byte
// record {ir, KP1 w ir} per proxy, so a lane's PC pair costs one 16-byte load (against 32 bytes for
// an (x, y, z, w) record: the texture units, not the FP64 pipes, bounded that form).  t.p is formed
// as the reference forms it, (t_x p_x + t_y p_y) + t_z p_z from the same rounded components: for
// the closest point-to-cluster pairs 1 - t.p is about 1e-5, so any other rounding of t.p moves the
// sums by about 1e-11 of their size.  On faces 1-4 (p_x, p_y, p_z) are (ir, X ir, Y ir) up to order
// and sign, p_z = +-Y ir; on faces 5 and 6 p_x = +-Y ir and p_z = +-ir.
constexpr int M6_HDR_FRAME = 9;
__host__ __device__ constexpr int m6_hdr_stride(int nd) { return (M6_HDR_FRAME + 2 * nd + 1) & ~1; }

// One thread per (panel, proxy k = a + j nd); the weight is read eta-major (w[k], as the reference
// reads it) when TRANSPOSED, else xi-major (w[a nd + j]).  KP1 goes with the weight: the PC sums
// evaluate the kernel factor over KP1.
template <bool TRANSPOSED>
__global__ void k_pack_far(int P, int nd, const double* __restrict__ bounds, const int* __restrict__ face,
                           const double* __restrict__ W, double* __restrict__ hdr, double2* __restrict__ rec) {
    const int m = nd * nd, deg = nd - 1, hs = m6_hdr_stride(nd);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)P * m) return;
    const int64_t S = i / m;
    const int k = (int)(i % m), a = k % nd, j = k / nd;
    const double* b = bounds + 6 * S;
    const double ax = tan(bli_node(b[0], b[2], deg, a)), ay = tan(bli_node(b[3], b[5], deg, j));
    const double ir = 1.0 / sqrt(1.0 + ax * ax + ay * ay);  // xyz_from_xieta's 1 / r
    const double w = TRANSPOSED ? W[i] : W[S * m + a * nd + j];
    rec[i] = make_double2(ir, KP1 * w * ir);
    double* h = hdr + S * hs;
    if (j == 0) h[M6_HDR_FRAME + a] = ax;
    if (a == 0) h[M6_HDR_FRAME + nd + j] = ay;
    if (k == 0) {
        // signed axes (c, e1, e2) of xyz_from_xieta's face cases, cube_sphere_module.f90:64-95
        static const signed char ax3[6][9] = {{1, 0, 0, 0, 1, 0, 0, 0, 1},  {0, 1, 0, -1, 0, 0, 0, 0, 1},
                                              {-1, 0, 0, 0, -1, 0, 0, 0, 1}, {0, -1, 0, 1, 0, 0, 0, 0, 1},
                                              {0, 0, 1, 0, 1, 0, -1, 0, 0},  {0, 0, -1, 0, 1, 0, 1, 0, 0}};
        const int f = face[S] - 1;
        for (int q = 0; q < M6_HDR_FRAME; ++q) h[q] = (double)ax3[f][q];
    }
}

// A point and its charge in one 32-byte record (the PP sources), so that a lane fetches a pair's
// operands with two 16-byte loads (the lanes of a wave mostly read different panels).
struct alignas(16) Pt4 {
    double x, y, z, w;
};

// pts[k] = (px, py, pz, q)[k], the points in tree order with their charge (ssh * area)
__global__ void k_pack_pts(int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                           const double* __restrict__ pz, const double* __restrict__ q, Pt4* __restrict__ pts) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    pts[k] = Pt4{px[k], py[k], pz[k], q[k]};
}

// One lane per target (splitting a target's list over 2 or 4 lanes measured no faster), ND = nd
// proxies a side, so that a row's X_a and records stay in registers.
template <int ND>
__global__ __launch_bounds__(64) void k_mom6_sums(int64_t nt, const int* __restrict__ order,
                                                 const double* __restrict__ tx, const double* __restrict__ ty,
                                                 const double* __restrict__ tz, const int* __restrict__ pc_off,
                                                 const int* __restrict__ npc, const int* __restrict__ pc_s,
                                                 const int* __restrict__ pp_off, const int* __restrict__ npp,
                                                 const int* __restrict__ pp_s, const int64_t* __restrict__ start,
                                                 const int* __restrict__ count, const Pt4* __restrict__ pts,
                                                 const double* __restrict__ hdr, const double2* __restrict__ rec,
                                                 double* __restrict__ gx, double* __restrict__ gy, int accumulate,
                                                 int do_pc, int do_pp) {
    constexpr int M = ND * ND, HS = m6_hdr_stride(ND);
    const int item = xcd_item((int)((nt + 63) / 64));
    if (item < 0) return;
    const int64_t th = (int64_t)item * 64 + threadIdx.x;
    if (th >= nt) return;
    const int64_t i = order ? order[th] : th;
    const Tgt T = make_tgt(tx[i], ty[i], tz[i]);
    double ax = 0.0, ay = 0.0;  // unscaled by T.kc
    if (do_pc) {  // pc_interact, :76-107
        // The proxies are well separated from the target (the far form, far_val), and the projections
        // pe = t.x s.y - t.y s.x, pn = t.a s.z - t.z (t.x s.x + t.y s.y) are linear in the source
        // point: the sums form v = sum f w p, per panel in its face frame (sums of F = f w ir, F X_a
        // and F Y_j, rotated by the signed axes), and project v once.  On the 1440 x 1080 grid this
        // stays within 9e-14 relative L2 of an 80-bit evaluation, against 6e-14 for per-pair
        // projections in the reference's order.
        double vx = 0.0, vy = 0.0, vz = 0.0;
        const int c0 = pc_off[i], c1 = c0 + npc[i];
        for (int e = c0; e < c1; ++e) {
            const int64_t S = pc_s[e];
            const double* h = hdr + S * HS;
            const double2* r = rec + S * M;
            double X[ND];
#pragma unroll
            for (int a = 0; a < ND; ++a) X[a] = h[M6_HDR_FRAME + a];
            // u = the signed target components that multiply ir, X ir and Y ir (exact: one term of each
            // sum is +-t_k, the others 0); t.p = (t_x p_x + t_y p_y) + t_z p_z adds the Y ir term last
            // on faces 1-4 and the ir term last on faces 5 and 6
            const double u0 = fma(T.x, h[0], fma(T.y, h[1], T.z * h[2]));
            const double u1 = fma(T.x, h[3], fma(T.y, h[4], T.z * h[5]));
            const double u2 = fma(T.x, h[6], fma(T.y, h[7], T.z * h[8]));
            const bool z_is_ir = h[2] != 0.0;  // faces 5 and 6
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int j = 0; j < ND; ++j) {
                double2 q[ND];
#pragma unroll
                for (int a = 0; a < ND; ++a) q[a] = r[j * ND + a];
                const double Y = h[M6_HDR_FRAME + ND + j];
                double r0 = 0.0, r1 = 0.0;
#pragma unroll
                for (int a = 0; a < ND; ++a) {
                    const double ir = q[a].x;
                    double g;  // the lanes of a wave mostly agree on the face group
                    if (z_is_ir) g = (u2 * (Y * ir) + u1 * (X[a] * ir)) + u0 * ir;
                    else g = (u0 * ir + u1 * (X[a] * ir)) + u2 * (Y * ir);
                    const double F = far_val(g) * q[a].y;
                    r0 += F;
                    r1 = fma(F, X[a], r1);
                }
                s0 += r0;
                s1 += r1;
                s2 = fma(r0, Y, s2);
            }
            vx = fma(s0, h[0], fma(s1, h[3], fma(s2, h[6], vx)));
            vy = fma(s0, h[1], fma(s1, h[4], fma(s2, h[7], vy)));
            vz = fma(s0, h[2], fma(s1, h[5], fma(s2, h[8], vz)));
        }
        ax = fma(T.x, vy, -(T.y * vx));
        ay = fma(T.a, vz, -(T.z * fma(T.x, vx, T.y * vy)));
    }
    if (do_pp) {  // pp_interact, :109-150
        // the points of a near leaf, the reference's expression as written (gfunc_ref_u), each pair
        // projected on the target's east and north in the reference's point order.  Points are taken
        // two at a time: the pair's second point is loaded before its first is evaluated and the next
        // pair's first before the second (two register sets, no copies).
        auto add = [&](const Pt4& c) {
            double pe, pn;
            const double f = gfunc_ref_u(T, c.x, c.y, c.z, pe, pn) * c.w;
            ax = fma(pe, f, ax);
            ay = fma(pn, f, ay);
        };
        const int p0 = pp_off[i], p1 = p0 + npp[i];
        for (int e = p0; e < p1; ++e) {
            const int S = pp_s[e], n = count[S];
            if (n <= 0) continue;
            const Pt4* c = pts + start[S];
            Pt4 ca = c[0];
            int k = 0;
            for (; k + 1 < n; k += 2) {
                const Pt4 cb = c[k + 1];
                add(ca);
                ca = c[k + 2 < n ? k + 2 : k + 1];
                add(cb);
            }
            if (k < n) add(ca);
        }
    }
    ax *= T.kc;
    ay *= T.kc;
    if (accumulate) { gx[i] += ax; gy[i] += ay; }
    else { gx[i] = ax; gy[i] = ay; }
}

// reduce P2M partials of all panels into the xi-major layout of proxy_source_compute
__global__ void k_p2m_reduce_ximajor(int P, int nd, const int* __restrict__ ck_off, const double* __restrict__ wpart,
                                     double* __restrict__ W) {
    int A = blockIdx.x, m = nd * nd;
    for (int w = threadIdx.x; w < m; w += blockDim.x) {
        double a = 0.0;
        for (int c = ck_off[A]; c < ck_off[A + 1]; ++c) a += wpart[(int64_t)c * m + w];
        int i = w % nd, j = w / nd;  // w = xi + eta * nd
        W[(int64_t)A * m + i * nd + j] = a;
    }
}

// P2M partial weights of one chunk of a panel's points (every panel, xi + eta * nd order); 128
// threads, 64 points a step.  The panel's nodes are computed once per block; thread t < 64 forms the
// xi basis of point t times its charge, thread 64 + t its eta basis, and thread w owns weights w,
// w + 128, w + 256 (two LDS reads and one FMA per point).
__global__ __launch_bounds__(128) void k_p2m_all(int nd, const int* __restrict__ ck_panel,
                                                 const int* __restrict__ ck_first, const int64_t* __restrict__ start,
                                                 const int* __restrict__ count, const double* __restrict__ bounds,
                                                 const int* __restrict__ perm, const double* __restrict__ xi,
                                                 const double* __restrict__ eta, const double* __restrict__ q,
                                                 double* __restrict__ wpart) {
    __shared__ double fb[2][64][MAX_DEG + 1], nodes[2][MAX_DEG + 1];
    const int A = ck_panel[blockIdx.x], f0 = ck_first[blockIdx.x];
    const int m = nd * nd, deg = nd - 1;
    const int cnt = min(P2M_CHUNK, count[A] - f0);
    const int64_t st = start[A] + f0;
    const double* b = bounds + 6 * (int64_t)A;
    const int tid = threadIdx.x, d = tid >> 6, tp = tid & 63;
    if (tid < 2 * nd) {
        const int dd = tid / nd, i = tid % nd;
        nodes[dd][i] = bli_node(b[3 * dd], b[3 * dd + 2], deg, i);
    }
    double acc[3] = {0.0, 0.0, 0.0};  // m <= 289 < 3 * 128
    for (int c0 = 0; c0 < cnt; c0 += 64) {
        const int nc = min(64, cnt - c0);
        __syncthreads();
        if (tp < nc) {
            const int64_t k = st + c0 + tp;
            const int p = perm[k];
            bli_basis_1d(d ? eta[p] : xi[p], nodes[d], deg, fb[d][tp]);
            if (d == 0) {  // the xi basis times the point's charge
                const double qk = q[k];
                for (int i = 0; i <= deg; ++i) fb[0][tp][i] *= qk;
            }
        }
        __syncthreads();
        for (int r = 0; r < 3; ++r) {
            const int w = tid + 128 * r;
            if (w < m) {
                const int i = w % nd, j = w / nd;
                double a = acc[r];
                for (int t = 0; t < nc; ++t) a = fma(fb[0][t][i], fb[1][t][j], a);
                acc[r] = a;
            }
        }
    }
    for (int r = 0; r < 3; ++r) {
        const int w = tid + 128 * r;
        if (w < m) wpart[(int64_t)blockIdx.x * m + w] = acc[r];
    }
}

__global__ void k_gather_q1(int64_t n, const int* __restrict__ perm, const double* __restrict__ area,
                            const double* __restrict__ ssh, double* __restrict__ q) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = perm[k];
    q[k] = ssh[p] * area[p];
}

static void mom6_weights(sal_geom g, const double* d_q, int degree, DBuf<double>& W, hipStream_t s) {
    const int nd = degree + 1, m = nd * nd, P = g->P;
    std::vector<int> ck_panel, ck_first, ck_off(P + 1, 0);
    for (int i = 0; i < P; ++i) {
        for (int f = 0; f < g->count[i]; f += P2M_CHUNK) {
            ck_panel.push_back(i);
            ck_first.push_back(f);
        }
        ck_off[i + 1] = (int)ck_panel.size();
    }
    int nck = (int)ck_panel.size();
    DBuf<int> d_ckp(std::max(nck, 1)), d_ckf(std::max(nck, 1)), d_cko(P + 1);
    h2d(d_ckp.p, ck_panel.data(), nck, s);
    h2d(d_ckf.p, ck_first.data(), nck, s);
    h2d(d_cko.p, ck_off.data(), P + 1, s);
    DBuf<double> wpart((size_t)std::max(nck, 1) * m);
    W.alloc((size_t)P * m);
    SAL_HIP(hipMemsetAsync(W.p, 0, (size_t)P * m * sizeof(double), s));
    if (nck > 0) {
        ProfScope ps("p2m_all", s);
        k_p2m_all<<<nck, 128, 0, s>>>(nd, d_ckp.p, d_ckf.p, g->d_start.p, g->d_count.p, g->d_bounds.p, g->perm.p,
                                      g->xi.p, g->eta.p, d_q, wpart.p);
    }
    k_p2m_reduce_ximajor<<<P, 128, 0, s>>>(P, nd, d_cko.p, wpart.p, W.p);
    SAL_HIP(hipGetLastError());
    SAL_HIP(hipStreamSynchronize(s));
}

}  // namespace sal

// STT target buffers live with the handle (one set per handle)
namespace sal {
static std::vector<std::pair<sal_geom, SttBuf*>> g_stt_bufs;
static SttBuf* g_stt_buf(sal_geom g) {
    for (auto& p : g_stt_bufs)
        if (p.first == g) return p.second;
    auto b = new SttBuf();
    g_stt_bufs.push_back({g, b});
    return b;
}
void stt_get(sal_geom g, SttState& st) {
    SttBuf* sb = nullptr;
    for (auto& p : g_stt_bufs)
        if (p.first == g) sb = p.second;
    st.nt = sb && g->stt_counts.p ? sb->nt : 0;
    const size_t nt = (size_t)st.nt;
    st.tx.resize(nt); st.ty.resize(nt); st.tz.resize(nt); st.pc_off.resize(nt); st.pp_off.resize(nt);
    if (!nt) return;
    hipStream_t s = g->stream;
    d2h(st.tx.data(), sb->tx.p, nt, s);
    d2h(st.ty.data(), sb->ty.p, nt, s);
    d2h(st.tz.data(), sb->tz.p, nt, s);
    d2h(st.pc_off.data(), sb->pc_off.p, nt, s);
    d2h(st.pp_off.data(), sb->pp_off.p, nt, s);
    SAL_HIP(hipStreamSynchronize(s));
}
void stt_set(sal_geom g, const SttState& st) {
    const size_t nt = (size_t)st.nt;
    if (!nt) return;
    SttBuf* sb = g_stt_buf(g);
    hipStream_t s = g->stream;
    sb->nt = st.nt;
    sb->tx.alloc(nt); sb->ty.alloc(nt); sb->tz.alloc(nt); sb->pc_off.alloc(nt); sb->pp_off.alloc(nt);
    h2d(sb->tx.p, st.tx.data(), nt, s);
    h2d(sb->ty.p, st.ty.data(), nt, s);
    h2d(sb->tz.p, st.tz.data(), nt, s);
    h2d(sb->pc_off.p, st.pc_off.data(), nt, s);
    h2d(sb->pp_off.p, st.pp_off.data(), nt, s);
    SAL_HIP(hipStreamSynchronize(s));
    stt_order(sb, s);
}
void stt_release(sal_geom g) {
    for (size_t k = 0; k < g_stt_bufs.size(); ++k)
        if (g_stt_bufs[k].first == g) {
            delete g_stt_bufs[k].second;
            g_stt_bufs.erase(g_stt_bufs.begin() + k);
            return;
        }
}
}  // namespace sal

using namespace sal;

extern "C" {

int sal_stt(sal_geom g, double theta, const double* x, const double* y, const double* z, int64_t n_own,
            int64_t* n_pc, int64_t* n_pp) {
    return guard([&] {
        SAL_REQUIRE(g && x && y && z && n_pc && n_pp, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(theta > 0 && n_own >= 0 && n_own < INT32_MAX, SAL_ERR_ARG, "bad arguments");
        hipStream_t s = g->stream;
        SttBuf* sb = g_stt_buf(g);
        sb->nt = n_own;
        g->npc = g->npp = 0;
        *n_pc = *n_pp = 0;
        if (n_own == 0) return;
        sb->tx.alloc(n_own); sb->ty.alloc(n_own); sb->tz.alloc(n_own);
        h2d(sb->tx.p, x, n_own, s);
        h2d(sb->ty.p, y, n_own, s);
        h2d(sb->tz.p, z, n_own, s);
        DBuf<int> npc(n_own), npp(n_own), overflow(1);
        DBuf<unsigned char> tie(n_own), skip(n_own);
        SAL_HIP(hipMemsetAsync(tie.p, 0, n_own, s));
        SAL_HIP(hipMemsetAsync(skip.p, 0, n_own, s));
        const unsigned nb = grid_for(n_own, 128);
        const int64_t nthreads = (int64_t)nb * 128;
        int cap = 64;
        DBuf<int> queue;
        for (;;) {  // count pass; grow the per-target ring until no walk overflows it
            queue.alloc((size_t)cap * nthreads);
            SAL_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
            k_stt<false><<<nb, 128, 0, s>>>(n_own, sb->tx.p, sb->ty.p, sb->tz.p, g->d_count.p, g->d_is_leaf.p,
                                            g->d_child1.p, g->d_radius.p, g->d_centre.p, theta, cap, queue.p, npc.p,
                                            npp.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, skip.p,
                                            tie.p, overflow.p, stt_tie_ulps());
            SAL_HIP(hipGetLastError());
            int ov = 0;
            d2h(&ov, overflow.p, 1, s);
            SAL_HIP(hipStreamSynchronize(s));
            if (ov == 0) break;
            SAL_REQUIRE(cap < (1 << 16), SAL_ERR_LIMIT, "single tree traversal queue exceeds 65536 panels");
            while (cap < ov) cap *= 2;
            cap *= 2;
        }
        // host walks for near-theta targets
        std::vector<unsigned char> htie(n_own);
        d2h(htie.data(), tie.p, n_own, s);
        SAL_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> tidx;
        std::vector<std::vector<int>> hpc, hpp;
        for (int64_t i = 0; i < n_own; ++i)
            if (htie[i]) {
                tidx.push_back(i);
                hpc.emplace_back();
                hpp.emplace_back();
                host_stt(g, theta, x[i], y[i], z[i], hpc.back(), hpp.back());
                int a = (int)hpc.back().size(), b = (int)hpp.back().size();
                h2d(npc.p + i, &a, 1, s);
                h2d(npp.p + i, &b, 1, s);
                SAL_HIP(hipStreamSynchronize(s));
            }
        if (!tidx.empty()) SAL_HIP(hipMemcpyAsync(skip.p, tie.p, n_own, hipMemcpyDeviceToDevice, s));
        sb->pc_off.alloc(n_own);
        sb->pp_off.alloc(n_own);
        int64_t tpc = exclusive_scan_i32(npc.p, sb->pc_off.p, n_own, s);
        int64_t tpp = exclusive_scan_i32(npp.p, sb->pp_off.p, n_own, s);
        g->pc_t.alloc(std::max<int64_t>(tpc, 1)); g->pc_s.alloc(std::max<int64_t>(tpc, 1));
        g->pp_t.alloc(std::max<int64_t>(tpp, 1)); g->pp_s.alloc(std::max<int64_t>(tpp, 1));
        k_stt<true><<<nb, 128, 0, s>>>(n_own, sb->tx.p, sb->ty.p, sb->tz.p, g->d_count.p, g->d_is_leaf.p,
                                       g->d_child1.p, g->d_radius.p, g->d_centre.p, theta, cap, queue.p, npc.p, npp.p,
                                       sb->pc_off.p, sb->pp_off.p, g->pc_t.p, g->pc_s.p, g->pp_t.p, g->pp_s.p, skip.p,
                                       tie.p, overflow.p, stt_tie_ulps());
        SAL_HIP(hipGetLastError());
        for (size_t k = 0; k < tidx.size(); ++k) {
            int64_t i = tidx[k];
            int a = 0, b = 0;
            d2h(&a, sb->pc_off.p + i, 1, s);
            d2h(&b, sb->pp_off.p + i, 1, s);
            SAL_HIP(hipStreamSynchronize(s));
            std::vector<int> ti(std::max(hpc[k].size(), hpp[k].size()), (int)i);
            h2d(g->pc_t.p + a, ti.data(), hpc[k].size(), s);
            h2d(g->pc_s.p + a, hpc[k].data(), hpc[k].size(), s);
            h2d(g->pp_t.p + b, ti.data(), hpp[k].size(), s);
            h2d(g->pp_s.p + b, hpp[k].data(), hpp[k].size(), s);
            SAL_HIP(hipStreamSynchronize(s));
        }
        // keep the per-target counts with the offsets for the sums
        sb->pc_off.ensure(n_own);
        SAL_HIP(hipStreamSynchronize(s));
        g->npc = tpc;
        g->npp = tpp;
        stt_order(sb, s);
        g->pair_ties += (int64_t)tidx.size();
        // counts are re-derived from offsets in the sums: store them after the offsets
        DBuf<int> cnts(2 * (size_t)n_own);
        SAL_HIP(hipMemcpyAsync(cnts.p, npc.p, n_own * sizeof(int), hipMemcpyDeviceToDevice, s));
        SAL_HIP(hipMemcpyAsync(cnts.p + n_own, npp.p, n_own * sizeof(int), hipMemcpyDeviceToDevice, s));
        SAL_HIP(hipStreamSynchronize(s));
        g->stt_counts = std::move(cnts);
        *n_pc = tpc;
        *n_pp = tpp;
    });
}

int sal_stt_counts(sal_geom g, int64_t* n_targets, int64_t* n_pc, int64_t* n_pp) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        if (n_targets) {
            int64_t nt = 0;
            for (auto& p : g_stt_bufs)
                if (p.first == g && g->stt_counts.p) nt = p.second->nt;
            *n_targets = nt;
        }
        if (n_pc) *n_pc = g->npc;
        if (n_pp) *n_pp = g->npp;
    });
}

int sal_stt_export(sal_geom g, int* pc_target, int* pc_source, int* pp_target, int* pp_source) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        hipStream_t s = g->stream;
        if (pc_target) d2h(pc_target, g->pc_t.p, g->npc, s);
        if (pc_source) d2h(pc_source, g->pc_s.p, g->npc, s);
        if (pp_target) d2h(pp_target, g->pp_t.p, g->npp, s);
        if (pp_source) d2h(pp_source, g->pp_s.p, g->npp, s);
        SAL_HIP(hipStreamSynchronize(s));
    });
}

int sal_mom6_proxy_source(sal_geom g, const double* area, const double* ssh, int degree, double* w) {
    return guard([&] {
        SAL_REQUIRE(g && area && ssh && w, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        hipStream_t s = g->stream;
        const int64_t n = g->n;
        DBuf<double> da(n), dsh(n), q(n), W;
        h2d(da.p, area, n, s);
        h2d(dsh.p, ssh, n, s);
        k_gather_q1<<<grid_for(n, 256), 256, 0, s>>>(n, g->perm.p, da.p, dsh.p, q.p);
        mom6_weights(g, q.p, degree, W, s);
        d2h(w, W.p, (size_t)g->P * (degree + 1) * (degree + 1), s);
        SAL_HIP(hipStreamSynchronize(s));
    });
}

static void mom6_run(sal_geom g, const double* W, const double* q, int degree, int transposed, int do_pc, int do_pp,
                     double* gx, double* gy, int accumulate) {
    SttBuf* sb = g_stt_buf(g);
    SAL_REQUIRE(sb->nt > 0 && g->stt_counts.p, SAL_ERR_STATE, "no single-tree lists: call sal_stt first");
    hipStream_t s = g->stream;
    const int64_t nt = sb->nt;
    DBuf<double> ox(nt), oy(nt);
    const int nd = degree + 1, m = nd * nd;
    DBuf<double> hdr(do_pc ? (size_t)g->P * m6_hdr_stride(nd) : 1);
    DBuf<double2> rec(do_pc ? (size_t)g->P * m : 1);
    DBuf<Pt4> pts(do_pp ? (size_t)g->n : 1);
    if (do_pc) {
        const unsigned grid = grid_for((int64_t)g->P * m, 256);
        if (transposed)
            k_pack_far<true><<<grid, 256, 0, s>>>(g->P, nd, g->d_bounds.p, g->d_face.p, W, hdr.p, rec.p);
        else
            k_pack_far<false><<<grid, 256, 0, s>>>(g->P, nd, g->d_bounds.p, g->d_face.p, W, hdr.p, rec.p);
        SAL_HIP(hipGetLastError());
    }
    if (do_pp) {
        k_pack_pts<<<grid_for(g->n, 256), 256, 0, s>>>(g->n, g->px.p, g->py.p, g->pz.p, q, pts.p);
        SAL_HIP(hipGetLastError());
    }
    {
        ProfScope ps(do_pc ? (do_pp ? "mom6_sums" : "mom6_pc") : "mom6_pp", s);
        SAL_REQUIRE((nt + 63) / 64 < INT32_MAX / 2, SAL_ERR_LIMIT, "too many targets for the per-point sums");
        const dim3 grid = xcd_grid((int)((nt + 63) / 64));
        auto launch = [&](auto nd_c) {
            k_mom6_sums<decltype(nd_c)::value><<<grid, 64, 0, s>>>(
                nt, sb->order.p, sb->tx.p, sb->ty.p, sb->tz.p, sb->pc_off.p, g->stt_counts.p, g->pc_s.p,
                sb->pp_off.p, g->stt_counts.p + nt, g->pp_s.p, g->d_start.p, g->d_count.p, pts.p, hdr.p, rec.p,
                ox.p, oy.p, 0, do_pc, do_pp);
        };
        static_assert(MAX_DEG == 16, "one instantiation per degree");
        switch (nd) {
#define SAL_M6_ND(k) case k: launch(std::integral_constant<int, k>{}); break;
            SAL_M6_ND(1) SAL_M6_ND(2) SAL_M6_ND(3) SAL_M6_ND(4) SAL_M6_ND(5) SAL_M6_ND(6) SAL_M6_ND(7)
            SAL_M6_ND(8) SAL_M6_ND(9) SAL_M6_ND(10) SAL_M6_ND(11) SAL_M6_ND(12) SAL_M6_ND(13) SAL_M6_ND(14)
            SAL_M6_ND(15) SAL_M6_ND(16) SAL_M6_ND(17)
#undef SAL_M6_ND
            default: SAL_REQUIRE(false, SAL_ERR_ARG, "degree must be in [0, 16]");
        }
    }
    SAL_HIP(hipGetLastError());
    std::vector<double> hx(nt), hy(nt);
    d2h(hx.data(), ox.p, nt, s);
    d2h(hy.data(), oy.p, nt, s);
    SAL_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < nt; ++i) {
        gx[i] = accumulate ? gx[i] + hx[i] : hx[i];
        gy[i] = accumulate ? gy[i] + hy[i] : hy[i];
    }
}

static void mom6_q(sal_geom g, const double* area, const double* ssh, DBuf<double>& q) {
    hipStream_t s = g->stream;
    const int64_t n = g->n;
    DBuf<double> da(n), dsh(n);
    q.alloc(n);
    h2d(da.p, area, n, s);
    h2d(dsh.p, ssh, n, s);
    k_gather_q1<<<grid_for(n, 256), 256, 0, s>>>(n, g->perm.p, da.p, dsh.p, q.p);
    SAL_HIP(hipGetLastError());
    SAL_HIP(hipStreamSynchronize(s));
}

int sal_mom6_pc_compute(sal_geom g, const double* w, int degree, int transposed, double* gx, double* gy,
                        int accumulate) {
    return guard([&] {
        SAL_REQUIRE(g && w && gx && gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        size_t nw = (size_t)g->P * (degree + 1) * (degree + 1);
        DBuf<double> W(nw);
        h2d(W.p, w, nw, g->stream);
        mom6_run(g, W.p, nullptr, degree, transposed, 1, 0, gx, gy, accumulate);
    });
}

int sal_mom6_pp_compute(sal_geom g, const double* area, const double* ssh, double* gx, double* gy, int accumulate) {
    return guard([&] {
        SAL_REQUIRE(g && area && ssh && gx && gy, SAL_ERR_ARG, "null pointer");
        DBuf<double> q;
        mom6_q(g, area, ssh, q);
        mom6_run(g, nullptr, q.p, 0, 0, 0, 1, gx, gy, accumulate);
    });
}

int sal_mom6_sums(sal_geom g, const double* area, const double* ssh, int degree, int transposed, double* gx,
                  double* gy, int accumulate) {
    return guard([&] {
        SAL_REQUIRE(g && area && ssh && gx && gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        DBuf<double> q, W;
        mom6_q(g, area, ssh, q);
        mom6_weights(g, q.p, degree, W, g->stream);
        mom6_run(g, W.p, q.p, degree, transposed, 1, 1, gx, gy, accumulate);
    });
}

}  // extern "C"
