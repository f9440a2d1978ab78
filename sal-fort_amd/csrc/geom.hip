// Cubed-sphere tree and interaction lists on the GPU.
//
// Tree (replaces tree_traverse, tree_module.f90:59-246).  The reference refines panels first-in
// first-out, so panel ids are breadth-first and every level is a contiguous id range; children of
// the k-th refined panel of a level get ids level_end + 4k .. +3 (:139-173).  The build is
// level-synchronous: the O(N) work (face and face angles of every point, then one stable child
// decision per point per level) runs on the GPU, and the O(P) panel bookkeeping (refine flags,
// child bounds mid = 0.5*(a+b), exactly the reference's arithmetic) runs on the host between levels.
// Counts are integer atomics, so they are order independent.
//
// Face angles come from the device's atan.  A child decision can only differ from the reference's
// (glibc atan) when xi or eta lies within a few ulp of the panel midline it is compared with; such
// points are collected, their angles recomputed with the host's atan, and the tree rebuilt.
// Panel centres and radii (:223-244) are O(P) and computed with the host's tan/acos so they are the
// reference's bits.
//
// Interaction list (replaces DUAL_TREE_TRAVERSAL, tree_module.f90:277-398).  The reference's queue
// is a concatenation of generations (each refinement appends 4 pairs, :342-385), so the list order
// is reproduced generation by generation: classify every pair of a generation, exclusive-scan the
// "emit" and "refine" flags, scatter.  MAC decisions within a few ulp of theta are re-decided on
// the host with acos from libm.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

static thread_local std::string t_error;
void set_error(const std::string& msg) { t_error = msg; }
static double g_tie_ulps = 16.0;

int64_t exclusive_scan_i32(const int* d_in, int* d_out, int64_t n, hipStream_t s) {
    if (n <= 0) return 0;
    size_t tmp = 0;
    SAL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_in, d_out, (int)n, s));
    DBuf<char> t(tmp + 1);
    SAL_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, d_in, d_out, (int)n, s));
    int a = 0, b = 0;
    d2h(&a, d_out + n - 1, 1, s);
    d2h(&b, d_in + n - 1, 1, s);
    SAL_HIP(hipStreamSynchronize(s));
    return (int64_t)a + (int64_t)b;
}

void sort_pairs_i32(const int* kin, int* kout, const int* vin, int* vout, int64_t n, int bits, hipStream_t s) {
    if (n <= 0) return;
    size_t tmp = 0;
    SAL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, (int)n, 0, bits, s));
    DBuf<char> t(tmp + 1);
    SAL_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, kin, kout, vin, vout, (int)n, 0, bits, s));
}

static int bits_for(int64_t v) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) <= v) ++b;
    return b;
}

// ------------------------------------------------------------------------------ tree kernels

__global__ void k_face_xieta(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, int* __restrict__ face, double* __restrict__ xi,
                             double* __restrict__ eta, const unsigned char* __restrict__ resolved) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int f = face_from_xyz(x[p], y[p], z[p]);
    face[p] = f;
    if (resolved && resolved[p]) return;  // host-computed angles already in place
    double a, b;
    xieta_from_xyz(x[p], y[p], z[p], f, a, b);
    xi[p] = a;
    eta[p] = b;
}

// root membership and per-face counts (a block histogram first: six counters would serialise)
__global__ void k_roots(int64_t n, const int* __restrict__ face, int* __restrict__ anc0, int* __restrict__ count) {
    __shared__ int h[6];
    if (threadIdx.x < 6) h[threadIdx.x] = 0;
    __syncthreads();
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        int f = face[p] - 1;
        anc0[p] = f;
        atomicAdd(&h[f], 1);
    }
    __syncthreads();
    if (threadIdx.x < 6 && h[threadIdx.x]) atomicAdd(&count[threadIdx.x], h[threadIdx.x]);
}

__device__ inline bool near_mid(double v, double m, double ulps) {
    // m == 0: the comparison v < 0 is the sign of a quotient, exact on both sides
    return m != 0.0 && fabs(v - m) <= ulps * 2.220446049250313e-16 * fabs(m);
}

// Child of every point of a refined level-L panel, and the children's point counts.  The new
// level's panels are [lo_new, lo_new + n_new).  HIST: the level has at most PART_HIST panels, so a
// block counts its points in an LDS histogram and adds each non-zero bin to the global count once
// (the top levels have a few dozen panels: one global atomic per point would serialise on them);
// the grid is then a fixed number of blocks striding over the points.  Otherwise one global atomic
// per point, spread over many counters.
constexpr int PART_HIST = 16384;  // 64 KB of LDS
constexpr int PART_BLOCKS = 1024;
template <bool HIST>
__global__ __launch_bounds__(256) void k_partition(int64_t n, const int* __restrict__ anc_l, int* __restrict__ anc_n,
                                                   const int* __restrict__ child1, const double* __restrict__ bounds,
                                                   const double* __restrict__ xi, const double* __restrict__ eta,
                                                   int* __restrict__ count, int lo_new, int n_new,
                                                   const unsigned char* __restrict__ resolved, int* __restrict__ tie_n,
                                                   unsigned char* __restrict__ tie_flag, double ulps) {
    __shared__ int h[HIST ? PART_HIST : 1];
    if (HIST) {
        for (int i = threadIdx.x; i < n_new; i += blockDim.x) h[i] = 0;
        __syncthreads();
    }
    const int64_t step = HIST ? (int64_t)gridDim.x * blockDim.x : n;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += step) {
        const int a = anc_l[p];
        const int c1 = a >= 0 ? child1[a] : -1;
        if (c1 < 0) {
            anc_n[p] = -1;
            continue;
        }
        const double mx = bounds[6 * (int64_t)a + 1], me = bounds[6 * (int64_t)a + 4];
        const double u = xi[p], v = eta[p];
        const int w = (u < mx) ? ((v < me) ? 0 : 3) : ((v < me) ? 1 : 2);  // tree_module.f90:180-192
        anc_n[p] = c1 + w;
        const int bin = c1 + w - lo_new;  // in [0, n_new) for a panel refined at this level
        if (HIST && (unsigned)bin < (unsigned)n_new) atomicAdd(&h[bin], 1);
        else atomicAdd(&count[c1 + w], 1);
        if (!resolved[p] && (near_mid(u, mx, ulps) || near_mid(v, me, ulps))) {
            tie_flag[p] = 1;
            atomicAdd(tie_n, 1);
        }
    }
    if (HIST) {
        __syncthreads();
        for (int i = threadIdx.x; i < n_new; i += blockDim.x)
            if (h[i]) atomicAdd(&count[lo_new + i], h[i]);
    }
}

__global__ void k_scatter_angles(int m, const int* __restrict__ idx, const double* __restrict__ vxi,
                                 const double* __restrict__ veta, double* __restrict__ xi, double* __restrict__ eta,
                                 unsigned char* __restrict__ resolved) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int p = idx[k];
    xi[p] = vxi[k];
    eta[p] = veta[k];
    resolved[p] = 1;
}

// coordinates of the flagged points, packed: out[3 k + c] = (x, y, z)[idx[k]]
__global__ void k_gather_xyz(int m, const int* __restrict__ idx, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int p = idx[k];
    out[3 * k] = x[p];
    out[3 * k + 1] = y[p];
    out[3 * k + 2] = z[p];
}

__global__ void k_leaf_key(int64_t n, int nlev, const int* __restrict__ anc, const int64_t* __restrict__ start,
                           int* __restrict__ key, int* __restrict__ val) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int leaf = anc[p];
    for (int l = 1; l < nlev; ++l) {
        int a = anc[(int64_t)l * n + p];
        if (a < 0) break;
        leaf = a;
    }
    key[p] = (int)start[leaf];
    val[p] = (int)p;
}

__global__ void k_layout(int64_t n, const int* __restrict__ perm, int* __restrict__ iperm,
                         const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                         double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = perm[k];
    iperm[p] = (int)k;
    px[k] = x[p];
    py[k] = y[p];
    pz[k] = z[p];
}

__global__ void k_level_key(int64_t n, const int* __restrict__ anc_l, int sentinel, int* __restrict__ key,
                            int* __restrict__ val) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = anc_l[p];
    key[p] = a >= 0 ? a : sentinel;
    val[p] = (int)p;
}

static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------------------------ tree build (host driver)

// Host: radius and centre of every panel with the host libm (tree_module.f90:223-244).
static void host_panel_geometry(sal_geom g) {
    g->radius.assign(g->P, 0.0);
    g->centre.assign(3 * (size_t)g->P, 0.0);
    for (int i = 0; i < g->P; ++i) {
        const double* b = &g->bounds[6 * (size_t)i];
        double x1, x2, x3, y1, y2, y3;
        xyz_from_xieta(b[1], b[4], g->face[i], x1, x2, x3);
        g->centre[3 * (size_t)i] = x1;
        g->centre[3 * (size_t)i + 1] = x2;
        g->centre[3 * (size_t)i + 2] = x3;
        const double cx[4] = {b[0], b[0], b[2], b[2]};
        const double ce[4] = {b[3], b[5], b[5], b[3]};
        double r = 0.0;
        for (int k = 0; k < 4; ++k) {
            xyz_from_xieta(cx[k], ce[k], g->face[i], y1, y2, y3);
            double dot = x1 * y1 + x2 * y2 + x3 * y3;
            dot = dot < 1.0 ? dot : 1.0;
            dot = dot > -1.0 ? dot : -1.0;
            double d = std::acos(dot);
            r = (k == 0 || d > r) ? d : r;
        }
        g->radius[i] = r;
    }
}

static void build_levels(sal_geom g, DBuf<unsigned char>& resolved, DBuf<int>& tie_n,
                         DBuf<unsigned char>& tie_flag, int64_t& n_ties) {
    const int64_t n = g->n;
    hipStream_t s = g->stream;
    const double pi = 4.0 * 0.78539816339744830962;
    g->P = 6;
    g->level.assign(6, 0);
    g->is_leaf.assign(6, 1);
    g->parent.assign(6, -1);
    g->child1.assign(6, -1);
    g->face.resize(6);
    g->bounds.assign(36, 0.0);
    for (int i = 0; i < 6; ++i) {  // tree_module.f90:76-88
        g->face[i] = i + 1;
        double* b = &g->bounds[6 * i];
        b[0] = -pi / 4.0; b[1] = 0.0; b[2] = pi / 4.0;
        b[3] = -pi / 4.0; b[4] = 0.0; b[5] = pi / 4.0;
    }
    g->level_lo = {0, 6};

    const int max_levels = 48;
    int lev_cap = 16;
    g->anc.alloc((size_t)lev_cap * (size_t)std::max<int64_t>(n, 1));
    size_t pcap = 1024;
    DBuf<int> d_cnt(pcap), d_c1(pcap);
    DBuf<double> d_b(6 * pcap);
    SAL_HIP(hipMemsetAsync(d_cnt.p, 0, pcap * sizeof(int), s));
    SAL_HIP(hipMemsetAsync(tie_n.p, 0, sizeof(int), s));
    SAL_HIP(hipMemsetAsync(tie_flag.p, 0, (size_t)n, s));
    k_face_xieta<<<grid_for(n, 256), 256, 0, s>>>(n, g->x.p, g->y.p, g->z.p, g->pface.p, g->xi.p, g->eta.p,
                                                  resolved.p);
    k_roots<<<grid_for(n, 256), 256, 0, s>>>(n, g->pface.p, g->anc.p, d_cnt.p);
    SAL_HIP(hipGetLastError());
    g->count.assign(6, 0);
    d2h(g->count.data(), d_cnt.p, 6, s);
    SAL_HIP(hipStreamSynchronize(s));

    for (int L = 0;; ++L) {
        int lo = g->level_lo[L], hi = g->level_lo[L + 1];
        int nref = 0;
        for (int i = lo; i < hi; ++i)
            if (g->count[i] >= g->thresh) ++nref;  // tree_module.f90:130
        if (nref == 0) break;
        SAL_REQUIRE(L + 1 < max_levels, SAL_ERR_LIMIT,
                    "tree deeper than 48 levels: at least cluster_thresh coincident points");
        int newP = g->P + 4 * nref;
        for (int i = lo; i < hi; ++i) {
            if (g->count[i] < g->thresh) continue;
            int c = g->P;
            g->is_leaf[i] = 0;
            g->child1[i] = c;
            const double* b = &g->bounds[6 * (size_t)i];
            double mnx = b[0], mdx = b[1], mxx = b[2], mne = b[3], mde = b[4], mxe = b[5];
            g->bounds.resize(6 * (size_t)(c + 4));
            double* q = &g->bounds[6 * (size_t)c];  // tree_module.f90:150-168
            double cb[4][6] = {{mnx, 0.5 * (mnx + mdx), mdx, mne, 0.5 * (mne + mde), mde},
                               {mdx, 0.5 * (mdx + mxx), mxx, mne, 0.5 * (mne + mde), mde},
                               {mdx, 0.5 * (mdx + mxx), mxx, mde, 0.5 * (mde + mxe), mxe},
                               {mnx, 0.5 * (mnx + mdx), mdx, mde, 0.5 * (mde + mxe), mxe}};
            std::memcpy(q, cb, sizeof(cb));
            for (int k = 0; k < 4; ++k) {
                g->level.push_back(g->level[i] + 1);
                g->is_leaf.push_back(1);
                g->parent.push_back(i);
                g->child1.push_back(-1);
                g->face.push_back(g->face[i]);
                g->count.push_back(0);
            }
            g->P += 4;
        }
        if ((size_t)newP > pcap) {
            size_t nc = pcap;
            while (nc < (size_t)newP) nc *= 2;
            DBuf<int> nc_cnt(nc);
            SAL_HIP(hipMemsetAsync(nc_cnt.p, 0, nc * sizeof(int), s));
            SAL_HIP(hipMemcpyAsync(nc_cnt.p, d_cnt.p, pcap * sizeof(int), hipMemcpyDeviceToDevice, s));
            d_cnt = std::move(nc_cnt);
            d_c1.alloc(nc);
            d_b.alloc(6 * nc);
            pcap = nc;
        }
        h2d(d_c1.p, g->child1.data(), (size_t)g->P, s);
        h2d(d_b.p, g->bounds.data(), 6 * (size_t)g->P, s);
        if (L + 2 > lev_cap) {
            int nl = lev_cap * 2;
            DBuf<int> na((size_t)nl * (size_t)n);
            SAL_HIP(hipMemcpyAsync(na.p, g->anc.p, (size_t)lev_cap * n * sizeof(int), hipMemcpyDeviceToDevice, s));
            g->anc = std::move(na);
            lev_cap = nl;
        }
        const int n_new = newP - hi;
        if (n_new <= PART_HIST)
            k_partition<true><<<std::min<unsigned>(grid_for(n, 256), PART_BLOCKS), 256, 0, s>>>(
                n, g->anc.p + (int64_t)L * n, g->anc.p + (int64_t)(L + 1) * n, d_c1.p, d_b.p, g->xi.p, g->eta.p,
                d_cnt.p, hi, n_new, resolved.p, tie_n.p, tie_flag.p, g_tie_ulps);
        else
            k_partition<false><<<grid_for(n, 256), 256, 0, s>>>(
                n, g->anc.p + (int64_t)L * n, g->anc.p + (int64_t)(L + 1) * n, d_c1.p, d_b.p, g->xi.p, g->eta.p,
                d_cnt.p, hi, n_new, resolved.p, tie_n.p, tie_flag.p, g_tie_ulps);
        SAL_HIP(hipGetLastError());
        d2h(g->count.data() + hi, d_cnt.p + hi, (size_t)(newP - hi), s);
        SAL_HIP(hipStreamSynchronize(s));
        g->level_lo.push_back(newP);
    }
    g->nlev = (int)g->level_lo.size() - 1;
    int nt = 0;
    d2h(&nt, tie_n.p, 1, s);
    SAL_HIP(hipStreamSynchronize(s));
    n_ties = nt;
}

// device copies of the host panel arrays
void upload_panels(sal_geom g) {
    const int P = g->P;
    hipStream_t s = g->stream;
    g->d_level.alloc(P); g->d_is_leaf.alloc(P); g->d_parent.alloc(P); g->d_child1.alloc(P);
    g->d_face.alloc(P); g->d_count.alloc(P); g->d_bounds.alloc(6 * (size_t)P); g->d_radius.alloc(P);
    g->d_centre.alloc(3 * (size_t)P); g->d_start.alloc(P);
    h2d(g->d_level.p, g->level.data(), P, s);
    h2d(g->d_is_leaf.p, g->is_leaf.data(), P, s);
    h2d(g->d_parent.p, g->parent.data(), P, s);
    h2d(g->d_child1.p, g->child1.data(), P, s);
    h2d(g->d_face.p, g->face.data(), P, s);
    h2d(g->d_count.p, g->count.data(), P, s);
    h2d(g->d_bounds.p, g->bounds.data(), 6 * (size_t)P, s);
    h2d(g->d_radius.p, g->radius.data(), P, s);
    h2d(g->d_centre.p, g->centre.data(), 3 * (size_t)P, s);
    h2d(g->d_start.p, g->start.data(), P, s);
    SAL_HIP(hipStreamSynchronize(s));
}

static void build_tree(sal_geom g) {
    const int64_t n = g->n;
    hipStream_t s = g->stream;
    g->xi.alloc(n);
    g->eta.alloc(n);
    g->pface.alloc(n);
    DBuf<unsigned char> resolved(n);
    SAL_HIP(hipMemsetAsync(resolved.p, 0, n, s));
    DBuf<int> tie_n(1);
    DBuf<unsigned char> tie_flag(n);
    int64_t nt = 0;
    build_levels(g, resolved, tie_n, tie_flag, nt);
    if (nt > 0) {
        // re-decide near-midline points with the host libm and rebuild
        std::vector<unsigned char> flag(n);
        d2h(flag.data(), tie_flag.p, (size_t)n, s);
        SAL_HIP(hipStreamSynchronize(s));
        std::vector<int> idx;
        for (int64_t p = 0; p < n; ++p)
            if (flag[p]) idx.push_back((int)p);
        int m = (int)idx.size();
        std::vector<double> hxyz(3 * (size_t)m), vxi(m), veta(m);
        DBuf<int> d_idx(m);
        if (m > 0) {  // one gather and one copy for all flagged points
            DBuf<double> d_xyz(3 * (size_t)m);
            h2d(d_idx.p, idx.data(), m, s);
            k_gather_xyz<<<grid_for(m, 256), 256, 0, s>>>(m, d_idx.p, g->x.p, g->y.p, g->z.p, d_xyz.p);
            SAL_HIP(hipGetLastError());
            d2h(hxyz.data(), d_xyz.p, 3 * (size_t)m, s);
            SAL_HIP(hipStreamSynchronize(s));
        }
        for (int k = 0; k < m; ++k) {
            const double hx = hxyz[3 * k], hy = hxyz[3 * k + 1], hz = hxyz[3 * k + 2];
            int f = face_from_xyz(hx, hy, hz);
            xieta_from_xyz(hx, hy, hz, f, vxi[k], veta[k]);  // host atan: the reference's
        }
        DBuf<double> d_vxi(m), d_veta(m);
        h2d(d_vxi.p, vxi.data(), m, s);
        h2d(d_veta.p, veta.data(), m, s);
        k_scatter_angles<<<grid_for(m, 256), 256, 0, s>>>(m, d_idx.p, d_vxi.p, d_veta.p, g->xi.p, g->eta.p,
                                                          resolved.p);
        SAL_HIP(hipGetLastError());
        int64_t again = 0;
        build_levels(g, resolved, tie_n, tie_flag, again);
        SAL_REQUIRE(again == 0, SAL_ERR_STATE, "tree rebuild after tie resolution flagged new ties");
        g->point_ties = m;
    }
    host_panel_geometry(g);

    // DFS layout: each panel's points contiguous, children nested in child order, leaves ascending
    g->start.assign(g->P, 0);
    int64_t acc = 0;
    for (int i = 0; i < 6; ++i) {
        g->start[i] = acc;
        acc += g->count[i];
    }
    for (int i = 0; i < g->P; ++i) {
        if (g->child1[i] < 0) continue;
        int64_t a = g->start[i];
        for (int k = 0; k < 4; ++k) {
            g->start[g->child1[i] + k] = a;
            a += g->count[g->child1[i] + k];
        }
    }
    upload_panels(g);

    DBuf<int> key(n), val(n), skey(n);
    g->perm.alloc(n);
    g->iperm.alloc(n);
    k_leaf_key<<<grid_for(n, 256), 256, 0, s>>>(n, g->nlev, g->anc.p, g->d_start.p, key.p, val.p);
    SAL_HIP(hipGetLastError());
    sort_pairs_i32(key.p, skey.p, val.p, g->perm.p, n, bits_for(n), s);
    g->px.alloc(n); g->py.alloc(n); g->pz.alloc(n);
    k_layout<<<grid_for(n, 256), 256, 0, s>>>(n, g->perm.p, g->iperm.p, g->x.p, g->y.p, g->z.p, g->px.p, g->py.p,
                                              g->pz.p);
    SAL_HIP(hipGetLastError());
    SAL_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------------------ assign_points_to_panels

// Walk one point down the tree the way contains_point probes it (tree_module.f90:39-56, :257-271):
// the root of its face if xi, eta < pi/4 (half-open), then the child whose half-open box holds it.
// A point within a few ulp of a midline is flagged for the host walk.
__global__ void k_assign(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, int nlev, const int* __restrict__ child1,
                         const double* __restrict__ bounds, int* __restrict__ out, int* __restrict__ tie_n,
                         int* __restrict__ tie_list, double ulps) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int l = 0; l < nlev; ++l) out[p * nlev + l] = -1;
    int f = face_from_xyz(x[p], y[p], z[p]);
    double u, v;
    xieta_from_xyz(x[p], y[p], z[p], f, u, v);
    const double q = 0.78539816339744830962;
    if (!(u >= -q && u < q && v >= -q && v < q)) return;  // on a face's max edge: no panel holds it
    int a = f - 1, l = 0;
    bool tie = false;
    while (a >= 0 && l < nlev) {
        out[p * nlev + l] = a;
        ++l;
        int c1 = child1[a];
        if (c1 < 0) break;
        double mx = bounds[6 * (int64_t)a + 1], me = bounds[6 * (int64_t)a + 4];
        tie = tie || near_mid(u, mx, ulps) || near_mid(v, me, ulps);
        a = c1 + ((u < mx) ? ((v < me) ? 0 : 3) : ((v < me) ? 1 : 2));
    }
    if (tie) tie_list[atomicAdd(tie_n, 1)] = (int)p;
}

// ------------------------------------------------------------------------------ dual tree traversal

enum : int { C_SKIP = 0, C_EMIT = 1, C_REF_T = 2, C_REF_S = 3 };

__global__ void k_dtt_classify(int ng, const int* __restrict__ gt, const int* __restrict__ gs,
                               const int* __restrict__ count, const int* __restrict__ is_leaf,
                               const double* __restrict__ radius, const double* __restrict__ centre, double theta,
                               int thresh, int* __restrict__ code, int* __restrict__ emit, int* __restrict__ refine,
                               int* __restrict__ tie_n, int* __restrict__ tie_list, double ulps) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng) return;
    int t = gt[k], s = gs[k];
    int ct = count[t], cs = count[s];
    int c = C_SKIP, ty = 0;
    if (ct > 0 && cs > 0) {  // tree_module.f90:308
        double dot = centre[3 * t] * centre[3 * s] + centre[3 * t + 1] * centre[3 * s + 1] +
                     centre[3 * t + 2] * centre[3 * s + 2];
        dot = dot > -1.0 ? dot : -1.0;
        dot = dot < 1.0 ? dot : 1.0;
        double dist = acos(dot);
        double sep = (radius[t] + radius[s]) / dist;
        if (dist > 0 && fabs(sep - theta) <= ulps * 2.220446049250313e-16 * theta)
            tie_list[atomicAdd(tie_n, 1)] = k;
        if (dist > 0 && sep < theta) {  // :313-325
            c = C_EMIT;
            ty = (ct > thresh ? 2 : 0) + (cs > thresh ? 1 : 0);
        } else if ((ct < thresh && cs < thresh) || (is_leaf[t] && is_leaf[s])) {  // :328-339
            c = C_EMIT;
        } else if (is_leaf[t] || (!is_leaf[s] && ct < cs)) {  // :340-350, :375-385
            c = C_REF_S;
        } else {  // :351-374
            c = C_REF_T;
        }
    }
    code[k] = c | (ty << 4);
    emit[k] = (c == C_EMIT);
    refine[k] = (c == C_REF_T || c == C_REF_S);
}

__global__ void k_dtt_scatter(int ng, const int* __restrict__ gt, const int* __restrict__ gs,
                              const int* __restrict__ code, const int* __restrict__ eoff, const int* __restrict__ roff,
                              const int* __restrict__ child1, int* __restrict__ ot, int* __restrict__ os,
                              int* __restrict__ oty, int* __restrict__ nt, int* __restrict__ ns) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng) return;
    int c = code[k] & 15, ty = code[k] >> 4;
    int t = gt[k], s = gs[k];
    if (c == C_EMIT) {
        int o = eoff[k];
        ot[o] = t;
        os[o] = s;
        oty[o] = ty;
    } else if (c == C_REF_S) {
        int o = 4 * roff[k], c1 = child1[s];
        for (int j = 0; j < 4; ++j) { nt[o + j] = t; ns[o + j] = c1 + j; }
    } else if (c == C_REF_T) {
        int o = 4 * roff[k], c1 = child1[t];
        for (int j = 0; j < 4; ++j) { nt[o + j] = c1 + j; ns[o + j] = s; }
    }
}

// host re-decision of one DTT pair with the host libm (tree_module.f90:306-388)
static int host_dtt_code(sal_geom g, int t, int s, double theta) {
    int ct = g->count[t], cs = g->count[s], thr = g->list_thresh;
    if (!(ct > 0 && cs > 0)) return C_SKIP;
    const double* a = &g->centre[3 * (size_t)t];
    const double* b = &g->centre[3 * (size_t)s];
    double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    dot = dot > -1.0 ? dot : -1.0;
    dot = dot < 1.0 ? dot : 1.0;
    double dist = std::acos(dot);
    double sep = (g->radius[t] + g->radius[s]) / dist;
    if (dist > 0 && sep < theta) return C_EMIT | (((ct > thr ? 2 : 0) + (cs > thr ? 1 : 0)) << 4);
    if ((ct < thr && cs < thr) || (g->is_leaf[t] && g->is_leaf[s])) return C_EMIT;
    if (g->is_leaf[t] || (!g->is_leaf[s] && ct < cs)) return C_REF_S;
    return C_REF_T;
}

__global__ void k_patch_codes(int m, const int* __restrict__ idx, const int* __restrict__ val, int* __restrict__ code,
                              int* __restrict__ emit, int* __restrict__ refine) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int i = idx[k], c = val[k];
    code[i] = c;
    emit[i] = ((c & 15) == C_EMIT);
    refine[i] = ((c & 15) == C_REF_T || (c & 15) == C_REF_S);
}

static void run_dtt(sal_geom g, double theta) {
    hipStream_t s = g->stream;
    std::vector<int> t0(36), s0(36);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) { t0[6 * i + j] = i; s0[6 * i + j] = j; }  // tree_module.f90:293-299
    DBuf<int> gt(36), gs(36), nt, ns, code, emit, refine, eoff, roff, tie_n(1), tie_list;
    size_t lcap = 1 << 20;
    DBuf<int> lt(lcap), ls(lcap), lty(lcap);
    h2d(gt.p, t0.data(), 36, s);
    h2d(gs.p, s0.data(), 36, s);
    int64_t ng = 36, total = 0;
    g->pair_ties = 0;
    while (ng > 0) {
        SAL_REQUIRE(ng < INT_MAX / 4, SAL_ERR_LIMIT, "dual tree traversal generation too large");
        code.ensure(ng); emit.ensure(ng); refine.ensure(ng); eoff.ensure(ng); roff.ensure(ng); tie_list.ensure(ng);
        SAL_HIP(hipMemsetAsync(tie_n.p, 0, sizeof(int), s));
        k_dtt_classify<<<grid_for(ng, 256), 256, 0, s>>>((int)ng, gt.p, gs.p, g->d_count.p, g->d_is_leaf.p,
                                                         g->d_radius.p, g->d_centre.p, theta, g->list_thresh, code.p,
                                                         emit.p, refine.p, tie_n.p, tie_list.p, g_tie_ulps);
        SAL_HIP(hipGetLastError());
        int nties = 0;
        d2h(&nties, tie_n.p, 1, s);
        SAL_HIP(hipStreamSynchronize(s));
        if (nties > 0) {  // re-decide near-theta pairs with the host libm
            std::vector<int> idx(nties), ht(ng), hs(ng), val(nties);
            d2h(idx.data(), tie_list.p, nties, s);
            d2h(ht.data(), gt.p, ng, s);
            d2h(hs.data(), gs.p, ng, s);
            SAL_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < nties; ++k) val[k] = host_dtt_code(g, ht[idx[k]], hs[idx[k]], theta);
            DBuf<int> di(nties), dv(nties);
            h2d(di.p, idx.data(), nties, s);
            h2d(dv.p, val.data(), nties, s);
            k_patch_codes<<<grid_for(nties, 256), 256, 0, s>>>(nties, di.p, dv.p, code.p, emit.p, refine.p);
            SAL_HIP(hipGetLastError());
            SAL_HIP(hipStreamSynchronize(s));
            g->pair_ties += nties;
        }
        int64_t ne = exclusive_scan_i32(emit.p, eoff.p, ng, s);
        int64_t nr = exclusive_scan_i32(refine.p, roff.p, ng, s);
        if ((size_t)(total + ne) > lcap) {
            size_t nc = lcap;
            while (nc < (size_t)(total + ne)) nc *= 2;
            DBuf<int> a(nc), b(nc), c(nc);
            SAL_HIP(hipMemcpyAsync(a.p, lt.p, total * sizeof(int), hipMemcpyDeviceToDevice, s));
            SAL_HIP(hipMemcpyAsync(b.p, ls.p, total * sizeof(int), hipMemcpyDeviceToDevice, s));
            SAL_HIP(hipMemcpyAsync(c.p, lty.p, total * sizeof(int), hipMemcpyDeviceToDevice, s));
            lt = std::move(a); ls = std::move(b); lty = std::move(c);
            lcap = nc;
        }
        nt.ensure(std::max<int64_t>(4 * nr, 1));
        ns.ensure(std::max<int64_t>(4 * nr, 1));
        k_dtt_scatter<<<grid_for(ng, 256), 256, 0, s>>>((int)ng, gt.p, gs.p, code.p, eoff.p, roff.p, g->d_child1.p,
                                                        lt.p + total, ls.p + total, lty.p + total, nt.p, ns.p);
        SAL_HIP(hipGetLastError());
        SAL_HIP(hipStreamSynchronize(s));
        total += ne;
        ng = 4 * nr;
        std::swap(gt, nt);
        std::swap(gs, ns);
    }
    g->nint = total;
    g->it.alloc(std::max<int64_t>(total, 1));
    g->is.alloc(std::max<int64_t>(total, 1));
    g->ty.alloc(std::max<int64_t>(total, 1));
    SAL_HIP(hipMemcpyAsync(g->it.p, lt.p, total * sizeof(int), hipMemcpyDeviceToDevice, s));
    SAL_HIP(hipMemcpyAsync(g->is.p, ls.p, total * sizeof(int), hipMemcpyDeviceToDevice, s));
    SAL_HIP(hipMemcpyAsync(g->ty.p, lty.p, total * sizeof(int), hipMemcpyDeviceToDevice, s));
    std::vector<int> hty(total);
    d2h(hty.data(), lty.p, total, s);
    SAL_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 4; ++k) g->type_count[k] = 0;
    for (int64_t k = 0; k < total; ++k) g->type_count[hty[k] & 3]++;
}

void invalidate_plans(sal_geom g) {
    for (Plan* p : g->plans) plan_free(p);
    g->plans.clear();
}

int64_t geom_device_bytes(sal_geom g) {
    int64_t b = 0;
    b += g->x.bytes() + g->y.bytes() + g->z.bytes() + g->xi.bytes() + g->eta.bytes() + g->pface.bytes();
    b += g->anc.bytes() + g->perm.bytes() + g->iperm.bytes() + g->px.bytes() + g->py.bytes() + g->pz.bytes();
    b += g->d_level.bytes() + g->d_is_leaf.bytes() + g->d_parent.bytes() + g->d_child1.bytes() + g->d_face.bytes();
    b += g->d_count.bytes() + g->d_bounds.bytes() + g->d_radius.bytes() + g->d_centre.bytes() + g->d_start.bytes();
    b += g->it.bytes() + g->is.bytes() + g->ty.bytes();
    b += g->pc_t.bytes() + g->pc_s.bytes() + g->pp_t.bytes() + g->pp_s.bytes();
    for (const Plan* p : g->plans) b += plan_device_bytes(p);
    return b;
}

}  // namespace sal

sal_geom_s::~sal_geom_s() {
    sal::invalidate_plans(this);
    sal::stt_release(this);
    if (stream) (void)hipStreamDestroy(stream);
}

using namespace sal;

extern "C" {

const char* sal_last_error(void) { return t_error.c_str(); }

int sal_version(int* major, int* minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return SAL_OK;
}

int sal_set_tie_ulps(double ulps) {
    return guard([&] {
        SAL_REQUIRE(ulps >= 0 && ulps < 1e15, SAL_ERR_ARG, "tie window must be in [0, 1e15) ulps");
        g_tie_ulps = ulps;
    });
}

static void create_common(sal_geom g, int64_t n, int thresh) {
    SAL_REQUIRE(n >= 1 && n < INT_MAX, SAL_ERR_ARG, "point count must be in [1, 2^31)");
    SAL_REQUIRE(thresh >= 1, SAL_ERR_ARG, "cluster_thresh must be >= 1");
    g->n = n;
    g->thresh = thresh;
    SAL_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
}

int sal_geom_create(const double* x, const double* y, const double* z, int64_t n, int cluster_thresh,
                    sal_geom* out) {
    return guard([&] {
        SAL_REQUIRE(x && y && z && out, SAL_ERR_ARG, "null pointer");
        *out = nullptr;
        auto g = new sal_geom_s();
        try {
            create_common(g, n, cluster_thresh);
            g->x.alloc(n); g->y.alloc(n); g->z.alloc(n);
            h2d(g->x.p, x, n, g->stream);
            h2d(g->y.p, y, n, g->stream);
            h2d(g->z.p, z, n, g->stream);
            build_tree(g);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int sal_geom_create_dev(const double* d_x, const double* d_y, const double* d_z, int64_t n, int cluster_thresh,
                        void* stream, sal_geom* out) {
    return guard([&] {
        SAL_REQUIRE(d_x && d_y && d_z && out, SAL_ERR_ARG, "null pointer");
        *out = nullptr;
        auto g = new sal_geom_s();
        try {
            create_common(g, n, cluster_thresh);
            if (stream) SAL_HIP(hipStreamSynchronize((hipStream_t)stream));
            g->x.alloc(n); g->y.alloc(n); g->z.alloc(n);
            SAL_HIP(hipMemcpyAsync(g->x.p, d_x, n * 8, hipMemcpyDeviceToDevice, g->stream));
            SAL_HIP(hipMemcpyAsync(g->y.p, d_y, n * 8, hipMemcpyDeviceToDevice, g->stream));
            SAL_HIP(hipMemcpyAsync(g->z.p, d_z, n * 8, hipMemcpyDeviceToDevice, g->stream));
            build_tree(g);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int sal_geom_destroy(sal_geom g) {
    return guard([&] {
        if (!g) return;
        if (g->stream) SAL_HIP(hipStreamSynchronize(g->stream));
        delete g;
    });
}

int sal_geom_counts(sal_geom g, int* n_panels, int* max_level, int64_t* n_inside_total) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        if (n_panels) *n_panels = g->P;
        if (max_level) *max_level = g->nlev - 1;
        if (n_inside_total) {
            int64_t t = 0;
            for (int c : g->count) t += c;
            *n_inside_total = t;
        }
    });
}

int sal_geom_size(sal_geom g, int64_t* n_points, int* cluster_thresh) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        if (n_points) *n_points = g->n;
        if (cluster_thresh) *cluster_thresh = g->thresh;
    });
}

int sal_geom_ties(sal_geom g, int64_t* point_ties, int64_t* pair_ties) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        if (point_ties) *point_ties = g->point_ties;
        if (pair_ties) *pair_ties = g->pair_ties;
    });
}

int sal_geom_device_bytes(sal_geom g, int64_t* bytes) {
    return guard([&] {
        SAL_REQUIRE(g && bytes, SAL_ERR_ARG, "null pointer");
        *bytes = geom_device_bytes(g);
    });
}

int sal_geom_export(sal_geom g, int* level, int* is_leaf, int* parent, int* child4, int* face, double* bounds6,
                    double* radius, int* count, int64_t* offsets, int* points_inside) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        const int P = g->P;
        for (int i = 0; i < P; ++i) {
            if (level) level[i] = g->level[i];
            if (is_leaf) is_leaf[i] = g->is_leaf[i];
            if (parent) parent[i] = g->parent[i];
            if (child4)
                for (int k = 0; k < 4; ++k) child4[4 * i + k] = g->child1[i] < 0 ? -1 : g->child1[i] + k;
            if (face) face[i] = g->face[i];
            if (bounds6) std::memcpy(bounds6 + 6 * (size_t)i, &g->bounds[6 * (size_t)i], 6 * sizeof(double));
            if (radius) radius[i] = g->radius[i];
            if (count) count[i] = g->count[i];
        }
        std::vector<int64_t> off(P + 1, 0);
        for (int i = 0; i < P; ++i) off[i + 1] = off[i] + g->count[i];
        if (offsets) std::memcpy(offsets, off.data(), (P + 1) * sizeof(int64_t));
        if (!points_inside) return;
        // per level: stable sort of the points by their level-L panel gives each panel's points
        // ascending, panels of the level in id order, i.e. the reference's points_inside arrays
        const int64_t n = g->n;
        hipStream_t s = g->stream;
        DBuf<int> key(n), val(n), skey(n), sval(n);
        for (int L = 0; L < g->nlev; ++L) {
            int lo = g->level_lo[L], hi = g->level_lo[L + 1];
            int64_t m = off[hi] - off[lo];
            k_level_key<<<grid_for(n, 256), 256, 0, s>>>(n, g->anc.p + (int64_t)L * n, P, key.p, val.p);
            SAL_HIP(hipGetLastError());
            sort_pairs_i32(key.p, skey.p, val.p, sval.p, n, bits_for(P + 1), s);
            d2h(points_inside + off[lo], sval.p, (size_t)m, s);
            SAL_HIP(hipStreamSynchronize(s));
        }
    });
}

int sal_assign_points(sal_geom g, const double* x, const double* y, const double* z, int64_t n_own, int nlev,
                      int* points_panels) {
    return guard([&] {
        SAL_REQUIRE(g && x && y && z && points_panels, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(n_own >= 0 && nlev >= 1, SAL_ERR_ARG, "bad sizes");
        if (n_own == 0) return;
        hipStream_t s = g->stream;
        DBuf<double> dx(n_own), dy(n_own), dz(n_own);
        DBuf<int> out((size_t)n_own * nlev), tie_n(1), tie_list(n_own);
        h2d(dx.p, x, n_own, s);
        h2d(dy.p, y, n_own, s);
        h2d(dz.p, z, n_own, s);
        SAL_HIP(hipMemsetAsync(tie_n.p, 0, sizeof(int), s));
        k_assign<<<grid_for(n_own, 256), 256, 0, s>>>(n_own, dx.p, dy.p, dz.p, nlev, g->d_child1.p, g->d_bounds.p,
                                                      out.p, tie_n.p, tie_list.p, g_tie_ulps);
        SAL_HIP(hipGetLastError());
        d2h(points_panels, out.p, (size_t)n_own * nlev, s);
        int nt = 0;
        d2h(&nt, tie_n.p, 1, s);
        SAL_HIP(hipStreamSynchronize(s));
        if (nt > 0) {
            std::vector<int> idx(nt);
            d2h(idx.data(), tie_list.p, nt, s);
            SAL_HIP(hipStreamSynchronize(s));
            const double q = 0.78539816339744830962;
            for (int p : idx) {  // host walk with the host libm
                int* o = points_panels + (size_t)p * nlev;
                for (int l = 0; l < nlev; ++l) o[l] = -1;
                int f = face_from_xyz(x[p], y[p], z[p]);
                double u, v;
                xieta_from_xyz(x[p], y[p], z[p], f, u, v);
                if (!(u >= -q && u < q && v >= -q && v < q)) continue;
                int a = f - 1, l = 0;
                while (a >= 0 && l < nlev) {
                    o[l++] = a;
                    int c1 = g->child1[a];
                    if (c1 < 0) break;
                    double mx = g->bounds[6 * (size_t)a + 1], me = g->bounds[6 * (size_t)a + 4];
                    a = c1 + ((u < mx) ? ((v < me) ? 0 : 3) : ((v < me) ? 1 : 2));
                }
            }
        }
    });
}

int sal_dtt(sal_geom g, double theta, int cluster_thresh, int64_t* n_interactions) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        SAL_REQUIRE(theta > 0 && std::isfinite(theta), SAL_ERR_ARG, "theta must be positive");
        g->list_thresh = cluster_thresh > 0 ? cluster_thresh : g->thresh;
        invalidate_plans(g);
        run_dtt(g, theta);
        if (n_interactions) *n_interactions = g->nint;
    });
}

int sal_dtt_counts(sal_geom g, int64_t counts4[4]) {
    return guard([&] {
        SAL_REQUIRE(g && counts4, SAL_ERR_ARG, "null pointer");
        for (int k = 0; k < 4; ++k) counts4[k] = g->type_count[k];
    });
}

int sal_dtt_export(sal_geom g, int* target, int* source, int* type) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null handle");
        if (target) d2h(target, g->it.p, g->nint, g->stream);
        if (source) d2h(source, g->is.p, g->nint, g->stream);
        if (type) d2h(type, g->ty.p, g->nint, g->stream);
        SAL_HIP(hipStreamSynchronize(g->stream));
    });
}

int sal_dtt_import(sal_geom g, const int* target, const int* source, const int* type, int64_t n) {
    return guard([&] {
        SAL_REQUIRE(g && (n == 0 || (target && source && type)), SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(n >= 0 && n < INT_MAX, SAL_ERR_ARG, "bad list length");
        for (int64_t k = 0; k < n; ++k) {
            SAL_REQUIRE(target[k] >= 0 && target[k] < g->P && source[k] >= 0 && source[k] < g->P &&
                            type[k] >= 0 && type[k] <= 3,
                        SAL_ERR_ARG, "interaction list entry out of range");
        }
        invalidate_plans(g);
        g->nint = n;
        g->it.alloc(std::max<int64_t>(n, 1));
        g->is.alloc(std::max<int64_t>(n, 1));
        g->ty.alloc(std::max<int64_t>(n, 1));
        h2d(g->it.p, target, n, g->stream);
        h2d(g->is.p, source, n, g->stream);
        h2d(g->ty.p, type, n, g->stream);
        for (int k = 0; k < 4; ++k) g->type_count[k] = 0;
        for (int64_t k = 0; k < n; ++k) g->type_count[type[k]]++;
        SAL_HIP(hipStreamSynchronize(g->stream));
    });
}

}  // extern "C"
