// FP64 evaluation of the dual-tree interaction list (replaces fast_sum, fast_sum_module.f90:291-332)
// and the FP64 direct sum used as the accuracy oracle.
//
// The reference evaluates the list entry by entry and redoes the source anterpolation (P2M) and the
// target interpolation for every entry (:99-111, :186-199, :229-241, :274-287).  Those are linear in
// the sources / in the proxy values, so here they are done once per panel:
//
//   1. proxies   z_k of every panel holding more than cluster_thresh points (the only panels that
//                act as clusters, :317-322), xyz_from_xieta of the Chebyshev grid;
//   2. P2M       W_S[k] = sum_{s in S} l_k(s) q_s, q = area * ssh, for those panels;
//   3. proxy     Phi_T[j] = sum over T's CP entries sum_{s in S} K(z_j, s) q_s
//      targets           + sum over T's CC entries sum_k K(z_j, y_k) W_S[k];
//   4. leaves    for each leaf and each panel A on its root path: A's PP entries (direct sums), A's
//                PC entries (proxies of the source with W), and L2P of Phi_A at the leaf's points.
//
// Every target point is written by exactly one thread (owner computes), with no atomics: results
// are deterministic.  Points live in the tree's DFS order so every panel's points are one
// contiguous range; the sums are gathered from / scattered to the caller's order with the tree's
// permutation.
//
// SAL_MODE_REFERENCE keeps the reference's semantics (projected east/north components at the
// proxies, self pair included, kernel expression order); SAL_MODE_ACCURATE uses the
// cancellation-free kernel, skips the self pair (its value is exactly 0 there) and interpolates the
// tangential Cartesian vector at cluster targets.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>

#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------------------------ plan

// uv[k] = (xi, eta) of the point at tree position k
__global__ void k_gather_uv(int64_t n, const int* __restrict__ perm, const double* __restrict__ xi,
                            const double* __restrict__ eta, double2* __restrict__ uv) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = perm[k];
    uv[k] = make_double2(xi[p], eta[p]);
}

void plan_free(Plan* p) { delete p; }

// PP entries between panels of at most PAIR_MAX points are summed by k_pair64: the kernel factor of
// the reference's expression depends on the pair only through g = t.s, and t.s == s.t bit for bit
// (the same products added in the same order), so one evaluation serves the entry A<-B (targets in
// A) and its mirror B<-A (targets in B).  The DTT list is symmetric up to traversal ties: an entry
// whose mirror is not a PP entry of the list gets a one-directional task.  Each entry's partial sums
// (one double2 per point of its target panel) go to pp_part at pp_off[position]; the leaf kernel
// adds them in list order.
constexpr int PAIR_MAX = 128;

// Mirrored entries matched in list order: keyed = (target << 32 | source, position) sorted; for each
// unused entry the first unused entry with the mirrored key in `mirror` (the same vector for PP and
// CC, the PC entries for a CP entry), or -1.
static int take_mirror(const std::vector<std::pair<uint64_t, int>>& mirror, std::vector<char>& used_m, int A, int B) {
    const uint64_t mk = ((uint64_t)B << 32) | (uint32_t)A;
    auto it = std::lower_bound(mirror.begin(), mirror.end(), std::make_pair(mk, INT32_MIN));
    while (it != mirror.end() && it->first == mk && used_m[it->second]) ++it;
    if (it == mirror.end() || it->first != mk) return -1;
    used_m[it->second] = 1;
    return it->second;
}

static void build_pair_tasks(sal_geom g, Plan* pl, const std::vector<int>& pt_off, const std::vector<int>& pt_ent,
                             const std::vector<int>& xt_off, const std::vector<int>& xt_ent, hipStream_t s) {
    const int P = g->P, m = pl->m;
    const int64_t ne = pt_off[P], nx = xt_off[P];
    auto small = [&](int A) { return g->count[A] > 0 && g->count[A] <= PAIR_MAX; };
    std::vector<std::pair<uint64_t, int>> kpp, kpc, kcp, kcc;  // (target << 32 | source, position)
    for (int A = 0; A < P; ++A) {
        for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
            const int B = pt_ent[e] >> 2, ty = pt_ent[e] & 3;
            const uint64_t key = ((uint64_t)A << 32) | (uint32_t)B;
            if (ty == 0 && small(A) && small(B)) kpp.push_back({key, e});
            if (ty == 1 && small(A) && m <= PAIR_MAX) kpc.push_back({key, e});
        }
        for (int e = xt_off[A]; e < xt_off[A + 1]; ++e) {
            const int B = xt_ent[e] >> 2, ty = xt_ent[e] & 3;
            const uint64_t key = ((uint64_t)A << 32) | (uint32_t)B;
            if (m > PAIR_MAX) continue;
            if (ty == 2 && small(B)) kcp.push_back({key, e});
            if (ty == 3) kcc.push_back({key, e});
        }
    }
    for (auto* v : {&kpp, &kpc, &kcp, &kcc}) std::sort(v->begin(), v->end());
    std::vector<int4> tasks, far;
    std::vector<char> upp(std::max<int64_t>(ne, 1), 0), upc(std::max<int64_t>(ne, 1), 0),
        ux(std::max<int64_t>(nx, 1), 0);
    std::vector<int64_t> off(std::max<int64_t>(ne, 1), -1), xoff(std::max<int64_t>(nx, 1), -1);
    int64_t total = 0, xtotal = 0;
    for (const auto& k : kpp) {  // every eligible PP entry: offsets in list order below
        const int A = (int)(k.first >> 32), B = (int)(k.first & 0xffffffffu), e = k.second;
        if (upp[e]) continue;
        upp[e] = 1;
        if (A == B) { tasks.push_back(make_int4(A, A, e, e)); continue; }
        tasks.push_back(make_int4(A, B, e, take_mirror(kpp, upp, A, B)));
    }
    for (const auto& k : kcc) {  // CC entries with their CC mirror
        const int A = (int)(k.first >> 32), B = (int)(k.first & 0xffffffffu), e = k.second;
        if (ux[e] || A == B) continue;
        const int f = take_mirror(kcc, ux, A, B);
        if (f < 0) continue;
        ux[e] = 1;
        far.push_back(make_int4(A, B, e, f));
    }
    for (const auto& k : kcp) {  // CP entries B <- A (B's proxies, A's points) with their PC mirror A <- B
        const int B = (int)(k.first >> 32), A = (int)(k.first & 0xffffffffu), e = k.second;
        const int f = take_mirror(kpc, upc, B, A);
        if (f < 0) continue;
        ux[e] = 1;
        far.push_back(make_int4(A | (1 << 30), B, f, e));
    }
    for (int A = 0; A < P; ++A) {  // partial-sum slots in list order
        for (int e = pt_off[A]; e < pt_off[A + 1]; ++e)
            if (upp[e] || upc[e]) { off[e] = total; total += g->count[A]; }
        for (int e = xt_off[A]; e < xt_off[A + 1]; ++e)
            if (ux[e]) { xoff[e] = xtotal; xtotal += m; }
    }
    // tasks in target order keep neighbouring panels together (the launch deals chunks of them to XCDs)
    auto by_target = [](const int4& a, const int4& b) { return (a.x & 0x3fffffff) < (b.x & 0x3fffffff); };
    std::stable_sort(tasks.begin(), tasks.end(), by_target);
    std::stable_sort(far.begin(), far.end(), by_target);
    pl->n_pair_tasks = (int)tasks.size();
    pl->n_far_tasks = (int)far.size();
    if (tasks.empty() && far.empty()) return;
    if (!tasks.empty()) {
        pl->pair_task.alloc(tasks.size());
        h2d(pl->pair_task.p, tasks.data(), tasks.size(), s);
    }
    if (!far.empty()) {
        pl->far_task.alloc(far.size());
        h2d(pl->far_task.p, far.data(), far.size(), s);
    }
    pl->pp_off.alloc(off.size());
    h2d(pl->pp_off.p, off.data(), off.size(), s);
    pl->pp_part.alloc(std::max<int64_t>(total, 1));
    pl->xp_off.alloc(xoff.size());
    h2d(pl->xp_off.p, xoff.data(), xoff.size(), s);
    pl->xp_part.alloc(std::max<int64_t>(xtotal, 1));
}

int64_t plan_device_bytes(const Plan* p) {
    int64_t b = p->pt_off.bytes() + p->pt_ent.bytes() + p->xt_off.bytes() + p->xt_ent.bytes() + p->slot.bytes() +
                p->slot_panel.bytes() + p->xt_panels.bytes() + p->leaves.bytes() + p->ck_slot.bytes() +
                p->ck_first.bytes() + p->ck_off.bytes() + p->prox.bytes() + p->uv.bytes() + p->pair_task.bytes() +
                p->pp_off.bytes() + p->pp_part.bytes() + p->far_task.bytes() + p->xp_off.bytes() +
                p->xp_part.bytes() + p->q.bytes() + p->W.bytes() +
                p->wpart.bytes() + p->phi.bytes();
    if (p->mixed && p->mixed_bytes) b += p->mixed_bytes(p->mixed);
    if (p->work) b += work_held((sal_work)p->work);
    return b;
}

Plan* get_plan(sal_geom g, int degree, int kind, int near64) {
    std::lock_guard<std::mutex> lk(g->mu);
    for (Plan* p : g->plans)
        if (p->degree == degree && p->kind == kind && p->near64 == near64) return p;
    SAL_REQUIRE(g->nint > 0, SAL_ERR_STATE, "no interaction list: call sal_dtt (or sal_dtt_import) first");
    hipStream_t s = g->stream;
    std::vector<int> it(g->nint), is(g->nint), ty(g->nint);
    d2h(it.data(), g->it.p, g->nint, s);
    d2h(is.data(), g->is.p, g->nint, s);
    d2h(ty.data(), g->ty.p, g->nint, s);
    SAL_HIP(hipStreamSynchronize(s));
    Plan* pl = build_plan(g, degree, kind, it, is, ty);
    pl->near64 = near64;
    g->plans.push_back(pl);
    return pl;
}

Plan* build_plan(sal_geom g, int degree, int kind, const std::vector<int>& it, const std::vector<int>& is,
                 const std::vector<int>& ty) {
    hipStream_t s = g->stream;
    const int64_t nint = (int64_t)it.size();
    static std::atomic<uint64_t> next_serial{1};
    Plan* pl = new Plan();
    pl->serial = next_serial++;
    try {
        pl->degree = degree;
        pl->kind = kind;
        const int nd = degree + 1, m = nd * nd, P = g->P;
        pl->nd = nd;
        pl->m = m;
        // CSR by target for point targets (PP, PC) and proxy targets (CP, CC)
        std::vector<int> pt_off(P + 1, 0), xt_off(P + 1, 0);
        for (int64_t e = 0; e < nint; ++e) {
            if (ty[e] <= 1) pt_off[it[e] + 1]++;
            else xt_off[it[e] + 1]++;
        }
        for (int i = 0; i < P; ++i) {
            pt_off[i + 1] += pt_off[i];
            xt_off[i + 1] += xt_off[i];
        }
        std::vector<int> pt_ent(pt_off[P]), xt_ent(xt_off[P]);
        {
            std::vector<int> a(pt_off.begin(), pt_off.end() - 1), b(xt_off.begin(), xt_off.end() - 1);
            for (int64_t e = 0; e < nint; ++e) {
                int v = (is[e] << 2) | ty[e];  // source id and type packed
                if (ty[e] <= 1) pt_ent[a[it[e]]++] = v;
                else xt_ent[b[it[e]]++] = v;
            }
        }
        // proxy slots: every panel the list uses as a cluster (PC / CC source, CP / CC target)
        std::vector<int> slot(P, -1);
        {
            std::vector<char> used(P, 0);
            for (int64_t e = 0; e < nint; ++e) {
                if (ty[e] == 1 || ty[e] == 3) used[is[e]] = 1;
                if (ty[e] >= 2) used[it[e]] = 1;
            }
            int Q = 0;
            for (int i = 0; i < P; ++i)
                if (used[i]) slot[i] = Q++;
            pl->Q = Q;
        }
        const int Q = pl->Q;
        std::vector<int> xt_panels, leaves;
        for (int i = 0; i < P; ++i) {
            if (xt_off[i + 1] > xt_off[i]) xt_panels.push_back(i);
            if (g->is_leaf[i] && g->count[i] > 0) leaves.push_back(i);
        }
        // P2M chunks: (slot, first point) pairs of at most P2M_CHUNK points
        std::vector<int> ck_slot, ck_first, ck_off(Q + 1, 0);
        for (int i = 0; i < P; ++i) {
            if (slot[i] < 0) continue;
            for (int f = 0; f < g->count[i]; f += P2M_CHUNK) {
                ck_slot.push_back(i);
                ck_first.push_back(f);
            }
            ck_off[slot[i] + 1] = (int)ck_slot.size();
        }
        pl->n_chunks = (int)ck_slot.size();
        pl->n_xt = (int)xt_panels.size();
        pl->n_leaves = (int)leaves.size();
        auto up = [&](DBuf<int>& d, const std::vector<int>& h) {
            d.alloc(std::max<size_t>(h.size(), 1));
            h2d(d.p, h.data(), h.size(), s);
        };
        up(pl->pt_off, pt_off); up(pl->pt_ent, pt_ent); up(pl->xt_off, xt_off); up(pl->xt_ent, xt_ent);
        up(pl->slot, slot); up(pl->xt_panels, xt_panels); up(pl->leaves, leaves);
        up(pl->ck_slot, ck_slot); up(pl->ck_first, ck_first); up(pl->ck_off, ck_off);
        std::vector<int> slot_panel(std::max(Q, 1));
        for (int i = 0; i < P; ++i)
            if (slot[i] >= 0) slot_panel[slot[i]] = i;
        up(pl->slot_panel, slot_panel);
        // proxy points of every slot
        pl->prox.alloc((size_t)std::max(Q, 1) * m * 3);
        build_proxies(g, pl, s);
        if (kind != PLAN_MIXED) {  // the FP64 kernels' P2M and L2P read the face angles in tree order
            pl->uv.alloc(std::max<int64_t>(g->n, 1));
            k_gather_uv<<<grid_for(g->n, 256), 256, 0, s>>>(g->n, g->perm.p, g->xi.p, g->eta.p, pl->uv.p);
            SAL_HIP(hipGetLastError());
        }
        if (kind == PLAN_FP64_REF) build_pair_tasks(g, pl, pt_off, pt_ent, xt_off, xt_ent, s);
        // per-leaf list of ancestors with work, for statistics
        pl->pt_entries = pt_off[P];
        pl->xt_entries = xt_off[P];
        SAL_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete pl;
        throw;
    }
    return pl;
}

__global__ void k_proxies(int Q, int nd, const int* __restrict__ slot_panel, const double* __restrict__ bounds,
                          const int* __restrict__ face, double* __restrict__ prox) {
    int q = blockIdx.x;
    int m = nd * nd;
    int A = slot_panel[q];
    const double* b = bounds + 6 * (int64_t)A;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        int i = k % nd, j = k / nd;
        double xi = bli_node(b[0], b[2], nd - 1, i), eta = bli_node(b[3], b[5], nd - 1, j);
        double x, y, z;
        xyz_from_xieta(xi, eta, face[A], x, y, z);
        double* o = prox + ((int64_t)q * m + k) * 3;
        o[0] = x;
        o[1] = y;
        o[2] = z;
    }
}

void build_proxies(sal_geom g, Plan* pl, hipStream_t s) {
    if (pl->Q == 0) return;
    k_proxies<<<pl->Q, 128, 0, s>>>(pl->Q, pl->nd, pl->slot_panel.p, g->d_bounds.p, g->d_face.p, pl->prox.p);
    SAL_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------ kernels

// q in DFS order: q[k] = area[perm[k]] * ssh[perm[k]]
__global__ void k_gather_q(int64_t n, const int* __restrict__ perm, const double* __restrict__ area,
                           const double* __restrict__ ssh, double* __restrict__ q) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = perm[k];
    q[k] = area[p] * ssh[p];
}

// P2M partial weights of one chunk of a slot's points; 128 threads, 64 points staged at a time.
// The panel's Chebyshev nodes are computed once per block; thread t < 64 forms the xi basis of
// point t of a step times its charge and thread 64 + t its eta basis, then each weight's thread
// sums over the step (two LDS reads and one FMA per point).
__global__ __launch_bounds__(128) void k_p2m64(int nd, const int* __restrict__ ck_slot, const int* __restrict__ ck_first,
                                               const int64_t* __restrict__ start, const int* __restrict__ count,
                                               const double* __restrict__ bounds, const double2* __restrict__ uv,
                                               const double* __restrict__ q, double* __restrict__ wpart) {
    __shared__ double fb[2][64][MAX_DEG + 1], nodes[2][MAX_DEG + 1];
    const int A = ck_slot[blockIdx.x], f0 = ck_first[blockIdx.x];
    const int m = nd * nd, deg = nd - 1;
    const int cnt = min(P2M_CHUNK, count[A] - f0);
    const int64_t st = start[A] + f0;
    const double* b = bounds + 6 * (int64_t)A;
    const int tid = threadIdx.x, d = tid >> 6, tp = tid & 63;
    if (tid < 2 * nd) {
        const int dd = tid / nd, i = tid % nd;
        nodes[dd][i] = bli_node(b[3 * dd], b[3 * dd + 2], deg, i);
    }
    __syncthreads();
    double acc[3] = {0.0, 0.0, 0.0};  // m <= 3 * 128 weights per block
    for (int c0 = 0; c0 < cnt; c0 += 64) {
        int nc = min(64, cnt - c0);
        if (tp < nc) {
            const int64_t k = st + c0 + tp;
            bli_basis_1d(d ? uv[k].y : uv[k].x, nodes[d], deg, fb[d][tp]);
            if (d == 0) {
                const double qk = q[k];
                for (int i = 0; i <= deg; ++i) fb[0][tp][i] *= qk;
            }
        }
        __syncthreads();
        for (int r = 0; r < 3; ++r) {
            int w = tid + 128 * r;
            if (w < m) {
                int i = w % nd, j = w / nd;
                double a = acc[r];
                for (int t = 0; t < nc; ++t) a = fma(fb[0][t][i], fb[1][t][j], a);
                acc[r] = a;
            }
        }
        __syncthreads();
    }
    for (int r = 0; r < 3; ++r) {
        int w = tid + 128 * r;
        if (w < m) wpart[(int64_t)blockIdx.x * m + w] = acc[r];
    }
}

__global__ void k_p2m_reduce(int Q, int m, const int* __restrict__ ck_off, const double* __restrict__ wpart,
                             double* __restrict__ W) {
    int q = blockIdx.x;
    for (int w = threadIdx.x; w < m; w += blockDim.x) {
        double a = 0.0;
        for (int c = ck_off[q]; c < ck_off[q + 1]; ++c) a += wpart[(int64_t)c * m + w];
        W[(int64_t)q * m + w] = a;
    }
}

// Proxy targets: one 256-thread block per target panel.  A pass over mm <= 256 proxies runs
// k = 256 / mm threads per proxy (three at d = 8, where m = 81): thread i takes proxy i % mm and every
// k-th source from i / mm on, and the k partial sums are added in thread order through LDS.
template <int MODE>
__global__ __launch_bounds__(256) void k_proxy64(int m, int nxt, const int* __restrict__ xt_panels, const int* __restrict__ xt_off,
                                                 const int* __restrict__ xt_ent, const int* __restrict__ slot,
                                                 const int64_t* __restrict__ start, const int* __restrict__ count,
                                                 const double* __restrict__ px, const double* __restrict__ py,
                                                 const double* __restrict__ pz, const double* __restrict__ q,
                                                 const double* __restrict__ prox, const double* __restrict__ W,
                                                 double* __restrict__ phi, const int64_t* __restrict__ xp_off,
                                                 const double2* __restrict__ xp_part) {
    constexpr int NC = (MODE == SAL_MODE_REFERENCE) ? 2 : 3;
    constexpr int CH = 256, NT = 256;
    __shared__ double sx[CH], sy[CH], sz[CH], sq[CH];
    __shared__ double red[NC][NT];
    const int item = xcd_item(nxt);
    if (item < 0) return;
    const int T = xt_panels[item];
    const int tq = slot[T];
    const int tid = threadIdx.x;
    for (int j0 = 0; j0 < m; j0 += NT) {
        const int mm = min(NT, m - j0), k = NT / mm, sub = tid / mm;
        const int j = j0 + tid % mm;
        const bool act = sub < k;
        double zx, zy, zz;
        {
            const double* z = prox + ((int64_t)tq * m + j) * 3;
            zx = z[0]; zy = z[1]; zz = z[2];
        }
        Tgt tz = make_tgt(zx, zy, zz);
        double acc[NC];
        for (int c = 0; c < NC; ++c) acc[c] = 0.0;
        for (int e = xt_off[T]; e < xt_off[T + 1]; ++e) {
            if (xp_off && xp_off[e] >= 0) continue;  // summed by k_farpair64, added below
            const int S = xt_ent[e] >> 2, ty = xt_ent[e] & 3;
            int ns;
            int64_t s0 = 0;
            const double* sprox = nullptr;
            const double* sw = nullptr;
            if (ty == 2) {  // CP: source points
                ns = count[S];
                s0 = start[S];
            } else {  // CC: source proxies and weights
                ns = m;
                sprox = prox + (int64_t)slot[S] * m * 3;
                sw = W + (int64_t)slot[S] * m;
            }
            for (int c0 = 0; c0 < ns; c0 += CH) {
                int nc = min(CH, ns - c0);
                __syncthreads();
                for (int t = tid; t < nc; t += NT) {
                    if (ty == 2) {
                        int64_t kk = s0 + c0 + t;
                        sx[t] = px[kk]; sy[t] = py[kk]; sz[t] = pz[kk]; sq[t] = q[kk];
                    } else {
                        const double* yk = sprox + (int64_t)(c0 + t) * 3;
                        sx[t] = yk[0]; sy[t] = yk[1]; sz[t] = yk[2]; sq[t] = sw[c0 + t];
                    }
                }
                __syncthreads();
                if (act) {
#pragma unroll 2
                    for (int t = sub; t < nc; t += k) {
                        if constexpr (MODE == SAL_MODE_REFERENCE) {  // acc scaled by tz.kc below
                            double pe, pn;
                            // CP, CC: the factor over KP1 (KP1 goes into the final scaling)
                            const double f = gfunc_ref_far_u(tz, sx[t], sy[t], sz[t], pe, pn) * sq[t];
                            acc[0] = fma(pe, f, acc[0]);
                            acc[1] = fma(pn, f, acc[1]);
                        } else {
                            double dx = sx[t] - zx, dy = sy[t] - zy, dz = sz[t] - zz;
                            double v = val_stable(dx * dx + dy * dy + dz * dz) * sq[t];
                            double r = dx * zx + dy * zy + dz * zz;  // tangential part P_z (s - z)
                            acc[0] = fma(v, dx - r * zx, acc[0]);
                            acc[1] = fma(v, dy - r * zy, acc[1]);
                            acc[2] = fma(v, dz - r * zz, acc[2]);
                        }
                    }
                }
            }
        }
        if (MODE == SAL_MODE_REFERENCE) {
            acc[0] *= KP1;
            acc[1] *= KP1;
            if (xp_off && act && sub == 0)  // entries k_farpair64 summed (KP1 included), in list order
                for (int e = xt_off[T]; e < xt_off[T + 1]; ++e) {
                    const int64_t po = xp_off[e];
                    if (po < 0) continue;
                    const double2 v = xp_part[po + j];
                    acc[0] += v.x;
                    acc[1] += v.y;
                }
            acc[0] *= tz.kc;
            acc[1] *= tz.kc;
        }
        __syncthreads();
        for (int c = 0; c < NC; ++c) red[c][tid] = act ? acc[c] : 0.0;
        __syncthreads();
        if (tid < mm)
            for (int c = 0; c < NC; ++c) {
                double a = 0.0;
                for (int i = 0; i < k; ++i) a += red[c][tid + i * mm];
                phi[((int64_t)tq * NC + c) * m + j] = a;
            }
    }
}

// L2P of one expansion at one point: acc[c] = sum_j sum_i l_i(u) l_j(v) ph[c m + i + j nd] in that
// order (fast_sum_module.f90:186-199).  l2p_reg keeps both bases in registers for nd <= ND (fully
// unrolled loops guarded by i < nd, the same operations in the same order as bli_basis_1d);
// l2p_mem is the general form with the bases in per-lane memory.
constexpr int L2P_REG_ND = 9;
template <int NM>
__device__ __forceinline__ void basis_reg(double x, const double* nodes, int nd, double (&f)[NM]) {
    bool found = false;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        f[i] = 0.0;
        if (i < nd && fabs(x - nodes[i]) < 1.0e-16) found = true;
    }
    if (found) {
#pragma unroll
        for (int i = 0; i < NM; ++i) f[i] = (i < nd && fabs(x - nodes[i]) < 1.0e-16) ? 1.0 : 0.0;
        return;
    }
    double den = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i < nd) {
            const double v = bli_coeff(nd - 1, i) / (x - nodes[i]);
            f[i] = v;
            den = den + v;
        }
    const double inv = 1.0 / den;
#pragma unroll
    for (int i = 0; i < NM; ++i) f[i] = f[i] * inv;
}
template <int NC, int NM>
__device__ __forceinline__ void l2p_reg(double u, double v, const double* nx, const double* ny, int nd,
                                        const double* __restrict__ ph, double (&acc)[NC]) {
    double fx[NM], fy[NM];
    basis_reg<NM>(u, nx, nd, fx);
    basis_reg<NM>(v, ny, nd, fy);
    const int m = nd * nd;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) acc[cc] = 0.0;
#pragma unroll
    for (int j = 0; j < NM; ++j)
        if (j < nd)
#pragma unroll
            for (int i = 0; i < NM; ++i)
                if (i < nd) {
                    const double bv = fx[i] * fy[j];
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc) acc[cc] = fma(bv, ph[cc * m + i + j * nd], acc[cc]);
                }
}
template <int NC>
__device__ __forceinline__ void l2p_mem(double u, double v, const double* nx, const double* ny, int nd,
                                        const double* __restrict__ ph, double (&acc)[NC]) {
    double fx[MAX_DEG + 1], fy[MAX_DEG + 1];
    bli_basis_1d(u, nx, nd - 1, fx);
    bli_basis_1d(v, ny, nd - 1, fy);
    const int m = nd * nd;
    for (int cc = 0; cc < NC; ++cc) acc[cc] = 0.0;
    for (int j = 0; j < nd; ++j)
        for (int i = 0; i < nd; ++i) {
            const double bv = fx[i] * fy[j];
            for (int cc = 0; cc < NC; ++cc) acc[cc] = fma(bv, ph[cc * m + i + j * nd], acc[cc]);
        }
}

// One pair (t, s) of a PP task in the reference's expression (gfunc_ref_u's operations, order and
// rounding): t = (tx, ty, tz) with ta = 1 - tz^2 and weight tq on this lane, s = point si of the
// staged panel.  Adds the term for target t (source s) to (ax, ay), and with `both` the term for
// target s (source t) to the staged panel's sums (bx, by)[si]; the lanes of one call name distinct
// si, and the LDS adds of a wave take effect in issue order, so every sum has a fixed order.  With
// DISTINCT (t and s in different panels, so distinct points: equal points fall in one leaf) g stays
// below 1 and mp above 0, so the upper clamp and the root's 1e-300 guard are left out (the same
// bits); the lower clamp stays (small root panels of a tiny grid can pair antipodal points).
template <bool DISTINCT = false>
__device__ __forceinline__ void pair_ref(double tx, double ty, double tz, double ta, double tq, int si,
                                         const double* sx, const double* sy, const double* sz, const double* sq,
                                         double* bx, double* by, bool both, double& ax, double& ay, int grp = 0,
                                         int ngrp = 1) {
    const double ux = sx[si], uy = sy[si], uz = sz[si];
    const double dxy = tx * ux + ty * uy;  // == ux * tx + uy * ty: the same products, the same sum
    double gg = dxy + tz * uz;
    if (!DISTINCT) gg = gg < 1.0 ? gg : 1.0;
    gg = gg > -1.0 ? gg : -1.0;
    const double mp = fma(-2.0, gg, 2.0);
    const double sqp = sqrt_pos(DISTINCT ? mp : mp + 1e-300);
    const double d1 = fma(sqp, mp, 1e-16);
    const double d2 = fma(2.0, gg * gg - 1.0, 1e-16);
    const double num = fma(KP1, d2, KP2 * fma(2.0, gg, sqp) * d1);
    const double f = num * rcp_nr(d1 * d2);
    const double fs = f * sq[si];
    ax = fma(tx * uy - ty * ux, fs, ax);           // pe(t <- s)
    ay = fma(uz * ta - tz * dxy, fs, ay);          // pn(t <- s)
    if (both) {
        const double ua = 1.0 - uz * uz;
        const double ft = f * tq;
        const double vx = (ux * ty - uy * tx) * ft;  // pe(s <- t)
        const double vy = (tz * ua - uz * dxy) * ft;  // pn(s <- t)
        // lanes that may name the same staged point add in separate instructions, group by group
        if (ngrp == 1) {
            atomicAdd(&bx[si], vx);
            atomicAdd(&by[si], vy);
        } else {
            for (int gi = 0; gi < ngrp; ++gi)
                if (grp == gi) {
                    atomicAdd(&bx[si], vx);
                    atomicAdd(&by[si], vy);
                }
        }
    }
}

// Wave steps of a two-way task with cl points on the lanes (chunks of 64) and cs staged, weighted by
// the LDS add groups a step needs: ks sub-lanes per point while ks na <= cs (no two lanes name the
// same staged point); a chunk of more points than cs has ceil(na / cs) groups.
__device__ __forceinline__ int pair_cost(int cl, int cs) {
    int c = 0;
    for (int a0 = 0; a0 < cl; a0 += 64) {
        const int na = min(64, cl - a0);
        int ks = 1;
        while (2 * ks * na <= 64 && 2 * ks * na <= cs) ks *= 2;
        const int len = cs - (ks - 1) * (cs / ks), ng = (na + cs - 1) / cs;
        c += len * (48 + 4 * ng);
    }
    return c;
}

// PP tasks of a reference-semantics plan (see build_pair_tasks), one wave each.  A two-way task puts
// the points of the smaller panel on the lanes (chunks of 64) and stages the other panel's points in
// LDS; at step u lane i pairs its point with staged point (i + u) mod cS (lanes name distinct staged
// points, which accumulate in LDS), and when a chunk has few points, ks sub-lanes per point split
// the steps into ranges starting h = cS / ks apart (still distinct staged points: ks na <= cS).  A
// one-way task (mirror missing) sums the lane panel's targets only.  A panel's task with itself
// pairs every point with every other once (steps u = 1 .. c/2 of the rotation, the last one by half
// the lanes when c is even) plus the reference's self pair (u = 0); above 64 points it runs as the
// blocks [0, 64) x [0, 64), [64, c) x [64, c) and [64, c) x [0, 64).  The sums are not scaled by the
// target's factor (the leaf kernel scales a point's total once).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
void k_pair64(int ntask, const int4* __restrict__ tasks, const int64_t* __restrict__ start,
              const int* __restrict__ count, const double* __restrict__ px, const double* __restrict__ py,
              const double* __restrict__ pz, const double* __restrict__ q, const int64_t* __restrict__ pp_off,
              double2* __restrict__ part) {
    __shared__ double sx[PAIR_MAX], sy[PAIR_MAX], sz[PAIR_MAX], sq[PAIR_MAX], bx[PAIR_MAX], by[PAIR_MAX];
    __shared__ double rx[64], ry[64];
    const int item = xcd_item(ntask);
    if (item < 0) return;
    const int4 tk = tasks[item];
    const int lane = threadIdx.x;
    const bool self = tk.x == tk.y, two = tk.w >= 0;
    // lanes on the panel that takes fewer wave steps
    const bool swap = two && !self && pair_cost(count[tk.y], count[tk.x]) < pair_cost(count[tk.x], count[tk.y]);
    const int Lp = swap ? tk.y : tk.x, Sp = swap ? tk.x : tk.y;
    const int64_t offL = pp_off[swap ? tk.w : tk.z], offS = two ? pp_off[swap ? tk.z : tk.w] : -1;
    const int cL = count[Lp], cS = count[Sp];
    const int64_t sL = start[Lp], sS = start[Sp];
    for (int i = lane; i < cS; i += 64) {
        const int64_t k = sS + i;
        sx[i] = px[k]; sy[i] = py[k]; sz[i] = pz[k]; sq[i] = q[k];
        bx[i] = 0.0; by[i] = 0.0;
    }
    __syncthreads();
    if (!self) {
        for (int a0 = 0; a0 < cL; a0 += 64) {
            const int na = min(64, cL - a0);
            int ks = 1;
            while (2 * ks * na <= 64 && (two ? 2 * ks * na : 2 * ks) <= cS) ks *= 2;
            const int h = cS / ks, j = lane / na, i = lane % na;
            // more lane points than staged ones (ks = 1): lanes i and i + cS name the same staged point
            const int ngrp = two ? (na + cS - 1) / cS : 1, grp = two ? i / cS : 0;
            const bool act = j < ks;
            const int64_t kt = sL + a0 + (act ? i : 0);
            const double tx = px[kt], ty = py[kt], tz = pz[kt], tq = q[kt];
            const double ta = 1.0 - tz * tz;
            const int len = cS - (ks - 1) * h;             // the last sub-lane's steps (the others: h)
            const int mylen = act ? (j == ks - 1 ? len : h) : 0;
            double ax = 0.0, ay = 0.0;
            int si = (two ? i % cS : 0) + j * h;  // i % cS < cS: one wrap at most
            for (int u = 0; u < len; ++u, ++si) {
                if (u < mylen) {
                    const int sj = si >= cS ? si - cS : si;
                    pair_ref<true>(tx, ty, tz, ta, tq, sj, sx, sy, sz, sq, bx, by, two, ax, ay, grp, ngrp);
                }
            }
            __syncthreads();
            rx[lane] = act ? ax : 0.0;
            ry[lane] = act ? ay : 0.0;
            __syncthreads();
            if (lane < na) {
                double ox = 0.0, oy = 0.0;
                for (int jj = 0; jj < ks; ++jj) {
                    ox += rx[lane + jj * na];
                    oy += ry[lane + jj * na];
                }
                part[offL + a0 + lane] = make_double2(ox, oy);
            }
        }
        if (two) {
            __syncthreads();
            for (int i = lane; i < cS; i += 64) part[offS + i] = make_double2(bx[i], by[i]);
        }
        return;
    }
    // a panel with itself: c = cL = cS points, staged; lanes hold point b0 + lane of a block of cb
    const int c = cL;
    // rotation within [b0, b0 + cb): self pair, then each other pair once
    auto self_block = [&](int b0, int cb, double& ax, double& ay) {
        if (lane >= cb) return;
        const int64_t kt = sL + b0 + lane;
        const double tx = px[kt], ty = py[kt], tz = pz[kt], tq = q[kt];
        const double ta = 1.0 - tz * tz;
        pair_ref(tx, ty, tz, ta, tq, b0 + lane, sx, sy, sz, sq, bx, by, false, ax, ay);
        for (int u = 1; 2 * u <= cb; ++u) {
            if (2 * u == cb && lane >= u) break;  // the middle step pairs lanes i and i + cb/2: once
            int sj = lane + u;
            if (sj >= cb) sj -= cb;
            pair_ref(tx, ty, tz, ta, tq, b0 + sj, sx, sy, sz, sq, bx, by, true, ax, ay);
        }
    };
    double ax = 0.0, ay = 0.0;
    if (c <= 64) {
        self_block(0, c, ax, ay);
        __syncthreads();
        if (lane < c) part[offL + lane] = make_double2(ax + bx[lane], ay + by[lane]);
        return;
    }
    const int c2 = c - 64;
    self_block(0, 64, ax, ay);  // points 0..63
    atomicAdd(&bx[lane], ax);   // after the block's own adds to bx[lane]
    atomicAdd(&by[lane], ay);
    ax = 0.0; ay = 0.0;
    self_block(64, c2, ax, ay);  // points 64..c-1 with each other
    if (lane < c2) {             // points 64..c-1 (lanes) with points 0..63 (staged)
        const int64_t kt = sL + 64 + lane;
        const double tx = px[kt], ty = py[kt], tz = pz[kt], tq = q[kt];
        const double ta = 1.0 - tz * tz;
        for (int u = 0; u < 64; ++u)
            pair_ref(tx, ty, tz, ta, tq, (lane + u) & 63, sx, sy, sz, sq, bx, by, true, ax, ay);
    }
    __syncthreads();
    part[offL + lane] = make_double2(bx[lane], by[lane]);
    if (lane < c2) part[offL + 64 + lane] = make_double2(ax + bx[64 + lane], ay + by[64 + lane]);
}

// Mirrored well-separated entries (see build_pair_tasks), one wave each: a CC entry with its CC
// mirror (both panels' proxies), or a PC entry (A's points <- B's proxies) with its CP mirror (B's
// proxies <- A's points).  The same rotation as k_pair64's two-way tasks (lanes on the side that
// takes fewer wave steps, LDS sums for the staged side in a fixed order) with the well-separated
// form of the reference expression (far_val) in both directions; the lane side projects its summed
// source vector once.  Partial sums include KP1 and exclude the target's factor KCONS / sqrt(1 - z^2),
// as the leaf kernel's and the proxy kernel's sums do before they scale.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
void k_farpair64(int ntask, const int4* __restrict__ tasks, int m, const int* __restrict__ slot,
                 const int64_t* __restrict__ start, const int* __restrict__ count, const double* __restrict__ px,
                 const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ q,
                 const double* __restrict__ prox, const double* __restrict__ W, const int64_t* __restrict__ pp_off,
                 const int64_t* __restrict__ xp_off, double2* __restrict__ pp_part, double2* __restrict__ xp_part) {
    __shared__ double sx[PAIR_MAX], sy[PAIR_MAX], sz[PAIR_MAX], sq[PAIR_MAX], bx[PAIR_MAX], by[PAIR_MAX];
    __shared__ double rx[64], ry[64], rz[64];
    const int item = xcd_item(ntask);
    if (item < 0) return;
    const int4 tk = tasks[item];
    const int lane = threadIdx.x;
    const bool pts = (tk.x >> 30) != 0;  // side A is a panel's points (else its proxies)
    const int A = tk.x & 0x3fffffff, B = tk.y;
    const int nA = pts ? count[A] : m, nB = m;
    const bool swap = pair_cost(nB, nA) < pair_cost(nA, nB);  // lanes on B
    // side descriptors: lane side L, staged side S
    const bool Lpts = pts && !swap, Spts = pts && swap;
    const int Lp = swap ? B : A, Sp = swap ? A : B;
    const int cL = swap ? nB : nA, cS = swap ? nA : nB;
    const int64_t offL = swap ? xp_off[tk.w] : (pts ? pp_off[tk.z] : xp_off[tk.z]);
    const int64_t offS = swap ? (pts ? pp_off[tk.z] : xp_off[tk.z]) : xp_off[tk.w];
    double2* const partL = Lpts ? pp_part : xp_part;
    double2* const partS = Spts ? pp_part : xp_part;
    auto entity = [&](bool isp, int X, int k, double& x, double& y, double& z, double& w) {
        if (isp) {
            const int64_t i = start[X] + k;
            x = px[i]; y = py[i]; z = pz[i]; w = q[i];
        } else {
            const int64_t i = (int64_t)slot[X] * m + k;
            x = prox[3 * i]; y = prox[3 * i + 1]; z = prox[3 * i + 2]; w = W[i];
        }
    };
    for (int i = lane; i < cS; i += 64) {
        double x, y, z, w;
        entity(Spts, Sp, i, x, y, z, w);
        sx[i] = x; sy[i] = y; sz[i] = z; sq[i] = w;
        bx[i] = 0.0; by[i] = 0.0;
    }
    __syncthreads();
    for (int a0 = 0; a0 < cL; a0 += 64) {
        const int na = min(64, cL - a0);
        int ks = 1;
        while (2 * ks * na <= 64 && 2 * ks * na <= cS) ks *= 2;
        const int h = cS / ks, j = lane / na, i = lane % na;
        const int ngrp = (na + cS - 1) / cS, grp = i / cS;
        const bool act = j < ks;
        double tx, ty, tz, tq;
        entity(Lpts, Lp, a0 + (act ? i : 0), tx, ty, tz, tq);
        const double ta = 1.0 - tz * tz;
        const int len = cS - (ks - 1) * h;
        const int mylen = act ? (j == ks - 1 ? len : h) : 0;
        // the lane side sums v = sum f w s and projects it once (pe(t <- s) = t.x s.y - t.y s.x and
        // pn(t <- s) = t.a s.z - t.z (t.x s.x + t.y s.y) are linear in s); the staged side takes
        // each pair's projections
        double ax = 0.0, ay = 0.0, az = 0.0;
        int si = i % cS + j * h;
        for (int u = 0; u < len; ++u, ++si) {
            if (u < mylen) {
                const int sj = si >= cS ? si - cS : si;
                const double ux = sx[sj], uy = sy[sj], uz = sz[sj];
                const double dxy = tx * ux + ty * uy;
                const double f = far_val(dxy + tz * uz);  // over KP1
                const double fs = f * sq[sj], ft = f * tq;
                ax = fma(fs, ux, ax);
                ay = fma(fs, uy, ay);
                az = fma(fs, uz, az);
                const double vx = fma(ux, ty, -(uy * tx)) * ft;            // pe(s <- t)
                const double vy = fma(tz, 1.0 - uz * uz, -(uz * dxy)) * ft;  // pn(s <- t)
                if (ngrp == 1) {
                    atomicAdd(&bx[sj], vx);
                    atomicAdd(&by[sj], vy);
                } else {
                    for (int gi = 0; gi < ngrp; ++gi)
                        if (grp == gi) {
                            atomicAdd(&bx[sj], vx);
                            atomicAdd(&by[sj], vy);
                        }
                }
            }
        }
        __syncthreads();
        rx[lane] = act ? ax : 0.0;
        ry[lane] = act ? ay : 0.0;
        rz[lane] = act ? az : 0.0;
        __syncthreads();
        if (lane < na) {  // lane < na holds point a0 + lane (j = 0)
            double ox = 0.0, oy = 0.0, oz = 0.0;
            for (int jj = 0; jj < ks; ++jj) {
                ox += rx[lane + jj * na];
                oy += ry[lane + jj * na];
                oz += rz[lane + jj * na];
            }
            const double pe = fma(tx, oy, -(ty * ox)), pn = fma(ta, oz, -(tz * fma(tx, ox, ty * oy)));
            partL[offL + a0 + lane] = make_double2(KP1 * pe, KP1 * pn);
        }
    }
    __syncthreads();
    for (int i = lane; i < cS; i += 64) partS[offS + i] = make_double2(KP1 * bx[i], KP1 * by[i]);
}

// Leaves: one block (one wave) per leaf, its points in batches of up to 64.  A batch of b points is
// split into groups of c = 64, or of c = 32, 16 and b mod 16 points (the binary digits of b), and a
// group runs k = 64 / c lanes per point: lane l takes point l % c of the group and every k-th source
// from l / c on, so a batch keeps all 64 lanes busy whatever b is.  The sources are staged in LDS
// once per batch for all its groups.  The k partial sums of a point are added in lane order through
// LDS and scaled by its KCONS / sqrt(1 - tz^2) once; then lane l interpolates the proxy values of
// every ancestor (L2P) at point l of the batch.
template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_leaf64(int nd, int m, int nleaves, const int* __restrict__ leaves,
                                               const int* __restrict__ parent, const int* __restrict__ pt_off,
                                               const int* __restrict__ pt_ent, const int* __restrict__ xt_off,
                                               const int* __restrict__ slot, const int64_t* __restrict__ start,
                                               const int* __restrict__ count, const double* __restrict__ bounds,
                                               const int* __restrict__ perm, const double* __restrict__ px,
                                               const double* __restrict__ py, const double* __restrict__ pz,
                                               const double2* __restrict__ uv, const double* __restrict__ q,
                                               const double* __restrict__ prox, const double* __restrict__ W,
                                               const double* __restrict__ phi, double* __restrict__ gx,
                                               double* __restrict__ gy, int accumulate,
                                               const int64_t* __restrict__ pp_off, const double2* __restrict__ part) {
    constexpr int NC = (MODE == SAL_MODE_REFERENCE) ? 2 : 3;
    constexpr int CH = 128;
    constexpr int NG = 3;  // groups per batch
    // LDS: the staged sources (x, y, z, weight); after a batch's sums its first half holds the lane
    // partial sums and its second half the point sums.  LDS, not registers, bounds the waves per
    // SIMD here (4 KiB a wave).
    __shared__ double lds[4 * CH];
    double* const sx = lds;
    double* const sy = lds + CH;
    double* const sz = lds + 2 * CH;
    double* const sq = lds + 3 * CH;
    double(*const red)[64] = reinterpret_cast<double(*)[64]>(lds);
    double(*const tsum)[64] = reinterpret_cast<double(*)[64]>(lds + 2 * CH);
    __shared__ double cheb[MAX_DEG + 1], nodes[2][MAX_DEG + 1];  // cos(pi i / deg); an ancestor's nodes
    const int item = xcd_item(nleaves);
    if (item < 0) return;
    const int L = leaves[item];
    const int cnt = count[L];
    const int64_t st = start[L];
    const int deg = nd - 1;
    const int lane = threadIdx.x;
    if (lane < nd) cheb[lane] = bli_cheb(deg, lane);  // read after the first barrier
    for (int b0 = 0; b0 < cnt; b0 += 64) {
        const int nb = min(64, cnt - b0);
        // groups: (first point in the batch, points); a group of c = 0 points is empty
        int g0[NG], gc[NG];
        if (nb == 64) {
            g0[0] = 0; gc[0] = 64;
            g0[1] = g0[2] = 64; gc[1] = gc[2] = 0;
        } else {
            gc[0] = nb & 32; gc[1] = nb & 16; gc[2] = nb & 15;
            g0[0] = 0; g0[1] = gc[0]; g0[2] = gc[0] + gc[1];
        }
        Tgt T[NG];
        int gk[NG], gs[NG];
        bool ga[NG];
        double ax[NG], ay[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            gk[g] = gc[g] > 0 ? 64 / gc[g] : 0;
            gs[g] = gc[g] > 0 ? lane / gc[g] : 0;
            ga[g] = gs[g] < gk[g];
            const int64_t kt = st + b0 + g0[g] + (gc[g] > 0 ? lane % gc[g] : 0);
            const int64_t kk = ga[g] ? kt : st;
            T[g].x = px[kk]; T[g].y = py[kk]; T[g].z = pz[kk];
            T[g].a = 1.0 - T[g].z * T[g].z;
            ax[g] = 0.0; ay[g] = 0.0;
        }
        for (int A = L; A >= 0; A = parent[A]) {
            for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
                const int S = pt_ent[e] >> 2, ty = pt_ent[e] & 3;
                if (pp_off && pp_off[e] >= 0) continue;  // summed by k_pair64, added below
                int ns;
                int64_t s0 = 0;
                const double* sprox = nullptr;
                const double* sw = nullptr;
                if (ty == 0) {
                    ns = count[S];
                    s0 = start[S];
                } else {
                    ns = m;
                    sprox = prox + (int64_t)slot[S] * m * 3;
                    sw = W + (int64_t)slot[S] * m;
                }
                for (int c0 = 0; c0 < ns; c0 += CH) {
                    const int nc = min(CH, ns - c0);
                    __syncthreads();
                    for (int t = lane; t < nc; t += 64) {
                        if (ty == 0) {
                            const int64_t kk = s0 + c0 + t;
                            sx[t] = px[kk]; sy[t] = py[kk]; sz[t] = pz[kk]; sq[t] = q[kk];
                        } else {  // PC: KP1 in the weight (gfunc_ref_far_u leaves it out)
                            const double* yk = sprox + (int64_t)(c0 + t) * 3;
                            sx[t] = yk[0]; sy[t] = yk[1]; sz[t] = yk[2];
                            sq[t] = MODE == SAL_MODE_REFERENCE ? KP1 * sw[c0 + t] : sw[c0 + t];
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        if (!ga[g]) continue;
                        double sxa = ax[g], sya = ay[g];
                        if (MODE == SAL_MODE_REFERENCE && ty != 0) {  // PC: a cluster's proxies
#pragma unroll 2
                            for (int t = gs[g]; t < nc; t += gk[g]) {
                                double pe, pn;
                                const double f = gfunc_ref_far_u(T[g], sx[t], sy[t], sz[t], pe, pn) * sq[t];
                                sxa = fma(pe, f, sxa);
                                sya = fma(pn, f, sya);
                            }
                        } else {
#pragma unroll 2
                            for (int t = gs[g]; t < nc; t += gk[g]) {
                                double pe, pn;
                                const double f = (MODE == SAL_MODE_REFERENCE
                                                      ? gfunc_ref_u(T[g], sx[t], sy[t], sz[t], pe, pn)
                                                      : gfunc_stable_u(T[g], sx[t], sy[t], sz[t], pe, pn)) *
                                                 sq[t];
                                sxa = fma(pe, f, sxa);
                                sya = fma(pn, f, sya);
                            }
                        }
                        ax[g] = sxa;
                        ay[g] = sya;
                    }
                }
            }
        }
        // PP entries k_pair64 summed: their partial sums at the points (sub-lane 0 of each point), in
        // list order, loads issued four entries at a time
        if (pp_off) {
            for (int A = L; A >= 0; A = parent[A]) {
                const int64_t rel = (st - start[A]) + b0;
                for (int e = pt_off[A]; e < pt_off[A + 1]; e += 4) {
                    double2 v[4][NG];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int64_t po = e + u < pt_off[A + 1] ? pp_off[e + u] : -1;
#pragma unroll
                        for (int g = 0; g < NG; ++g)
                            v[u][g] = (po >= 0 && ga[g] && gs[g] == 0) ? part[po + rel + g0[g] + lane % max(gc[g], 1)]
                                                                       : make_double2(0.0, 0.0);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int g = 0; g < NG; ++g) {
                            ax[g] += v[u][g].x;
                            ay[g] += v[u][g].y;
                        }
                }
            }
        }
        // per group: the k partial sums of each point, in lane order
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            if (gc[g] == 0) continue;
            __syncthreads();
            red[0][lane] = ga[g] ? ax[g] : 0.0;
            red[1][lane] = ga[g] ? ay[g] : 0.0;
            __syncthreads();
            if (lane < gc[g]) {
                double sxs = 0.0, sys = 0.0;
                for (int i = 0; i < gk[g]; ++i) {
                    sxs += red[0][lane + i * gc[g]];
                    sys += red[1][lane + i * gc[g]];
                }
                tsum[0][g0[g] + lane] = sxs;
                tsum[1][g0[g] + lane] = sys;
            }
        }
        // point b0 + lane: scale, L2P of every ancestor with proxy values (the ancestor's nodes formed
        // once in LDS), store
        __syncthreads();
        const bool own = lane < nb;
        const int64_t kt = st + b0 + (own ? lane : 0);
        const int p = perm[kt];
        const Tgt Tl = make_tgt(px[kt], py[kt], pz[kt]);
        double ox = tsum[0][lane] * Tl.kc, oy = tsum[1][lane] * Tl.kc;
        const double u = uv[kt].x, v = uv[kt].y;
        for (int A = L; A >= 0; A = parent[A]) {
            if (xt_off[A + 1] == xt_off[A]) continue;  // L2P of Phi_A (fast_sum_module.f90:186-199)
            const double* b = bounds + 6 * (int64_t)A;
            __syncthreads();
            if (lane < 2 * nd) {
                const int d = lane / nd, i = lane % nd;
                nodes[d][i] = deg == 0 ? 0.5 * (b[3 * d + 2] + b[3 * d]) : bli_node_cs(b[3 * d], b[3 * d + 2], cheb[i]);
            }
            __syncthreads();
            if (own) {
                const double* ph = phi + (int64_t)slot[A] * NC * m;
                double acc[NC];
                if (nd <= L2P_REG_ND) l2p_reg<NC, L2P_REG_ND>(u, v, nodes[0], nodes[1], nd, ph, acc);
                else l2p_mem<NC>(u, v, nodes[0], nodes[1], nd, ph, acc);
                if (MODE == SAL_MODE_REFERENCE) {
                    ox += acc[0];
                    oy += acc[1];
                } else {
                    double ge, gn;
                    project_en(Tl, acc[0], acc[1], acc[2], ge, gn);
                    ox += ge;
                    oy += gn;
                }
            }
        }
        if (own) {
            if (accumulate) {
                gx[p] += ox;
                gy[p] += oy;
            } else {
                gx[p] = ox;
                gy[p] = oy;
            }
        }
    }
}

void eval_fp64(sal_geom g, int degree, int mode, const double* area, const double* ssh, int R, double* gx, double* gy,
               int accumulate, hipStream_t s) {
    Plan* pl = get_plan(g, degree, mode == SAL_MODE_REFERENCE ? PLAN_FP64_REF : PLAN_FP64_ACC);
    eval_fp64_plan(g, pl, mode, area, ssh, R, gx, gy, accumulate, s);
}

// A leaf untouched by the plan's entries still gets its outputs written (zero, or unchanged when
// accumulating): every leaf has a block.
void eval_fp64_plan(sal_geom g, Plan* pl, int mode, const double* area, const double* ssh, int R, double* gx,
                    double* gy, int accumulate, hipStream_t s) {
    const int64_t n = g->n;
    const int m = pl->m, nd = pl->nd, NC = (mode == SAL_MODE_REFERENCE) ? 2 : 3;
    SAL_REQUIRE(pl->kind != PLAN_MIXED && pl->uv.p, SAL_ERR_STATE, "FP64 evaluation needs an FP64 plan");
    pl->q.ensure(n);
    pl->W.ensure((size_t)std::max(pl->Q, 1) * m);
    pl->wpart.ensure((size_t)std::max(pl->n_chunks, 1) * m);
    pl->phi.ensure((size_t)std::max(pl->Q, 1) * NC * m);
    for (int r = 0; r < R; ++r) {
        const double* sr = ssh + (int64_t)r * n;
        double* ox = gx + (int64_t)r * n;
        double* oy = gy + (int64_t)r * n;
        k_gather_q<<<grid_for(n, 256), 256, 0, s>>>(n, g->perm.p, area, sr, pl->q.p);
        if (pl->n_chunks > 0) {
            ProfScope ps("p2m64", s);
            k_p2m64<<<pl->n_chunks, 128, 0, s>>>(nd, pl->ck_slot.p, pl->ck_first.p, g->d_start.p, g->d_count.p,
                                                 g->d_bounds.p, pl->uv.p, pl->q.p, pl->wpart.p);
            k_p2m_reduce<<<pl->Q, 128, 0, s>>>(pl->Q, m, pl->ck_off.p, pl->wpart.p, pl->W.p);
        }
        // mirrored well-separated entries: partial sums for the proxy and the leaf kernels
        const bool far = mode == SAL_MODE_REFERENCE && pl->n_far_tasks > 0;
        if (far) {
            ProfScope pf("farpair64", s);
            k_farpair64<<<xcd_grid(pl->n_far_tasks), 64, 0, s>>>(
                pl->n_far_tasks, pl->far_task.p, m, pl->slot.p, g->d_start.p, g->d_count.p, g->px.p, g->py.p, g->pz.p,
                pl->q.p, pl->prox.p, pl->W.p, pl->pp_off.p, pl->xp_off.p, pl->pp_part.p, pl->xp_part.p);
        }
        if (pl->n_xt > 0) {
            ProfScope ps("proxy64", s);
            if (mode == SAL_MODE_REFERENCE)
                k_proxy64<SAL_MODE_REFERENCE><<<xcd_grid(pl->n_xt), 256, 0, s>>>(
                    m, pl->n_xt, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p, pl->slot.p, g->d_start.p, g->d_count.p, g->px.p,
                    g->py.p, g->pz.p, pl->q.p, pl->prox.p, pl->W.p, pl->phi.p, far ? pl->xp_off.p : nullptr,
                    far ? pl->xp_part.p : nullptr);
            else
                k_proxy64<SAL_MODE_ACCURATE><<<xcd_grid(pl->n_xt), 256, 0, s>>>(
                    m, pl->n_xt, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p, pl->slot.p, g->d_start.p, g->d_count.p, g->px.p,
                    g->py.p, g->pz.p, pl->q.p, pl->prox.p, pl->W.p, pl->phi.p, nullptr, nullptr);
        }
        const bool pairs = mode == SAL_MODE_REFERENCE && pl->n_pair_tasks > 0;
        const bool parts = pairs || far;  // the leaf kernel adds partial sums of PP and PC entries
        if (pairs) {
            ProfScope ps("pair64", s);
            k_pair64<<<xcd_grid(pl->n_pair_tasks), 64, 0, s>>>(pl->n_pair_tasks, pl->pair_task.p, g->d_start.p,
                                                                g->d_count.p, g->px.p, g->py.p, g->pz.p, pl->q.p,
                                                                pl->pp_off.p, pl->pp_part.p);
        }
        ProfScope ps("leaf64", s);
        if (mode == SAL_MODE_REFERENCE)
            k_leaf64<SAL_MODE_REFERENCE><<<xcd_grid(pl->n_leaves), 64, 0, s>>>(
                nd, m, pl->n_leaves, pl->leaves.p, g->d_parent.p, pl->pt_off.p, pl->pt_ent.p, pl->xt_off.p, pl->slot.p, g->d_start.p,
                g->d_count.p, g->d_bounds.p, g->perm.p, g->px.p, g->py.p, g->pz.p, pl->uv.p, pl->q.p,
                pl->prox.p, pl->W.p, pl->phi.p, ox, oy, accumulate, parts ? pl->pp_off.p : nullptr,
                parts ? pl->pp_part.p : nullptr);
        else
            k_leaf64<SAL_MODE_ACCURATE><<<xcd_grid(pl->n_leaves), 64, 0, s>>>(
                nd, m, pl->n_leaves, pl->leaves.p, g->d_parent.p, pl->pt_off.p, pl->pt_ent.p, pl->xt_off.p, pl->slot.p, g->d_start.p,
                g->d_count.p, g->d_bounds.p, g->perm.p, g->px.p, g->py.p, g->pz.p, pl->uv.p, pl->q.p,
                pl->prox.p, pl->W.p, pl->phi.p, ox, oy, accumulate, nullptr, nullptr);
        SAL_HIP(hipGetLastError());
    }
}

// ------------------------------------------------------------------------------ direct sum

constexpr int DS_TG = 64;        // targets per block
constexpr int DS_CHUNK = 16384;  // sources per block

template <int MODE>
__global__ __launch_bounds__(64) void k_direct64(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, const double* __restrict__ area,
                                                 const double* __restrict__ ssh, const int* __restrict__ tgt,
                                                 int64_t nt, double* __restrict__ part) {
    __shared__ double sx[256], sy[256], sz[256], sq[256];
    const int64_t a = (int64_t)blockIdx.x * DS_TG + threadIdx.x;
    const bool act = a < nt;
    const int t = act ? tgt[a] : 0;
    const double tx = x[t], ty = y[t], tz = z[t];
    Tgt T = make_tgt(tx, ty, tz);
    const int64_t c0 = (int64_t)blockIdx.y * DS_CHUNK, c1 = min(n, c0 + DS_CHUNK);
    double ax = 0.0, ay = 0.0;
    for (int64_t b0 = c0; b0 < c1; b0 += 256) {
        int nc = (int)min((int64_t)256, c1 - b0);
        __syncthreads();
        for (int k = threadIdx.x; k < nc; k += 64) {
            sx[k] = x[b0 + k]; sy[k] = y[b0 + k]; sz[k] = z[b0 + k]; sq[k] = area[b0 + k] * ssh[b0 + k];
        }
        __syncthreads();
        for (int k = 0; k < nc; ++k) {
            if (MODE != 0 && b0 + k == t) continue;
            double ge, gn;
            if (MODE == 2) gfunc_stable(T, sx[k], sy[k], sz[k], ge, gn);
            else gfunc_ref(T, sx[k], sy[k], sz[k], ge, gn);
            ax = fma(ge, sq[k], ax);
            ay = fma(gn, sq[k], ay);
        }
    }
    if (act) {
        part[((int64_t)blockIdx.y * nt + a) * 2] = ax;
        part[((int64_t)blockIdx.y * nt + a) * 2 + 1] = ay;
    }
}

__global__ void k_direct_reduce(int64_t nt, int nchunk, const double* __restrict__ part, double* __restrict__ gx,
                                double* __restrict__ gy) {
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nt) return;
    double sx = 0.0, sy = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        sx += part[((int64_t)c * nt + a) * 2];
        sy += part[((int64_t)c * nt + a) * 2 + 1];
    }
    gx[a] = sx;
    gy[a] = sy;
}

// The gradient as a tangential Cartesian vector: T = east e_lon + north e_lat, with the unit vectors
// the reference projects on (fast_sum_module.f90:25-27): e_lon = (-ty, tx, 0) / c and
// e_lat = (-tz tx, -tz ty, 1 - tz^2) / c = (-tz tx / c, -tz ty / c, c), c = cos(lat) taken as
// hypot(tx, ty) (no cancellation near the poles).  At a pole both components are zero
// (fast_sum_module.f90:19), and so is T.  g3 is [R][n][3].
__global__ void k_en_to_xyz(int64_t n, int R, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, const double* __restrict__ gx,
                            const double* __restrict__ gy, double* __restrict__ g3) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * R) return;
    const int64_t p = i % n;
    const double tx = x[p], ty = y[p], tz = z[p], e = gx[i], nn = gy[i];
    const double c = hypot(tx, ty);
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (c > 0.0) {
        ox = (-ty * e - tz * tx * nn) / c;
        oy = (tx * e - tz * ty * nn) / c;
        oz = c * nn;
    }
    g3[3 * i] = ox;
    g3[3 * i + 1] = oy;
    g3[3 * i + 2] = oz;
}

}  // namespace sal

using namespace sal;

extern "C" {

int sal_eval_opts_default(sal_eval_opts* o) {
    return guard([&] {
        SAL_REQUIRE(o, SAL_ERR_ARG, "null pointer");
        o->degree = 8;
        o->mode = SAL_MODE_ACCURATE;
        o->precision = SAL_PREC_MIXED;
        o->m2l = SAL_M2L_BF16X3;
        o->rhs_tile = 0;
        o->graph = 0;
        o->store = 0;
        o->near64 = 2;
    });
}

int sal_fast_sum(sal_geom g, const double* area, const double* ssh, int degree, double* gx, double* gy,
                 int accumulate) {
    return guard([&] {
        SAL_REQUIRE(g && area && ssh && gx && gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        const int64_t n = g->n;
        hipStream_t s = g->stream;
        DBuf<double> da(n), ds(n), ox(n), oy(n);
        h2d(da.p, area, n, s);
        h2d(ds.p, ssh, n, s);
        eval_fp64(g, degree, SAL_MODE_REFERENCE, da.p, ds.p, 1, ox.p, oy.p, 0, s);
        std::vector<double> hx(n), hy(n);
        d2h(hx.data(), ox.p, n, s);
        d2h(hy.data(), oy.p, n, s);
        SAL_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) {
            gx[i] = accumulate ? gx[i] + hx[i] : hx[i];
            gy[i] = accumulate ? gy[i] + hy[i] : hy[i];
        }
    });
}

int sal_work_counts(sal_geom g, int degree, int64_t K4[4]) {
    return guard([&] {
        SAL_REQUIRE(g && K4, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(degree >= 0 && degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        const int64_t m = (int64_t)(degree + 1) * (degree + 1);
        std::vector<int> it(g->nint), is(g->nint), ty(g->nint);
        d2h(it.data(), g->it.p, g->nint, g->stream);
        d2h(is.data(), g->is.p, g->nint, g->stream);
        d2h(ty.data(), g->ty.p, g->nint, g->stream);
        SAL_HIP(hipStreamSynchronize(g->stream));
        for (int k = 0; k < 4; ++k) K4[k] = 0;
        for (int64_t e = 0; e < g->nint; ++e) {
            int64_t ct = g->count[it[e]], cs = g->count[is[e]];
            switch (ty[e]) {
                case 0: K4[0] += ct * cs; break;
                case 1: K4[1] += ct * m; break;
                case 2: K4[2] += m * cs; break;
                default: K4[3] += m * m; break;
            }
        }
    });
}

int sal_eval_dev(sal_geom g, const sal_eval_opts* o, const double* d_area, const double* d_ssh, int R,
                 double* d_gx, double* d_gy, int accumulate, void* stream) {
    return guard([&] {
        SAL_REQUIRE(g && o && d_area && d_ssh && d_gx && d_gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->degree >= 0 && o->degree <= MAX_DEG, SAL_ERR_ARG, "degree must be in [0, 16]");
        SAL_REQUIRE(o->mode == SAL_MODE_REFERENCE || o->mode == SAL_MODE_ACCURATE, SAL_ERR_ARG, "bad mode");
        SAL_REQUIRE(R >= 1, SAL_ERR_ARG, "R must be >= 1");
        hipStream_t s = (hipStream_t)stream;
        if (o->precision == SAL_PREC_FP64) {
            eval_fp64(g, o->degree, o->mode, d_area, d_ssh, R, d_gx, d_gy, accumulate, s);
        } else {
            SAL_REQUIRE(o->precision == SAL_PREC_MIXED, SAL_ERR_ARG, "bad precision");
            SAL_REQUIRE(o->mode == SAL_MODE_ACCURATE, SAL_ERR_ARG, "mixed precision runs in SAL_MODE_ACCURATE");
            eval_mixed(g, *o, d_area, d_ssh, R, d_gx, d_gy, accumulate, s);
        }
    });
}

}  // extern "C"

// The direct sum's launches: per field, the per-chunk partial sums into part (nchunk x nt x 2, every
// element written), then their reduction into the outputs; part is the caller's (an allocation of
// the allocating entry point, or the static scratch below).
static size_t direct_scratch_doubles(int64_t n, int64_t nt) {
    return (size_t)((n + DS_CHUNK - 1) / DS_CHUNK) * (size_t)nt * 2;
}

static void direct_launch(const double* d_x, const double* d_y, const double* d_z, const double* d_area,
                          const double* d_ssh, int64_t n, int R, const int* d_targets, int64_t nt, int mode,
                          double* d_gx, double* d_gy, double* part, hipStream_t s) {
    const int nchunk = (int)((n + DS_CHUNK - 1) / DS_CHUNK);
    dim3 grid((unsigned)((nt + DS_TG - 1) / DS_TG), (unsigned)nchunk);
    for (int r = 0; r < R; ++r) {
        const double* sr = d_ssh + (int64_t)r * n;
        if (mode == 0) k_direct64<0><<<grid, 64, 0, s>>>(n, d_x, d_y, d_z, d_area, sr, d_targets, nt, part);
        else if (mode == 1) k_direct64<1><<<grid, 64, 0, s>>>(n, d_x, d_y, d_z, d_area, sr, d_targets, nt, part);
        else k_direct64<2><<<grid, 64, 0, s>>>(n, d_x, d_y, d_z, d_area, sr, d_targets, nt, part);
        k_direct_reduce<<<grid_for(nt, 256), 256, 0, s>>>(nt, nchunk, part, d_gx + (int64_t)r * nt,
                                                          d_gy + (int64_t)r * nt);
        SAL_HIP(hipGetLastError());
    }
}

// Scratch of sal_direct_sum_dev_noalloc: the partial sums in file-scope device memory, 64 MiB that
// every GPU the library runs on holds until the process ends, so the entry point allocates nothing.
// In host code the array's name is not a device address: its address on each device is looked up
// once with that device current (scratch_ptr) and only that pointer is passed to kernels.
constexpr size_t SCRATCH_DOUBLES = (size_t)8 << 20;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[MAX_DEVICES] = {};

static double* scratch_ptr() {
    int dev = 0;
    SAL_HIP(hipGetDevice(&dev));
    SAL_REQUIRE(dev >= 0 && dev < MAX_DEVICES, SAL_ERR_LIMIT, "device id beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        SAL_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
        SAL_REQUIRE(p != nullptr, SAL_ERR_HIP, "no device address for the direct sum's scratch");
        g_scratch_dev[dev] = (double*)p;
    }
    return g_scratch_dev[dev];
}

extern "C" {

int sal_direct_sum_dev(const double* d_x, const double* d_y, const double* d_z, const double* d_area,
                       const double* d_ssh, int64_t n, int R, const int* d_targets, int64_t nt, int mode,
                       double* d_gx, double* d_gy, void* stream) {
    return guard([&] {
        SAL_REQUIRE(d_x && d_y && d_z && d_area && d_ssh && d_targets && d_gx && d_gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(n >= 1 && nt >= 0 && R >= 1, SAL_ERR_ARG, "bad sizes");
        SAL_REQUIRE(mode >= 0 && mode <= 2, SAL_ERR_ARG, "direct-sum mode must be 0, 1 or 2");
        if (nt == 0) return;
        hipStream_t s = (hipStream_t)stream;
        DBuf<double> part(direct_scratch_doubles(n, nt));
        direct_launch(d_x, d_y, d_z, d_area, d_ssh, n, R, d_targets, nt, mode, d_gx, d_gy, part.p, s);
        SAL_HIP(hipStreamSynchronize(s));  // part is freed on return
    });
}

int sal_direct_sum_dev_noalloc(const double* d_x, const double* d_y, const double* d_z, const double* d_area,
                               const double* d_ssh, int64_t n, int R, const int* d_targets, int64_t nt, int mode,
                               double* d_gx, double* d_gy, void* stream) {
    return guard([&] {
        SAL_REQUIRE(d_x && d_y && d_z && d_area && d_ssh && d_targets && d_gx && d_gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(n >= 1 && nt >= 0 && R >= 1, SAL_ERR_ARG, "bad sizes");
        SAL_REQUIRE(mode >= 0 && mode <= 2, SAL_ERR_ARG, "direct-sum mode must be 0, 1 or 2");
        if (nt == 0) return;
        const size_t need = direct_scratch_doubles(n, nt);
        SAL_REQUIRE(need <= SCRATCH_DOUBLES, SAL_ERR_LIMIT,
                    "sal_direct_sum_dev_noalloc: " + std::to_string(nt) + " targets of " + std::to_string(n) +
                        " points need " + std::to_string(need) + " doubles of scratch (2 per target per " +
                        std::to_string(DS_CHUNK) + " points), more than its static " +
                        std::to_string(SCRATCH_DOUBLES) + "; use sal_direct_sum_dev or fewer targets");
        direct_launch(d_x, d_y, d_z, d_area, d_ssh, n, R, d_targets, nt, mode, d_gx, d_gy, scratch_ptr(),
                      (hipStream_t)stream);
    });
}

int sal_gradient_xyz_dev(const double* d_x, const double* d_y, const double* d_z, int64_t n, int R,
                         const double* d_gx, const double* d_gy, double* d_g3, void* stream) {
    return guard([&] {
        SAL_REQUIRE(d_x && d_y && d_z && d_gx && d_gy && d_g3, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(n >= 0 && R >= 1, SAL_ERR_ARG, "bad sizes");
        if (n == 0) return;
        hipStream_t s = (hipStream_t)stream;
        k_en_to_xyz<<<grid_for(n * R, 256), 256, 0, s>>>(n, R, d_x, d_y, d_z, d_gx, d_gy, d_g3);
        SAL_HIP(hipGetLastError());
    });
}

}  // extern "C"
