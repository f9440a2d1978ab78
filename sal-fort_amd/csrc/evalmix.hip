// Mixed-precision evaluation of R fields on one geometry (SAL_PREC_MIXED, SAL_MODE_ACCURATE).
//
// Same plan as the FP64 path (eval64.hip): per-panel P2M, cluster targets (CP, CC entries), then one
// block per leaf for the PP / PC entries of the leaf's root path and the L2P of the cluster-target
// values.  Fields are processed in tiles of RT (8, 16 or 32) so every source value and every kernel
// evaluation is shared by RT fields.
//
// Precision:
//   - FP64: P2M weights (barycentric basis values and their sums), every coordinate difference
//     s - t (then rounded to FP32, which keeps the chord's relative accuracy for near pairs), the
//     per-target running sums;
//   - FP32: the kernel (chord, regularised p1 + p2, projection), the products with the fields;
//     the near-field contributions of one source panel are summed in FP32 (at most a leaf's worth
//     of terms) and then promoted into the FP64 running sum (compensated leaf accumulation);
//   - cluster targets collect U = sum val W (y - c_T) and Phi0 = sum val W, relative to the target
//     panel's centre c_T, and form the tangential vector P_z (U - (z - c_T) Phi0) at each proxy z:
//     no FP32 cancellation (|y - c_T| and |z - c_T| are bounded by the separation).
//   - the proxy-to-proxy (CC) products run on MFMA when requested (m2l != SAL_M2L_FP32_VALU).
#include <algorithm>

#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct MixBuf {
    int R = 0;
    DBuf<double> Qd, Wd, wpart;  // [n][R], [Q][m][R], [chunks][m][R]
    DBuf<float> Qf, W32, phi;    // [n][R], [Q][m][R], [Q][3][m][R]
};

static void mix_free(void* p) { delete (MixBuf*)p; }

static MixBuf* mixbuf(Plan* pl) {
    if (!pl->mixed) {
        pl->mixed = new MixBuf();
        pl->mixed_free = mix_free;
    }
    return (MixBuf*)pl->mixed;
}

// FP32 kernel from a coordinate difference d = s - t (computed in FP64): ke, kn such that the
// east/north gradient contribution is (ke, kn) * q.  tz, a = 1 - tz^2, ic = 1/sqrt(a) (0 at poles).
__device__ __forceinline__ void kern32(float tx, float ty, float tz, float a, float ic, float dx, float dy, float dz,
                                       float& ke, float& kn) {
    float c2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    float c = __builtin_sqrtf(c2);
    float d1 = fmaf(c, c2, 1e-16f);
    float d2 = fmaf(-0.5f * c2, 4.0f - c2, 1e-16f);
    float num = fmaf((float)KP1, d2, (float)KP2 * ((2.0f - c2) + c) * d1);
    float f = num * __builtin_amdgcn_rcpf(d1 * d2) * ((float)KCONS) * ic;
    ke = (tx * dy - ty * dx) * f;
    kn = (dz * a - tz * fmaf(tx, dx, ty * dy)) * f;
}

// scalar val(c) in FP32 from d
__device__ __forceinline__ float val32(float dx, float dy, float dz) {
    float c2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    float c = __builtin_sqrtf(c2);
    float d1 = fmaf(c, c2, 1e-16f);
    float d2 = fmaf(-0.5f * c2, 4.0f - c2, 1e-16f);
    float num = fmaf((float)KP1, d2, (float)KP2 * ((2.0f - c2) + c) * d1);
    return num * __builtin_amdgcn_rcpf(d1 * d2) * ((float)KCONS);
}

// q[k][r] = area * ssh_r in tree (DFS) order, FP64 and FP32
__global__ void k_gather_qR(int64_t n, int R, const int* __restrict__ perm, const double* __restrict__ area,
                            const double* __restrict__ ssh, double* __restrict__ Qd, float* __restrict__ Qf) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = perm[k];
    double a = area[p];
    for (int r = 0; r < R; ++r) {
        double v = a * ssh[(int64_t)r * n + p];
        Qd[k * R + r] = v;
        Qf[k * R + r] = (float)v;
    }
}

// P2M for R fields in FP64: one block per chunk of a slot's points, one thread per proxy (j < m),
// RT fields per pass over the chunk.
template <int RT>
__global__ __launch_bounds__(128) void k_p2m_R(int nd, int R, int r0, const int* __restrict__ ck_slot,
                                               const int* __restrict__ ck_first, const int64_t* __restrict__ start,
                                               const int* __restrict__ count, const double* __restrict__ bounds,
                                               const int* __restrict__ perm, const double* __restrict__ xi,
                                               const double* __restrict__ eta, const double* __restrict__ Qd,
                                               double* __restrict__ wpart) {
    __shared__ double fx[64][MAX_DEG + 1], fy[64][MAX_DEG + 1], fq[64][RT];
    const int A = ck_slot[blockIdx.x], f0 = ck_first[blockIdx.x];
    const int m = nd * nd, deg = nd - 1;
    const int cnt = min(P2M_CHUNK, count[A] - f0);
    const int64_t st = start[A] + f0;
    const double* b = bounds + 6 * (int64_t)A;
    for (int j0 = 0; j0 < m; j0 += 128) {
        const int j = j0 + threadIdx.x;
        const int i = j % nd, jj = j / nd;
        double acc[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) acc[r] = 0.0;
        for (int c0 = 0; c0 < cnt; c0 += 64) {
            int nc = min(64, cnt - c0);
            __syncthreads();
            if (threadIdx.x < nc) {
                int64_t k = st + c0 + threadIdx.x;
                int p = perm[k];
                double nodes[MAX_DEG + 1];
                for (int t = 0; t < nd; ++t) nodes[t] = bli_node(b[0], b[2], deg, t);
                bli_basis_1d(xi[p], nodes, deg, fx[threadIdx.x]);
                for (int t = 0; t < nd; ++t) nodes[t] = bli_node(b[3], b[5], deg, t);
                bli_basis_1d(eta[p], nodes, deg, fy[threadIdx.x]);
            }
            for (int t = threadIdx.x; t < nc * RT; t += 128) {
                int pt = t / RT, r = t % RT;
                fq[pt][r] = (r0 + r < R) ? Qd[(st + c0 + pt) * R + r0 + r] : 0.0;
            }
            __syncthreads();
            if (j < m) {
                for (int t = 0; t < nc; ++t) {
                    double bv = fx[t][i] * fy[t][jj];
#pragma unroll
                    for (int r = 0; r < RT; ++r) acc[r] = fma(bv, fq[t][r], acc[r]);
                }
            }
        }
        if (j < m) {
#pragma unroll
            for (int r = 0; r < RT; ++r)
                if (r0 + r < R) wpart[((int64_t)blockIdx.x * m + j) * R + r0 + r] = acc[r];
        }
    }
}

__global__ void k_p2m_reduce_R(int Q, int m, int R, const int* __restrict__ ck_off, const double* __restrict__ wpart,
                               double* __restrict__ Wd, float* __restrict__ W32) {
    int q = blockIdx.x;
    for (int w = threadIdx.x; w < m * R; w += blockDim.x) {
        double a = 0.0;
        for (int c = ck_off[q]; c < ck_off[q + 1]; ++c) a += wpart[(int64_t)c * m * R + w];
        Wd[(int64_t)q * m * R + w] = a;
        W32[(int64_t)q * m * R + w] = (float)a;
    }
}

// Cluster targets (VALU): one block per target panel, one thread per proxy.
template <int RT>
__global__ __launch_bounds__(128) void k_prox_mix(int m, int R, int r0, const int* __restrict__ xt_panels,
                                                  const int* __restrict__ xt_off, const int* __restrict__ xt_ent,
                                                  const int* __restrict__ slot, const int64_t* __restrict__ start,
                                                  const int* __restrict__ count, const double* __restrict__ centre,
                                                  const double* __restrict__ px, const double* __restrict__ py,
                                                  const double* __restrict__ pz, const float* __restrict__ Qf,
                                                  const double* __restrict__ prox, const float* __restrict__ W32,
                                                  float* __restrict__ phi) {
    constexpr int CH = 64;
    __shared__ double sx[CH], sy[CH], sz[CH];
    __shared__ float se[CH][3];
    __shared__ float sq[CH][RT];
    const int T = xt_panels[blockIdx.x];
    const int tq = slot[T];
    const double cx = centre[3 * T], cy = centre[3 * T + 1], cz = centre[3 * T + 2];
    for (int j0 = 0; j0 < m; j0 += 128) {
        const int j = j0 + threadIdx.x;
        const bool act = j < m;
        double zx = cx, zy = cy, zz = cz;
        if (act) {
            const double* z = prox + ((int64_t)tq * m + j) * 3;
            zx = z[0]; zy = z[1]; zz = z[2];
        }
        float U[3][RT], P0[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) { U[0][r] = U[1][r] = U[2][r] = P0[r] = 0.f; }
        for (int e = xt_off[T]; e < xt_off[T + 1]; ++e) {
            const int S = xt_ent[e] >> 2, ty = xt_ent[e] & 3;
            const int ns = (ty == 2) ? count[S] : m;
            const int64_t s0 = (ty == 2) ? start[S] : 0;
            const double* sprox = prox + (int64_t)slot[S] * m * 3;
            const float* sw = W32 + (int64_t)slot[S] * m * R;
            for (int c0 = 0; c0 < ns; c0 += CH) {
                const int nc = min(CH, ns - c0);
                __syncthreads();
                for (int t = threadIdx.x; t < nc; t += blockDim.x) {
                    double a, b, c;
                    if (ty == 2) { int64_t k = s0 + c0 + t; a = px[k]; b = py[k]; c = pz[k]; }
                    else { const double* yk = sprox + (int64_t)(c0 + t) * 3; a = yk[0]; b = yk[1]; c = yk[2]; }
                    sx[t] = a; sy[t] = b; sz[t] = c;
                    se[t][0] = (float)(a - cx); se[t][1] = (float)(b - cy); se[t][2] = (float)(c - cz);
                }
                for (int t = threadIdx.x; t < nc * RT; t += blockDim.x) {
                    int pt = t / RT, r = t % RT;
                    float v = 0.f;
                    if (r0 + r < R) v = (ty == 2) ? Qf[(s0 + c0 + pt) * R + r0 + r] : sw[(int64_t)(c0 + pt) * R + r0 + r];
                    sq[pt][r] = v;
                }
                __syncthreads();
                if (act) {
                    for (int t = 0; t < nc; ++t) {
                        float v = val32((float)(sx[t] - zx), (float)(sy[t] - zy), (float)(sz[t] - zz));
                        float e0 = v * se[t][0], e1 = v * se[t][1], e2 = v * se[t][2];
#pragma unroll
                        for (int r = 0; r < RT; ++r) {
                            float q = sq[t][r];
                            U[0][r] = fmaf(e0, q, U[0][r]);
                            U[1][r] = fmaf(e1, q, U[1][r]);
                            U[2][r] = fmaf(e2, q, U[2][r]);
                            P0[r] = fmaf(v, q, P0[r]);
                        }
                    }
                }
            }
        }
        if (act) {
            // T = P_z (U - b P0), b = z - c_T
            float bx = (float)(zx - cx), by = (float)(zy - cy), bz = (float)(zz - cz);
            float fzx = (float)zx, fzy = (float)zy, fzz = (float)zz;
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                if (r0 + r >= R) break;
                float vx = U[0][r] - bx * P0[r], vy = U[1][r] - by * P0[r], vz = U[2][r] - bz * P0[r];
                float rr = vx * fzx + vy * fzy + vz * fzz;
                float* o = phi + ((int64_t)tq * 3 * m + j) * R + r0 + r;
                o[0] = vx - rr * fzx;
                o[(int64_t)m * R] = vy - rr * fzy;
                o[(int64_t)2 * m * R] = vz - rr * fzz;
            }
        }
    }
}

// Leaves: one wave per leaf, one lane per target point; PP / PC entries of the root path, then L2P.
template <int RT>
__global__ __launch_bounds__(64) void k_leaf_mix(int nd, int m, int R, int r0, const int* __restrict__ leaves,
                                                 const int* __restrict__ parent, const int* __restrict__ pt_off,
                                                 const int* __restrict__ pt_ent, const int* __restrict__ xt_off,
                                                 const int* __restrict__ slot, const int64_t* __restrict__ start,
                                                 const int* __restrict__ count, const double* __restrict__ bounds,
                                                 const int* __restrict__ perm, const double* __restrict__ px,
                                                 const double* __restrict__ py, const double* __restrict__ pz,
                                                 const double* __restrict__ xi, const double* __restrict__ eta,
                                                 const float* __restrict__ Qf, const double* __restrict__ prox,
                                                 const float* __restrict__ W32, const float* __restrict__ phi,
                                                 double* __restrict__ gx, double* __restrict__ gy, int64_t n,
                                                 int accumulate) {
    constexpr int CH = 64;
    __shared__ double sx[CH], sy[CH], sz[CH];
    __shared__ float sq[CH][RT];
    const int L = leaves[blockIdx.x];
    const int cnt = count[L];
    const int64_t st = start[L];
    const int deg = nd - 1;
    for (int t0 = 0; t0 < cnt; t0 += 64) {
        const int me = t0 + threadIdx.x;
        const bool act = me < cnt;
        const int64_t k = st + (act ? me : 0);
        const int p = perm[k];
        const double tx = px[k], ty = py[k], tz = pz[k];
        Tgt T = make_tgt(tx, ty, tz);
        const float ftx = (float)tx, fty = (float)ty, ftz = (float)tz, fa = (float)T.a, fic = (float)T.inv_c;
        double ae[RT], an[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) ae[r] = an[r] = 0.0;
        for (int A = L; A >= 0; A = parent[A]) {
            for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
                const int S = pt_ent[e] >> 2, ty_ = pt_ent[e] & 3;
                const int ns = (ty_ == 0) ? count[S] : m;
                const int64_t s0 = (ty_ == 0) ? start[S] : 0;
                const double* sprox = prox + (int64_t)slot[max(S, 0)] * m * 3;
                const float* sw = W32 + (int64_t)slot[max(S, 0)] * m * R;
                float be[RT], bn[RT];
#pragma unroll
                for (int r = 0; r < RT; ++r) be[r] = bn[r] = 0.f;
                for (int c0 = 0; c0 < ns; c0 += CH) {
                    const int nc = min(CH, ns - c0);
                    __syncthreads();
                    for (int t = threadIdx.x; t < nc; t += 64) {
                        if (ty_ == 0) { int64_t kk = s0 + c0 + t; sx[t] = px[kk]; sy[t] = py[kk]; sz[t] = pz[kk]; }
                        else { const double* yk = sprox + (int64_t)(c0 + t) * 3; sx[t] = yk[0]; sy[t] = yk[1]; sz[t] = yk[2]; }
                    }
                    for (int t = threadIdx.x; t < nc * RT; t += 64) {
                        int pt = t / RT, r = t % RT;
                        float v = 0.f;
                        if (r0 + r < R) v = (ty_ == 0) ? Qf[(s0 + c0 + pt) * R + r0 + r] : sw[(int64_t)(c0 + pt) * R + r0 + r];
                        sq[pt][r] = v;
                    }
                    __syncthreads();
                    for (int t = 0; t < nc; ++t) {
                        float ke, kn;
                        kern32(ftx, fty, ftz, fa, fic, (float)(sx[t] - tx), (float)(sy[t] - ty), (float)(sz[t] - tz), ke,
                               kn);
#pragma unroll
                        for (int r = 0; r < RT; ++r) {
                            be[r] = fmaf(ke, sq[t][r], be[r]);
                            bn[r] = fmaf(kn, sq[t][r], bn[r]);
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    ae[r] += (double)be[r];
                    an[r] += (double)bn[r];
                }
            }
            if (xt_off[A + 1] > xt_off[A]) {  // L2P of the tangential vector at A's proxies, then project
                const double* b = bounds + 6 * (int64_t)A;
                double nodes[MAX_DEG + 1], fx[MAX_DEG + 1], fy[MAX_DEG + 1];
                for (int i = 0; i < nd; ++i) nodes[i] = bli_node(b[0], b[2], deg, i);
                bli_basis_1d(xi[p], nodes, deg, fx);
                for (int i = 0; i < nd; ++i) nodes[i] = bli_node(b[3], b[5], deg, i);
                bli_basis_1d(eta[p], nodes, deg, fy);
                const float* ph = phi + (int64_t)slot[A] * 3 * m * R + r0;
                float vx[RT], vy[RT], vz[RT];
#pragma unroll
                for (int r = 0; r < RT; ++r) vx[r] = vy[r] = vz[r] = 0.f;
                for (int jj = 0; jj < nd; ++jj)
                    for (int i = 0; i < nd; ++i) {
                        const float bv = (float)(fx[i] * fy[jj]);
                        const float* q0 = ph + (int64_t)(i + jj * nd) * R;
#pragma unroll
                        for (int r = 0; r < RT; ++r) {
                            if (r0 + r < R) {
                                vx[r] = fmaf(bv, q0[r], vx[r]);
                                vy[r] = fmaf(bv, q0[(int64_t)m * R + r], vy[r]);
                                vz[r] = fmaf(bv, q0[(int64_t)2 * m * R + r], vz[r]);
                            }
                        }
                    }
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    double ge, gn;
                    project_en(T, vx[r], vy[r], vz[r], ge, gn);
                    ae[r] += ge;
                    an[r] += gn;
                }
            }
        }
        if (act) {
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                if (r0 + r >= R) break;
                double* ox = gx + (int64_t)(r0 + r) * n + p;
                double* oy = gy + (int64_t)(r0 + r) * n + p;
                if (accumulate) { *ox += ae[r]; *oy += an[r]; }
                else { *ox = ae[r]; *oy = an[r]; }
            }
        }
    }
}

template <int RT>
static void run_tiles(sal_geom g, Plan* pl, MixBuf* mb, int R, double* gx, double* gy, int accumulate, hipStream_t s) {
    const int m = pl->m, nd = pl->nd;
    for (int r0 = 0; r0 < R; r0 += RT) {
        if (pl->n_chunks > 0) {
            ProfScope ps("p2m_mix", s);
            k_p2m_R<RT><<<pl->n_chunks, 128, 0, s>>>(nd, R, r0, pl->ck_slot.p, pl->ck_first.p, g->d_start.p,
                                                     g->d_count.p, g->d_bounds.p, g->perm.p, g->xi.p, g->eta.p,
                                                     mb->Qd.p, mb->wpart.p);
        }
    }
    if (pl->Q > 0) {
        ProfScope ps("p2m_reduce", s);
        k_p2m_reduce_R<<<pl->Q, 256, 0, s>>>(pl->Q, m, R, pl->ck_off.p, mb->wpart.p, mb->Wd.p, mb->W32.p);
    }
    for (int r0 = 0; r0 < R; r0 += RT) {
        if (pl->n_xt > 0) {
            ProfScope ps("far_mix", s);
            k_prox_mix<RT><<<pl->n_xt, 128, 0, s>>>(m, R, r0, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p, pl->slot.p,
                                                    g->d_start.p, g->d_count.p, g->d_centre.p, g->px.p, g->py.p,
                                                    g->pz.p, mb->Qf.p, pl->prox.p, mb->W32.p, mb->phi.p);
        }
        {
            ProfScope ps("near_mix", s);
            k_leaf_mix<RT><<<pl->n_leaves, 64, 0, s>>>(nd, m, R, r0, pl->leaves.p, g->d_parent.p, pl->pt_off.p,
                                                       pl->pt_ent.p, pl->xt_off.p, pl->slot.p, g->d_start.p,
                                                       g->d_count.p, g->d_bounds.p, g->perm.p, g->px.p, g->py.p,
                                                       g->pz.p, g->xi.p, g->eta.p, mb->Qf.p, pl->prox.p, mb->W32.p,
                                                       mb->phi.p, gx, gy, g->n, accumulate);
        }
        SAL_HIP(hipGetLastError());
    }
}

void eval_mixed(sal_geom g, const sal_eval_opts& o, const double* area, const double* ssh, int R, double* gx,
                double* gy, int accumulate, hipStream_t s) {
    SAL_REQUIRE(o.degree >= 1 && o.degree <= 10, SAL_ERR_ARG, "mixed precision supports degree 1..10");
    Plan* pl = get_plan(g, o.degree, PLAN_MIXED);
    MixBuf* mb = mixbuf(pl);
    const int64_t n = g->n;
    const int m = pl->m;
    mb->Qd.ensure((size_t)n * R);
    mb->Qf.ensure((size_t)n * R);
    mb->Wd.ensure((size_t)std::max(pl->Q, 1) * m * R);
    mb->W32.ensure((size_t)std::max(pl->Q, 1) * m * R);
    mb->wpart.ensure((size_t)std::max(pl->n_chunks, 1) * m * R);
    mb->phi.ensure((size_t)std::max(pl->Q, 1) * 3 * m * R);
    {
        ProfScope ps("gather_q", s);
        k_gather_qR<<<grid_for(n, 256), 256, 0, s>>>(n, R, g->perm.p, area, ssh, mb->Qd.p, mb->Qf.p);
    }
    int rt = o.rhs_tile > 0 ? o.rhs_tile : (R >= 32 ? 32 : R >= 16 ? 16 : 8);
    SAL_REQUIRE(rt == 8 || rt == 16 || rt == 32, SAL_ERR_ARG, "rhs_tile must be 8, 16 or 32");
    if (rt == 8) run_tiles<8>(g, pl, mb, R, gx, gy, accumulate, s);
    else if (rt == 16) run_tiles<16>(g, pl, mb, R, gx, gy, accumulate, s);
    else run_tiles<32>(g, pl, mb, R, gx, gy, accumulate, s);
}

}  // namespace sal
