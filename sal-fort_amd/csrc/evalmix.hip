// Mixed-precision evaluation of R fields on one geometry (SAL_PREC_MIXED, SAL_MODE_ACCURATE).
//
// Same plan as the FP64 path (eval64.hip), fields processed in tiles of RT (16 or 32):
//
//   1. qprep   Qf[k][r] = area * ssh_r in tree (DFS) order, FP32
//   2. p2m     W_S[k][r] for every cluster source (FP64 sums of FP64 basis values), stored FP32
//   3. far     cluster targets T: U[c][j][r] = sum val(z_j, y) w^r (y - c_T)_c and
//              Phi0[j][r] = sum val(z_j, y) w^r over T's CP entries (y: source points, w: q) and CC
//              entries (y: source proxies, w: W); T_j = P_z (U - (z_j - c_T) Phi0) is the tangential
//              vector at proxy z_j.  These are GEMMs [proxies x sources] x [sources x 4 RT]: the
//              kernel matrix goes to LDS as split bf16 (hi + lo) or FP32, the right-hand side is
//              built from LDS, and MFMA accumulates in FP32.
//   4. l2l     top-down, level by level: every non-leaf panel below a cluster target receives its
//              parent's tangential-vector interpolant at its own proxy grid (separable 1-D
//              interpolation, exact for the degree-d polynomial), plus its own cluster values.
//              Leaves are not pushed down to: the near kernel evaluates their parent's expansion on
//              the parent's grid (and a leaf's own cluster values on its own grid).
//   5. near    one wave per 32 targets of a leaf: every PP entry (source points) and PC entry
//              (source proxies with W) of the leaf's root path, as FP32 MFMA tiles
//              [16 targets x 4 sources] x [4 sources x 16 fields], east and north separately; each
//              source panel's FP32 sum is flushed into FP64 accumulators (compensated leaf sums).
//              The same wave then adds L2P: the leaf's interpolant at its 32 targets (an MFMA GEMM
//              of basis products and the local expansion), projected on each target's east/north
//              basis, before its single write of the sums.
//   5b. close  PP pairs closer than near64 times the target leaf's mean point spacing are left out of
//              the near kernel's FP32 stream and summed here in FP64 (kernel value, products, sums),
//              one thread per target over its close-pair list (built once per plan).
//   6. out     scatter to the caller's order, FP64.
//
// Precision: coordinate differences in FP64 then FP32 (the chord keeps its relative accuracy even
// for close pairs); kernel values and products in FP32 except the close pairs (FP64); near-field
// sums in FP32 per source panel and FP64 across panels; cluster values are formed relative to the
// target centre so no FP32 cancellation occurs in the tangential projection.  The FP32 kernel
// value's relative error (about 1e-7) times a term, squared and summed over a target's near field,
// is dominated by the few closest pairs, so computing those in FP64 removes most of the near
// field's rounding error (which the cancelling sum of a smooth field's gradient amplifies).
#include <algorithm>
#include <memory>
#include <mutex>

#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
// row stride (floats) of the packed fields, proxy weights and local expansions: whole RHS tiles
static inline int row_stride(int R, int rt) { return (R + rt - 1) / rt * rt; }

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Read-only work lists of a mixed plan (shared by every workspace evaluating the plan).
struct MixLists {
    DBuf<int2> near_work;              // (leaf, first target)
    DBuf<int> lslot;                   // panel -> local-expansion slot (-1: none)
    DBuf<int> l2l_panels;              // panels receiving their parent's interpolant, by level
    DBuf<unsigned char> own;           // per panel: has cluster-target entries of its own
    std::vector<int> l2l_lo;           // level ranges in l2l_panels
    int n_near = 0, n_full = 0, nloc = 0, n_l2p = 0;  // near_work: n_full items of 17..32 targets first
    // upward pass: U panels, their direct P2M sources (panels whose whole point range they take), the
    // M2M parents by level (deepest level first)
    DBuf<int> u_panels, uslot, p2m_off, p2m_src, m2m_panels;
    std::vector<int> m2m_lo;
    int nU = 0;
    // stored operator (SAL_eval_opts.store): near-field kernel values per work item and the
    // split-bf16 far-field kernel matrices per proxy-target panel, in the order the kernels use them
    bool stored_near = false, stored_far = false, stored_far32 = false, stored_far3 = false;  // far parts:
                                                                  // split bf16, FP32, three-part bf16
    DBuf<int64_t> near_off, far_off;   // start of every near work item (floats) / proxy-target panel (steps)
    DBuf<float> near_op;               // per item: [steps][64 lanes][2 rows per item row tile]
    DBuf<bf8> far_op;                  // [k-steps][full row tiles][hi, lo][64 lanes], then the last
                                       // tile's proxy rows only: [hi, lo][k-group][row]
    DBuf<bf8> far_op32;                // FP32: [k-steps][full row tiles][k-groups of 4][64 lanes], then
                                       // the last tile's proxy rows only: [k-groups][lk][row]
    DBuf<bf8> far_op3;                 // three-part bf16: as far_op with three parts
    int64_t store_bytes() const {
        return (int64_t)(near_op.bytes() + far_op.bytes() + far_op32.bytes() + far_op3.bytes() + near_off.bytes() +
                         far_off.bytes());
    }
    // close pairs (summed in FP64 by k_close): per leaf the FP32 chord-squared cutoff (negative: none),
    // per target point (tree order) its close PP sources (tree order), CSR
    DBuf<float> cut2;
    DBuf<int> close_off, close_src;
    DBuf<double2> close_val;           // per close pair: FP64 (east, north) kernel factors
    int64_t n_close = 0;
};

static void lists_free(void* p) { delete (MixLists*)p; }
static int64_t lists_bytes(const void* p) {
    const MixLists* ml = (const MixLists*)p;
    return (int64_t)(ml->near_work.bytes() + ml->lslot.bytes() + ml->l2l_panels.bytes() + ml->own.bytes() +
                     ml->u_panels.bytes() + ml->uslot.bytes() + ml->p2m_off.bytes() +
                     ml->p2m_src.bytes() + ml->m2m_panels.bytes() + ml->cut2.bytes() + ml->close_off.bytes() +
                     ml->close_src.bytes() + ml->close_val.bytes()) + ml->store_bytes();
}

}  // namespace sal

// Per-caller buffers of one mixed evaluation; one workspace evaluates one request at a time.
struct sal_work_s {
    sal::DBuf<float> Qf, W32, loc;     // [n][R], [Q][m][R], [nloc][3][m][R]
    sal::DBuf<double> wpart, Ge, Gn;   // upward-pass expansions [U panels][m][R], [n][R], [n][R]
    // captured launch sequences, keyed by everything their kernel arguments depend on
    struct Graph {
        uint64_t plan_serial;
        const void* ptr[4];
        int R, rt, m2l, acc, store;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    hipStream_t cap = nullptr;         // capture stream (the caller's may be the null stream)
    void drop_graphs() {
        for (auto& gr : graphs) (void)hipGraphExecDestroy(gr.exec);
        graphs.clear();
    }
    int64_t bytes() const {
        return (int64_t)(Qf.bytes() + W32.bytes() + loc.bytes() + wpart.bytes() + Ge.bytes() + Gn.bytes());
    }
    ~sal_work_s() {
        drop_graphs();
        if (cap) (void)hipStreamDestroy(cap);
    }
};

namespace sal {

// FP32 chord squared |d|^2 of d = s - t (components rounded from their FP64 differences).  The
// near kernel's close-pair test and the close-pair list both use this value, so they agree bit for bit.
__device__ __forceinline__ float chord2_f32(float dx, float dy, float dz) { return fmaf(dx, dx, fmaf(dy, dy, dz * dz)); }

// FP32 kernel value val(c) from the chord squared
__device__ __forceinline__ float val32_c2(float c2) {
    float c = __builtin_amdgcn_sqrtf(c2);  // the hardware root (1 ulp), not the rounded library one
    float d1 = fmaf(c, c2, 1e-16f);
    float d2 = fmaf(-0.5f * c2, 4.0f - c2, 1e-16f);
    float num = fmaf((float)KP1, d2, (float)KP2 * ((2.0f - c2) + c) * d1);
    return num * __builtin_amdgcn_rcpf(d1 * d2) * ((float)KCONS);
}

// FP32 kernel value val(c) from d = s - t
__device__ __forceinline__ float val32(float dx, float dy, float dz) { return val32_c2(chord2_f32(dx, dy, dz)); }

// 16-byte load (p 16-byte aligned)
__device__ __forceinline__ float4 ld4(const float* p) { return *(const float4*)p; }



// ------------------------------------------------------------------------------ 1. qprep

// Transposes through LDS: a block reads QB consecutive points of every field in the caller's order
// (coalesced along points) and writes each point's row of fields in tree order (32 consecutive lanes
// write one row segment of up to 32 fields).  Rows are Rs >= R floats (whole RHS tiles, 16-byte
// aligned); the columns past R hold zeros.
constexpr int QB = 256;
__global__ __launch_bounds__(QB) void k_qprep(int64_t n, int R, int Rs, const int* __restrict__ iperm,
                                             const double* __restrict__ area, const double* __restrict__ ssh,
                                             float* __restrict__ Qf) {
    __shared__ float tile[QB][33];
    __shared__ int64_t dst[QB];
    const int64_t p0 = (int64_t)blockIdx.x * QB;
    const int np = (int)min((int64_t)QB, n - p0);
    const int t = threadIdx.x;
    double a = 0.0;
    if (t < np) {
        a = area[p0 + t];
        dst[t] = (int64_t)iperm[p0 + t] * Rs;
    }
    for (int r0 = 0; r0 < Rs; r0 += 32) {
        const int nr = min(32, Rs - r0);
        __syncthreads();
        if (t < np)
            for (int c = 0; c < nr; ++c)
                tile[t][c] = (r0 + c < R) ? (float)(a * ssh[(int64_t)(r0 + c) * n + p0 + t]) : 0.f;
        __syncthreads();
        const int c = t & 31;
        for (int pt = t >> 5; pt < np; pt += QB / 32)
            if (c < nr) Qf[dst[pt] + r0 + c] = tile[pt][c];
    }
}

// ------------------------------------------------------------------------------ 2. upward pass

// Expansions W[k][r] = sum_s l_k(s) q_s^r of the "U panels": every panel holding more than the
// list's cluster threshold of points (the cluster sources the list uses are among them) plus any
// other panel the list uses as a cluster source.  P2M anterpolates each point once, into its lowest
// U ancestor: a U panel takes the whole point range of each child that is not a U panel (its own
// range when it is a leaf).  M2M then adds every U child's expansion into its parent, bottom-up
// level by level, by separable interpolation of the child grid onto the parent basis, which is exact
// for the degree-d polynomial.  FP64 throughout; rows of R doubles.  P2M is a GEMM per U panel,
// [proxies x points] basis products times [points x RT] weights, on FP64 MFMA 16x16x4 (4 points a
// step): wave w takes the row tiles w, w + 2, ...; lane l holds proxy 16 rt + l % 16 and point l / 16
// of A, field l % 16 and point l / 16 of B; the FP64 result tiles hold proxies 16 rt + l / 16 + 4 i
// (rows interleaved, unlike the FP32 16x16x4 layout).
typedef double d4 __attribute__((ext_vector_type(4)));
template <int ND, int RT>
__global__ __launch_bounds__(128) void k_p2m_up(int R, int Rs, int r0, const int* __restrict__ u_panels,
                                                const int* __restrict__ src_off, const int* __restrict__ src,
                                                const int64_t* __restrict__ start, const int* __restrict__ count,
                                                const double* __restrict__ bounds, const int* __restrict__ perm,
                                                const double* __restrict__ xi, const double* __restrict__ eta,
                                                const float* __restrict__ Qf, double* __restrict__ Wup) {
    constexpr int M = ND * ND, CH = 64, NRT = (M + 15) / 16, RPW = (NRT + 1) / 2, NCT = RT / 16;
    __shared__ double fx[CH][ND], fy[CH][ND];
    __shared__ double fq[CH][RT];
    __shared__ double nodes_s[2][ND];  // the panel's grid nodes in xi and eta
    const int u = blockIdx.x, A = u_panels[u];
    const double* b = bounds + 6 * (int64_t)A;
    const int w = threadIdx.x, wv = w >> 6, lane = w & 63, lr = lane & 15, lk = lane >> 4;
    if (w < 2 * ND) nodes_s[w / ND][w % ND] = bli_node(b[3 * (w / ND)], b[3 * (w / ND) + 2], ND - 1, w % ND);
    d4 Wt[RPW][NCT];
#pragma unroll
    for (int q = 0; q < RPW; ++q)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) Wt[q][ct] = (d4){0.0, 0.0, 0.0, 0.0};
    for (int e = src_off[u]; e < src_off[u + 1]; ++e) {
        const int C = src[e], cnt = count[C];
        const int64_t st = start[C];
        for (int c0 = 0; c0 < cnt; c0 += CH) {
            const int nc = min(CH, cnt - c0);
            __syncthreads();
            {  // threads 0..63: xi basis of point c0 + thread; 64..127: eta basis
                const int pt = threadIdx.x & (CH - 1), d = threadIdx.x / CH;
                if (pt < nc) {
                    const int p = perm[st + c0 + pt];
                    double nodes[ND], f[ND];
#pragma unroll
                    for (int t = 0; t < ND; ++t) nodes[t] = nodes_s[d][t];
                    bli_basis_1d(d ? eta[p] : xi[p], nodes, ND - 1, f);
#pragma unroll
                    for (int t = 0; t < ND; ++t) (d ? fy : fx)[pt][t] = f[t];
                }
            }
            {  // the chunk's field rows: all 16-byte loads first, then the LDS stores
                constexpr int V = RT / 4, NV = CH * V / 128;
                float4 v[NV];
#pragma unroll
                for (int jv = 0; jv < NV; ++jv) {
                    const int idx = threadIdx.x + 128 * jv, pt = idx / V, c4 = idx % V;
                    v[jv] = pt < nc ? ld4(Qf + (st + c0 + pt) * Rs + r0 + 4 * c4) : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int jv = 0; jv < NV; ++jv) {
                    const int idx = threadIdx.x + 128 * jv, pt = idx / V, c4 = idx % V;
                    fq[pt][4 * c4] = v[jv].x;
                    fq[pt][4 * c4 + 1] = v[jv].y;
                    fq[pt][4 * c4 + 2] = v[jv].z;
                    fq[pt][4 * c4 + 3] = v[jv].w;
                }
            }
            __syncthreads();
            for (int s0 = 0; s0 < nc; s0 += 4) {
                const int pt = s0 + lk;
                const bool pv = pt < nc;  // rows past the chunk hold zero weights
                double bq[NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) bq[ct] = fq[pt][16 * ct + lr];
#pragma unroll
                for (int q = 0; q < RPW; ++q) {
                    const int rt = wv + 2 * q;
                    if (rt >= NRT) break;
                    const int k = 16 * rt + lr;
                    const double a = (pv && k < M) ? fx[pt][k % ND] * fy[pt][k / ND] : 0.0;
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct)
                        Wt[q][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bq[ct], Wt[q][ct], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int rt = wv + 2 * q;
        if (rt >= NRT) break;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int i4 = 0; i4 < 4; ++i4) {
                const int k = 16 * rt + lk + 4 * i4, r = 16 * ct + lr;
                if (k < M && r0 + r < R) Wup[((int64_t)u * M + k) * R + r0 + r] = Wt[q][ct][i4];
            }
    }
}

// One block per parent P of a level: W_P += sum over U children C of the interpolation of W_C's grid
// onto P's basis: W_P[a][b] += sum_{i,j} l^P_a(xi^C_i) l^P_b(eta^C_j) W_C[i][j].  Children in a
// fixed order, fields 32 at a time, in two passes through LDS: thread (field, j) reads column j of
// W_C once and forms u[a][j] = sum_i Lx[i][a] W_C[i][j]; then thread (field, a) adds
// sum_j Ly[j][b] u[a][j] to W_P[a][b].
constexpr int M2M_THREADS = 384;
template <int ND>
__global__ __launch_bounds__(M2M_THREADS) void k_m2m(int R, const int* __restrict__ parents, const int* __restrict__ child1,
                                             const int* __restrict__ uslot, const double* __restrict__ bounds,
                                             double* __restrict__ Wup) {
    constexpr int M = ND * ND;
    __shared__ double Lx[ND][ND], Ly[ND][ND];  // [child node][parent basis]
    __shared__ double U[ND][ND][32];           // u[a][j] of 32 fields
    __shared__ double cs[ND];                  // grid nodes on [-1, 1]
    const int P = parents[blockIdx.x];
    const double* bp = bounds + 6 * (int64_t)P;
    double* dst = Wup + (int64_t)uslot[P] * M * R;
    if (threadIdx.x < ND) cs[threadIdx.x] = bli_cheb(ND - 1, threadIdx.x);  // read after the loop's barrier
    for (int k = 0; k < 4; ++k) {
        const int C = child1[P] + k, uc = uslot[C];
        if (uc < 0) continue;  // uniform over the block
        __syncthreads();
        if (threadIdx.x < 2 * ND) {
            const int d = threadIdx.x / ND, c = threadIdx.x % ND;  // d 0: xi, 1: eta
            const double* bc = bounds + 6 * (int64_t)C;
            double nodes[ND], f[ND];
#pragma unroll
            for (int t = 0; t < ND; ++t) nodes[t] = bli_node_cs(bp[3 * d], bp[3 * d + 2], cs[t]);
            bli_basis_1d(bli_node_cs(bc[3 * d], bc[3 * d + 2], cs[c]), nodes, ND - 1, f);
#pragma unroll
            for (int t = 0; t < ND; ++t) {
                if (d == 0) Lx[c][t] = f[t];
                else Ly[c][t] = f[t];
            }
        }
        __syncthreads();
        const double* sp = Wup + (int64_t)uc * M * R;
        for (int f0 = 0; f0 < R; f0 += 32) {
            const int nf = min(32, R - f0);
            for (int w = threadIdx.x; w < 32 * ND; w += blockDim.x) {
                const int r = w % 32, j = w / 32;
                if (r >= nf) continue;
                double p[ND];
#pragma unroll
                for (int i = 0; i < ND; ++i) p[i] = sp[(int64_t)(i + j * ND) * R + f0 + r];
#pragma unroll 1
                for (int a = 0; a < ND; ++a) {  // rolled: the weights stay in LDS
                    double acc = 0.0;
#pragma unroll
                    for (int i = 0; i < ND; ++i) acc = fma(Lx[i][a], p[i], acc);
                    U[a][j][r] = acc;
                }
            }
            __syncthreads();
            for (int w = threadIdx.x; w < 32 * ND; w += blockDim.x) {
                const int r = w % 32, a = w / 32;
                if (r >= nf) continue;
                double t[ND];
#pragma unroll
                for (int j = 0; j < ND; ++j) t[j] = U[a][j][r];
#pragma unroll 1
                for (int bb = 0; bb < ND; ++bb) {
                    double acc = dst[(int64_t)(a + bb * ND) * R + f0 + r];
#pragma unroll
                    for (int j = 0; j < ND; ++j) acc = fma(Ly[j][bb], t[j], acc);
                    dst[(int64_t)(a + bb * ND) * R + f0 + r] = acc;
                }
            }
            __syncthreads();  // U is rewritten by the next pass; Lx, Ly by the next child
        }
    }
}

// FP32 proxy weights of every proxy slot (rows of Rs floats, zeros past R and for slots that are
// only proxy targets)
__global__ void k_w32(int m, int R, int Rs, const int* __restrict__ slot_panel, const int* __restrict__ uslot,
                      const double* __restrict__ Wup, float* __restrict__ W32) {
    const int q = blockIdx.x, u = uslot[slot_panel[q]];
    for (int w = threadIdx.x; w < m * Rs; w += blockDim.x) {
        const int j = w / Rs, r = w % Rs;
        W32[(int64_t)q * m * Rs + w] = (u >= 0 && r < R) ? (float)Wup[((int64_t)u * m + j) * R + r] : 0.f;
    }
}

// ------------------------------------------------------------------------------ 3. far (cluster targets)

// 16 zero bytes: the LDS copies' source for padding lanes
__device__ __attribute__((aligned(16))) float g_zero_row[4];

// LDS DMA of 16 bytes per lane: lane l's bytes land at dst + 16 l (dst is the same for every lane)
__device__ __forceinline__ void lds_copy16(const float* src, float* dst) {
    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}
// the same with the non-temporal hint (cache policy bit NT = 2): data read once per evaluation
__device__ __forceinline__ void lds_copy16_nt(const float* src, float* dst) {
    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 2);
}

constexpr int FAR_KC = 64;  // sources per far-field chunk (two 32-source MFMA steps)

// GEMM per target panel: C[j][c*RT + r] += sum_k A[j][k] B[k][c*RT + r], c = 0..3, with
// B[k][c*RT + r] = w_k^r o_c(y_k).  The column groups are chosen so no FP32 sum cancels: with (e1, e2)
// an orthonormal basis of the tangent plane at the panel centre c_T,
//   o_0 = y.e1, o_1 = y.e2 (the source's tangential coordinates at c_T), o_2 = |y - c_T|^2 / 2,
//   o_3 = 1,
// so y - c_T = o_0 e1 + o_1 e2 - o_2 c_T exactly.  Summing the Cartesian offsets y - c_T instead
// would add the radial parts -o_2 c_T of every source coherently while the tangential parts (the
// gradient) cancel, and the FP32 rounding of that radial sum would swamp the tangential result; here
// the radial sum Phi2 only enters the epilogue through P_z(c_T), at most the panel radius in size.
// 256 threads = 4 waves; wave w owns component c = w for all RT fields (RT/16 column tiles) and all
// MP/16 row tiles.  Sources are taken in chunks of KC = 64; only the 32-source steps that hold
// sources are multiplied (the same steps in every KIND, in source order, so the accumulation order
// of every C element is the same in all of them).
//   KIND 0  split bf16 (hi*hi + hi*lo + lo*hi, 16x16x32 MFMA), kernel matrix evaluated into LDS
//   KIND 1  FP32 MFMA (16x16x4), kernel matrix evaluated into LDS; the FP32 accumulators are added
//           into FP64 sums after every chunk of KC sources (compensated sums: the far sum of a
//           panel runs over thousands of sources whose terms cancel, and one FP32 accumulator
//           over all of them loses several times the products' own rounding)
//   KIND 3  evaluates the split-bf16 kernel matrix as KIND 0 does and stores it in the MFMA operand
//           layout instead of using it: per 32-source step and row tile, the 64 lanes' 8 hi values,
//           then their 8 lo values (2 KB; a step of all row tiles is contiguous, the panel's steps
//           from op_off)
//   KIND 4  evaluates the FP32 kernel matrix as KIND 1 does and stores it in the FP32 MFMA operand
//           order: per 32-source step, for each full row tile and group of 4 sources the 64 lanes'
//           values (256 B), then the last tile's proxy rows only, [group][lk][row] (4 B per matrix
//           entry, the size and step offsets of the split-bf16 operand)
//   KIND 5  copies that operand to LDS a chunk at a time and reproduces KIND 1 (FP64 chunk sums)
//   KIND 6  split bf16 in three parts (v = h + m + l, each a bf16 of the remainder: 24 bits of
//           mantissa, as FP32), six MFMA products (hh, hm, mh, hl, lh, mm) with FP64 chunk sums as
//           KIND 1: FP32 precision at twice the MFMA work of KIND 0 instead of FP32 MFMA's eight
//   KIND 7  evaluates the three-part kernel matrix as KIND 6 does and stores it as KIND 3 stores two
//           parts (6 B per matrix entry)
//   KIND 8  copies that operand to LDS a chunk at a time and reproduces KIND 6
//   KIND 2  copies the stored operand to LDS a chunk at a time and reproduces KIND 0.  At 64
//           proxies (d = 7) it is held to five waves per SIMD (96 registers): five blocks per CU
//           instead of four, 2.21 -> 2.07 ms (larger grids need more LDS than five blocks have, and
//           49 proxies would spill under that register bound).
// The weights w (field rows) are staged transposed, sT[field][source], so a lane reads its 8
// consecutive sources of one field as two 16-byte LDS reads.
template <int KIND, int RT, int MP>
__global__ __launch_bounds__(256, KIND == 2 && MP == 64 ? 5 : 1) void k_far_mfma(int m, int R, int Rs, int r0, const int* __restrict__ xt_panels,
                                                  const int* __restrict__ xt_off, const int* __restrict__ xt_ent,
                                                  const int* __restrict__ slot, const int* __restrict__ lslot,
                                                  const int64_t* __restrict__ start, const int* __restrict__ count,
                                                  const double* __restrict__ centre, const double* __restrict__ px,
                                                  const double* __restrict__ py, const double* __restrict__ pz,
                                                  const float* __restrict__ Qf, const double* __restrict__ prox,
                                                  const float* __restrict__ W32, const int64_t* __restrict__ op_off,
                                                  bf8* __restrict__ fop, float* __restrict__ loc) {
    constexpr int KC = FAR_KC;
    constexpr int NKS_MAX = KC / 32;
    constexpr int NRT = MP / 16;
    constexpr int NCT = RT / 16;
    constexpr int LDA = KC + 8;
    constexpr bool SPLIT = !(KIND == 1 || KIND == 4 || KIND == 5);  // split-bf16 operands (else FP32)
    constexpr int NP = (KIND >= 6) ? 3 : 2;                          // parts of a split value
    constexpr bool STORED = (KIND == 2 || KIND == 5 || KIND == 8);   // stored operand read here
    constexpr bool EVAL = !STORED;                   // kernel matrix evaluated here
    constexpr bool SUM = !(KIND == 3 || KIND == 4 || KIND == 7);  // products accumulated (else stored)
    constexpr int LDT = KC + 4;                      // sT row stride (floats): 16-byte aligned rows
    constexpr int ABYTES = (STORED && SPLIT) ? NKS_MAX * NRT * 64 * NP * 16
                         : STORED ? NKS_MAX * MP * 32 * 4
                                  : (SPLIT ? NP * MP * LDA * 2 : MP * LDA * 4);
    constexpr int EBYTES = SUM ? 4 * MP * 16 * 4 : 16;  // epilogue exchange, 16 fields at a time
    __shared__ __attribute__((aligned(16))) unsigned char smem[(ABYTES > EBYTES ? ABYTES : EBYTES)];
    __shared__ double zx[MP], zy[MP], zz[MP], yx[EVAL ? KC : 1], yy[EVAL ? KC : 1], yz[EVAL ? KC : 1];
    __shared__ __attribute__((aligned(16))) float sT[SUM ? RT : 1][LDT];
    __shared__ __attribute__((aligned(16))) float soT[3][KC];
    const int T = xt_panels[blockIdx.x];
    const int tq = slot[T];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const double cx = centre[3 * T], cy = centre[3 * T + 1], cz = centre[3 * T + 2];
    for (int j = tid; j < MP; j += 256) {
        if (j < m) {
            const double* z = prox + ((int64_t)tq * m + j) * 3;
            zx[j] = z[0]; zy[j] = z[1]; zz[j] = z[2];
        } else {
            zx[j] = cx; zy[j] = cy; zz[j] = cz;
        }
    }
    f4 C[NRT][NCT];
    constexpr bool F64SUM = (KIND == 1 || KIND == 5 || KIND == 6 || KIND == 8);  // FP64 sums of the
                                                                              // per-chunk FP32 accumulators
    double D[F64SUM ? NRT : 1][F64SUM ? NCT : 1][4];
#pragma unroll
    for (int a = 0; a < NRT; ++a)
#pragma unroll
        for (int b = 0; b < NCT; ++b) C[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
    if (F64SUM)
#pragma unroll
        for (int a = 0; a < NRT; ++a)
#pragma unroll
            for (int b = 0; b < NCT; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) D[F64SUM ? a : 0][F64SUM ? b : 0][i] = 0.0;
    const int comp = wv;  // 0, 1: tangential sums A, B; 2: Phi2 (radial); 3: Phi0
    // tangent basis at c_T (in LDS, read after the first barrier): e1 from the coordinate axis least
    // aligned with c_T, e2 = c_T x e1
    __shared__ double ebas[6];
    if (tid == 0) {
        double e1x, e1y, e1z;
        const double ax = fabs(cx), ay = fabs(cy), az = fabs(cz);
        const double ux = (ax <= ay && ax <= az) ? 1.0 : 0.0, uy = (ux == 0.0 && ay <= az) ? 1.0 : 0.0,
                     uz = (ux == 0.0 && uy == 0.0) ? 1.0 : 0.0;
        const double d = ux * cx + uy * cy + uz * cz;
        e1x = ux - d * cx; e1y = uy - d * cy; e1z = uz - d * cz;
        const double inv = 1.0 / sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
        e1x *= inv; e1y *= inv; e1z *= inv;
        ebas[0] = e1x; ebas[1] = e1y; ebas[2] = e1z;
        ebas[3] = cy * e1z - cz * e1y; ebas[4] = cz * e1x - cx * e1z; ebas[5] = cx * e1y - cy * e1x;
    }
    // stored operand: per 32-source step, NRT - 1 full row tiles (64 lanes of each part in turn) and
    // the last tile compact, only its vl rows that hold proxies; 16-byte units per step: 4 m per
    // bf16 part, 8 m for FP32
    const int vl = m - 16 * (NRT - 1), ups = SPLIT ? 4 * NP * m : 8 * m;
    bf8 bzero;
#pragma unroll
    for (int i = 0; i < 8; ++i) bzero[i] = (__bf16)0.f;
    const int kk = tid & 31, jb = tid >> 5;  // kernel-matrix fill: column kk (+32 s), rows jb (+8 s)
    int64_t kst = (STORED || !SUM) ? op_off[blockIdx.x] : 0;  // stored operand: this chunk's first step
    for (int e = xt_off[T]; e < xt_off[T + 1]; ++e) {
        const int S = xt_ent[e] >> 2, ty = xt_ent[e] & 3;
        const bool cp = (ty == 2);
        const int ns = cp ? count[S] : m;
        const int64_t s0 = cp ? start[S] : 0;
        const double* sprox = prox + (int64_t)slot[cp ? 0 : S] * m * 3;
        const float* swg = W32 + (int64_t)slot[cp ? 0 : S] * m * Rs;
        for (int c0 = 0; c0 < ns; c0 += KC) {
            const int nc = min(KC, ns - c0);
            const int nks = (nc + 31) / 32;  // steps holding sources
            __syncthreads();  // the previous chunk's readers are done
            for (int k = tid; k < KC; k += 256) {
                double a = 2.0 * cx, b = 2.0 * cy, c = 2.0 * cz;  // padding: far away, weight 0
                if (k < nc) {
                    if (cp) { const int64_t q = s0 + c0 + k; a = px[q]; b = py[q]; c = pz[q]; }
                    else { const double* y = sprox + (int64_t)(c0 + k) * 3; a = y[0]; b = y[1]; c = y[2]; }
                }
                if (EVAL) { yx[k] = a; yy[k] = b; yz[k] = c; }
                const double ox = a - cx, oy = b - cy, oz = c - cz;
                soT[0][k] = (float)(a * ebas[0] + b * ebas[1] + c * ebas[2]);
                soT[1][k] = (float)(a * ebas[3] + b * ebas[4] + c * ebas[5]);
                soT[2][k] = (float)(0.5 * (ox * ox + oy * oy + oz * oz));
            }
            if (SUM) {  // weight rows: all 16-byte loads first, then transposed LDS stores
                constexpr int V = RT / 4, NV = (KC * V + 255) / 256;
                float4 v[NV];
#pragma unroll
                for (int jv = 0; jv < NV; ++jv) {
                    const int idx = tid + 256 * jv, k = idx / V, c4 = idx % V;
                    v[jv] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (idx < KC * V && k < nc)
                        v[jv] = ld4((cp ? Qf + (s0 + c0 + k) * Rs : swg + (int64_t)(c0 + k) * Rs) + r0 + 4 * c4);
                }
#pragma unroll
                for (int jv = 0; jv < NV; ++jv) {
                    const int idx = tid + 256 * jv, k = idx / V, c4 = idx % V;
                    if (idx < KC * V) {
                        sT[4 * c4][k] = v[jv].x;
                        sT[4 * c4 + 1][k] = v[jv].y;
                        sT[4 * c4 + 2][k] = v[jv].z;
                        sT[4 * c4 + 3][k] = v[jv].w;
                    }
                }
            }
            if (STORED) {  // the chunk's stored operand is one contiguous run: LDS DMA, 1 KB per wave copy
                const int units = nks * ups;
                const bf8* src = fop + kst * ups;
                for (int u0 = 64 * wv; u0 < units; u0 += 256) {
                    const int u = u0 + lane;
                    lds_copy16_nt(u < units ? (const float*)(src + u) : g_zero_row, (float*)((bf8*)smem + u0));
                }
            }
            __syncthreads();
            if (EVAL) {
                // kernel matrix A[j][k] = val(z_j, y_k), zero outside the valid block
                for (int kb = 0; kb < KC; kb += 32) {
                    const int k = kb + kk;
                    const double ykx = yx[k], yky = yy[k], ykz = yz[k];
                    for (int j = jb; j < MP; j += 8) {
                        float v = 0.f;
                        if (j < m && k < nc)
                            v = val32((float)(ykx - zx[j]), (float)(yky - zy[j]), (float)(ykz - zz[j]));
                        if (SPLIT) {
                            __bf16* ap = (__bf16*)smem;  // part p at ap + p MP LDA
                            const __bf16 h = (__bf16)v;
                            const float r1 = v - (float)h;
                            const __bf16 mid = (__bf16)r1;
                            ap[j * LDA + k] = h;
                            ap[MP * LDA + j * LDA + k] = mid;
                            if (NP == 3) ap[2 * MP * LDA + j * LDA + k] = (__bf16)(r1 - (float)mid);
                        } else {
                            ((float*)smem)[j * LDA + k] = v;
                        }
                    }
                }
                __syncthreads();
            }
            if (KIND == 3 || KIND == 7) {
                const __bf16* ap = (const __bf16*)smem;
                for (int idx = tid; idx < nks * ups; idx += 256) {
                    const int ks = idx / ups, u = idx % ups;
                    int row, kg, part;
                    if (u < (NRT - 1) * 64 * NP) {  // full tiles: 64 lanes of each part in turn
                        const int l = u % 64;
                        part = (u / 64) % NP;
                        row = 16 * (u / (64 * NP)) + (l & 15);
                        kg = l >> 4;
                    } else {  // the last tile's vl proxy rows: [part][k-group][row]
                        const int v = u - (NRT - 1) * 64 * NP;
                        part = v / (4 * vl);
                        kg = (v / vl) % 4;
                        row = 16 * (NRT - 1) + v % vl;
                    }
                    const int off = part * MP * LDA + row * LDA + ks * 32 + 8 * kg;
                    __builtin_nontemporal_store(*(const bf8*)(ap + off), fop + kst * ups + idx);
                }
                kst += nks;
                continue;
            }
            if (KIND == 4) {  // 16-byte units of the FP32 operand, 4 floats of consecutive lanes / rows
                const float* af = (const float*)smem;
                for (int idx = tid; idx < nks * ups; idx += 256) {
                    const int ks = idx / ups, u = idx % ups;
                    f4 v;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int f = 4 * u + q;  // float within the step
                        int row, k;
                        if (f < 512 * (NRT - 1)) {  // full tile f / 512, group (f % 512) / 64, lane f % 64
                            const int l = f % 64;
                            row = 16 * (f / 512) + (l & 15);
                            k = ks * 32 + 4 * ((f % 512) / 64) + (l >> 4);
                        } else {  // the last tile's vl proxy rows: [group][lk][row]
                            const int g = f - 512 * (NRT - 1), r = g % (4 * vl);
                            row = 16 * (NRT - 1) + r % vl;
                            k = ks * 32 + 4 * (g / (4 * vl)) + r / vl;
                        }
                        v[q] = af[row * LDA + k];
                    }
                    __builtin_nontemporal_store(v, (f4*)(fop + kst * ups) + idx);
                }
                kst += nks;
                continue;
            }
#pragma unroll
            for (int ks = 0; ks < NKS_MAX; ++ks) {
                if (ks >= nks) break;
                if (SPLIT) {
                    bf8 bh[NCT], bl[NCT], bm[NCT];  // bm: the middle part (NP == 3)
                    {
                        const float4 o0 = (comp < 3) ? *(const float4*)&soT[comp][ks * 32 + 8 * lk]
                                                     : make_float4(1.f, 1.f, 1.f, 1.f);
                        const float4 o1 = (comp < 3) ? *(const float4*)&soT[comp][ks * 32 + 8 * lk + 4]
                                                     : make_float4(1.f, 1.f, 1.f, 1.f);
                        const float o[8] = {o0.x, o0.y, o0.z, o0.w, o1.x, o1.y, o1.z, o1.w};
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) {
                            const float* wr = &sT[16 * ct + lr][ks * 32 + 8 * lk];
                            const float4 w0 = *(const float4*)wr, w1 = *(const float4*)(wr + 4);
                            const float w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                const float v = w[i] * o[i];
                                const __bf16 h = (__bf16)v;
                                const float r1 = v - (float)h;
                                bh[ct][i] = h;
                                if (NP == 3) {
                                    const __bf16 mid = (__bf16)r1;
                                    bm[ct][i] = mid;
                                    bl[ct][i] = (__bf16)(r1 - (float)mid);
                                } else {
                                    bl[ct][i] = (__bf16)r1;
                                }
                            }
                        }
                    }
                    const __bf16* ap = (const __bf16*)smem;
#pragma unroll
                    for (int rt = 0; rt < NRT; ++rt) {
                        bf8 a_h, a_l, a_m;  // a_m: the middle part (NP == 3); a_l the last part
                        if (STORED) {
                            const bf8* fa = (const bf8*)smem + ks * ups + rt * 64 * NP;
                            if (rt < NRT - 1) {
                                a_h = fa[lane];
                                if (NP == 3) { a_m = fa[64 + lane]; a_l = fa[128 + lane]; }
                                else a_l = fa[64 + lane];
                            } else {  // compact last tile: rows past the proxies are zero
                                const bool ok = lr < vl;
                                const int q = ok ? lk * vl + lr : 0;
                                a_h = ok ? fa[q] : bzero;
                                if (NP == 3) {
                                    a_m = ok ? fa[4 * vl + q] : bzero;
                                    a_l = ok ? fa[8 * vl + q] : bzero;
                                } else {
                                    a_l = ok ? fa[4 * vl + q] : bzero;
                                }
                            }
                        } else {
                            const int off = (16 * rt + lr) * LDA + ks * 32 + 8 * lk;
                            a_h = *(const bf8*)(ap + off);
                            if (NP == 3) {
                                a_m = *(const bf8*)(ap + MP * LDA + off);
                                a_l = *(const bf8*)(ap + 2 * MP * LDA + off);
                            } else {
                                a_l = *(const bf8*)(ap + MP * LDA + off);
                            }
                        }
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) {
                            if (NP == 3) {  // the six products of order up to 2^-16: hh, hm, mh, hl, lh, mm
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bh[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bm[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_m, bh[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bl[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, bh[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_m, bm[ct], C[rt][ct], 0, 0, 0);
                            } else {
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bh[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bl[ct], C[rt][ct], 0, 0, 0);
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, bh[ct], C[rt][ct], 0, 0, 0);
                            }
                        }
                    }
                } else {
#pragma unroll
                    for (int k4 = 0; k4 < 8; ++k4) {  // the step's 32 sources, 4 per FP32 MFMA
                        const int k = ks * 32 + k4 * 4 + lk;
                        const float o = (comp < 3) ? soT[comp][k] : 1.f;
                        float bv[NCT];
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) bv[ct] = sT[16 * ct + lr][k] * o;
                        const float* af = (const float*)smem;
#pragma unroll
                        for (int rt = 0; rt < NRT; ++rt) {
                            float a;
                            if (KIND == 5) {  // the stored operand's order (KIND 4)
                                const float* fa = af + ks * 4 * ups;
                                if (rt < NRT - 1) a = fa[rt * 512 + k4 * 64 + lane];
                                else a = lr < vl ? fa[(NRT - 1) * 512 + k4 * 4 * vl + lk * vl + lr] : 0.f;
                            } else {
                                a = af[(16 * rt + lr) * LDA + k];
                            }
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct)
                                C[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv[ct], C[rt][ct], 0, 0, 0);
                        }
                    }
                }
            }
            if (STORED) kst += nks;
            if (F64SUM) {
#pragma unroll
                for (int a = 0; a < NRT; ++a)
#pragma unroll
                    for (int b = 0; b < NCT; ++b) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) D[F64SUM ? a : 0][F64SUM ? b : 0][i] += (double)C[a][b][i];
                        C[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
                    }
            }
        }
    }
    if (!SUM) return;
    // epilogue, 16 fields at a time: exchange through LDS, then in FP64
    // T_j = P_z (A e1 + B e2 - Phi2 c_T - Phi0 b_j), b_j = z_j - c_T
    const int ls = lslot[T];
    float* E = (float*)smem;  // [4][MP][16] (KIND 1: the FP64 sums rounded once)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
        __syncthreads();
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                E[(comp * MP + 16 * rt + 4 * lk + i) * 16 + lr] =
                    F64SUM ? (float)D[F64SUM ? rt : 0][F64SUM ? ct : 0][i] : C[rt][ct][i];
        __syncthreads();
        for (int idx = tid; idx < m * 16; idx += 256) {
            const int j = idx / 16, rr = idx % 16, r = 16 * ct + rr;
            float* o = loc + ((int64_t)ls * 3 * m + j) * Rs + r0 + r;
            if (r0 + r >= R) {  // padding columns of the tile
                o[0] = o[(int64_t)m * Rs] = o[(int64_t)2 * m * Rs] = 0.f;
                continue;
            }
            const double ta = E[(0 * MP + j) * 16 + rr], tb = E[(1 * MP + j) * 16 + rr];
            const double p2 = E[(2 * MP + j) * 16 + rr], p0 = E[(3 * MP + j) * 16 + rr];
            const double bx = zx[j] - cx, by = zy[j] - cy, bz = zz[j] - cz;
            const double vx = ta * ebas[0] + tb * ebas[3] - p2 * cx - p0 * bx;
            const double vy = ta * ebas[1] + tb * ebas[4] - p2 * cy - p0 * by;
            const double vz = ta * ebas[2] + tb * ebas[5] - p2 * cz - p0 * bz;
            const double q = vx * zx[j] + vy * zy[j] + vz * zz[j];
            o[0] = (float)(vx - q * zx[j]);
            o[(int64_t)m * Rs] = (float)(vy - q * zy[j]);
            o[(int64_t)2 * m * Rs] = (float)(vz - q * zz[j]);
        }
    }
}

// ------------------------------------------------------------------------------ 4. l2l

// One block per panel A of a level: loc[A] = (own ? loc[A] : 0) + parent's interpolant at A's grid.
// Child grids sit on one half of the parent's interval in each direction, so the interpolation is
// separable: tmp[a][j] = sum_i Lx[a][i] P[i][j], out[a][b] = sum_j Ly[b][j] tmp[a][j].  One thread
// per column (component, field) holds all ND x ND outputs and reads each parent value once, one
// grid row j at a time (consecutive lanes: consecutive fields, so every row read is coalesced).
template <int ND>
__global__ __launch_bounds__(256) void k_l2l(int R, int Rs, const int* __restrict__ panels, const int* __restrict__ parent,
                                             const int* __restrict__ lslot, const unsigned char* __restrict__ own,
                                             const double* __restrict__ bounds, float* __restrict__ loc) {
    constexpr int M = ND * ND;
    __shared__ float Lx[ND][ND], Ly[ND][ND];
    __shared__ double cs[ND];  // grid nodes on [-1, 1]
    const int A = panels[blockIdx.x], Pa = parent[A];
    const double* ba = bounds + 6 * (int64_t)A;
    const double* bp = bounds + 6 * (int64_t)Pa;
    if (threadIdx.x < ND) cs[threadIdx.x] = bli_cheb(ND - 1, threadIdx.x);
    __syncthreads();
    if (threadIdx.x < 2 * ND) {
        const int d = threadIdx.x / ND, a = threadIdx.x % ND;  // d 0: xi, 1: eta
        double nodes[ND], f[ND];
#pragma unroll
        for (int t = 0; t < ND; ++t) nodes[t] = bli_node_cs(bp[3 * d], bp[3 * d + 2], cs[t]);
        const double x = bli_node_cs(ba[3 * d], ba[3 * d + 2], cs[a]);
        bli_basis_1d(x, nodes, ND - 1, f);
#pragma unroll
        for (int t = 0; t < ND; ++t) {
            if (d == 0) Lx[a][t] = (float)f[t];
            else Ly[a][t] = (float)f[t];
        }
    }
    __syncthreads();
    const float* src = loc + (int64_t)lslot[Pa] * 3 * M * Rs;
    float* dst = loc + (int64_t)lslot[A] * 3 * M * Rs;
    const bool keep = own[A] != 0;
    for (int col = threadIdx.x; col < 3 * R; col += blockDim.x) {
        const int c = col / R, r = col % R;
        const float* sp = src + (int64_t)c * M * Rs + r;
        float* o = dst + (int64_t)c * M * Rs + r;
        float out[ND][ND];  // [a][b]
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
            for (int bb = 0; bb < ND; ++bb) out[a][bb] = keep ? o[(int64_t)(a + bb * ND) * Rs] : 0.f;
#pragma unroll 1
        for (int j = 0; j < ND; ++j) {  // rolled: one row of ND values in flight at a time
            float p[ND];
#pragma unroll
            for (int i = 0; i < ND; ++i) p[i] = sp[(int64_t)(i + j * ND) * Rs];
#pragma unroll
            for (int a = 0; a < ND; ++a) {
                float t = 0.f;
#pragma unroll
                for (int i = 0; i < ND; ++i) t = fmaf(Lx[a][i], p[i], t);
#pragma unroll
                for (int bb = 0; bb < ND; ++bb) out[a][bb] = fmaf(Ly[bb][j], t, out[a][bb]);
            }
        }
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
            for (int bb = 0; bb < ND; ++bb) o[(int64_t)(a + bb * ND) * Rs] = out[a][bb];
    }
}

// ------------------------------------------------------------------------------ 5. near

// L2P basis of one target in one direction (d 0: xi, 1: eta) on the panel's grid, as FP32, into
// sf[d][tt][0..ND-1] (sf rows of 12 floats); cs[i] = bli_cheb(ND - 1, i)
template <int ND>
__device__ __forceinline__ void l2p_basis(float* sf, const double* b, double x, int d, int tt, const double* cs) {
    double nodes[ND], f[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) nodes[i] = bli_node_cs(b[3 * d], b[3 * d + 2], cs[i]);
    bli_basis_1d(x, nodes, ND - 1, f);
#pragma unroll
    for (int i = 0; i < ND; ++i) sf[(d * 32 + tt) * 12 + i] = (float)f[i];
}

// One wave per work item: NRT = 2 for items of 17..32 targets of a leaf, NRT = 1 for items of at
// most 16 (the last item of a leaf, launched separately).  Lane l holds rows lr + 16 rt of the A
// tiles (targets) and source lk of each group of 4; the C tiles hold targets 16 rt + 4 lk + i,
// field 16 ct + lr.
//
// MODE 0 evaluates the kernel values on the fly.  MODE 1 evaluates them the same way and stores
// them instead of using them: per lane per step of 4 sources, (east, north) of each of its NRT
// target rows (2 NRT floats; a step is one coalesced record of 64 lanes), the item's steps
// contiguous from op_off[item] (in floats).  MODE 2 reads the stored values, with non-temporal loads:
// read once per evaluation, they should not evict the field rows every work item re-reads from L2;
// it copies the field rows into LDS by DMA rather than through registers.  All three walk the same
// entries in the same order, so MODE 2 reproduces MODE 0 bit for bit, and a row's values do not
// depend on NRT.  PP pairs whose FP32 chord squared is below the leaf's cut2 (close pairs) get the
// value 0 here (and in the stored operator): k_close sums them in FP64.
template <int RT, int MODE, int NRT>
__global__ __launch_bounds__(64, MODE == 2 ? 3 : 1)
void k_near(int nx, int m, int R, int Rs, int r0, const int2* __restrict__ work,
            const int* __restrict__ parent, const int* __restrict__ pt_off, const int* __restrict__ pt_ent,
            const int* __restrict__ slot, const int64_t* __restrict__ start, const int* __restrict__ count,
            const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
            const float* __restrict__ Qf, const double* __restrict__ prox, const float* __restrict__ W32,
            const int64_t* __restrict__ op_off, float* __restrict__ op, double* __restrict__ Ge,
            double* __restrict__ Gn, int nd, const int* __restrict__ lslot, const double* __restrict__ bounds,
            const int* __restrict__ perm, const double* __restrict__ xi, const double* __restrict__ eta,
            const float* __restrict__ loc, const unsigned char* __restrict__ own, const float* __restrict__ cut2) {
    constexpr int NCT = RT / 16, SCH = 64, NB = 4;  // NB: steps per batch of stored values
    constexpr bool EVAL = (MODE != 2), SUM = (MODE != 1);
    typedef float rec_t __attribute__((ext_vector_type(2 * NRT)));  // one lane's values of one step
    __shared__ double sx[SCH], sy[SCH], sz[SCH];
    constexpr int LDQ = RT + 16;  // 16-byte aligned rows; the two rows a half-wave reads use disjoint banks
    __shared__ __attribute__((aligned(16))) float sq[MODE == 0 ? SCH : 1][LDQ];
    // MODE 2 copies the rows with LDS DMA (no registers): unpadded rows of RT floats, and for RT = 32
    // the two 16-field halves of odd rows swapped, so the two rows a half-wave reads use disjoint banks
    __shared__ __attribute__((aligned(16))) float sqd[MODE == 2 ? SCH * RT : 1];
    __shared__ double cheb[12];  // L2P grid nodes on [-1, 1]
    const int item = xcd_item(nx);
    if (item < 0) return;
    const int2 wk = work[item];
    const int L = wk.x, t0 = wk.y;
    const int cnt = count[L];
    const int64_t st = start[L];
    const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
    const float cutL = (EVAL && cut2) ? cut2[L] : -1.0f;
    rec_t* __restrict__ opq = (rec_t*)(op + (MODE != 0 ? op_off[item] : 0)) + lane;
    double tx[NRT], ty[NRT], tz[NRT];
    float fx[NRT], fy[NRT], fz[NRT], fa[NRT], fic[NRT];
    if (EVAL) {
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) {
            int t = t0 + lr + 16 * rt;
            if (t >= cnt) t = t0;
            const int64_t k = st + t;
            tx[rt] = px[k]; ty[rt] = py[k]; tz[rt] = pz[k];
            Tgt T = make_tgt(tx[rt], ty[rt], tz[rt]);
            fx[rt] = (float)tx[rt]; fy[rt] = (float)ty[rt]; fz[rt] = (float)tz[rt];
            fa[rt] = (float)T.a; fic[rt] = (float)T.inv_c;
        }
    }
    double De[NRT][NCT][4], Dn[NRT][NCT][4];
#pragma unroll
    for (int a = 0; a < NRT; ++a)
#pragma unroll
        for (int b = 0; b < NCT; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) De[a][b][i] = Dn[a][b][i] = 0.0;
    for (int A = L; A >= 0; A = parent[A]) {
        for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
            const int S = pt_ent[e] >> 2, kind = pt_ent[e] & 3;
            const bool pp = (kind == 0);
            const int ns = pp ? count[S] : m;
            const int64_t s0 = pp ? start[S] : 0;
            const double* sprox = prox + (int64_t)slot[pp ? 0 : S] * m * 3;
            const float* sw = W32 + (int64_t)slot[pp ? 0 : S] * m * Rs;
            f4 Ce[NRT][NCT], Cn[NRT][NCT];
#pragma unroll
            for (int a = 0; a < NRT; ++a)
#pragma unroll
                for (int b = 0; b < NCT; ++b) { Ce[a][b] = (f4){0.f, 0.f, 0.f, 0.f}; Cn[a][b] = Ce[a][b]; }
            for (int c0 = 0; c0 < ns; c0 += SCH) {
                const int nc = min(SCH, ns - c0);
                rec_t va[NB];  // MODE 2: the chunk's first batch of stored values, in flight during staging
                if (MODE == 2) {
#pragma unroll
                    for (int i = 0; i < NB; ++i)
                        if (4 * i < nc) va[i] = __builtin_nontemporal_load(opq + 64 * i);
                }
                __syncthreads();
                if (EVAL) {
                    const int t = lane;
                    double a = 10.0, b = 0.0, c = 0.0;  // padding: far away, weight 0
                    if (t < nc) {
                        if (pp) { const int64_t q = s0 + c0 + t; a = px[q]; b = py[q]; c = pz[q]; }
                        else { const double* y = sprox + (int64_t)(c0 + t) * 3; a = y[0]; b = y[1]; c = y[2]; }
                    }
                    sx[t] = a; sy[t] = b; sz[t] = c;
                }
                if (MODE == 2) {  // weight rows: lane l's 16 bytes of copy j land at sqd[256 j + 4 l]
                    constexpr int V = RT / 4, RPL = 64 / V, NLD = SCH / RPL;
                    const float* base = pp ? Qf + (s0 + c0) * Rs : sw + (int64_t)c0 * Rs;
#pragma unroll
                    for (int j = 0; j < NLD; ++j) {
                        if (j * RPL >= nc) break;  // rows up to the next multiple of RPL (>= 4) are written
                        const int r = j * RPL + lane / V, p = lane % V;
                        const int c4 = (RT == 32) ? (p ^ ((r & 1) << 2)) : p;
                        const float* src = r < nc ? base + (int64_t)r * Rs + r0 + 4 * c4 : g_zero_row;
                        lds_copy16(src, sqd + j * 256);
                    }
                } else if (SUM) {  // weight rows: all 16-byte loads first, then the LDS stores
                    constexpr int V = RT / 4, NV = SCH * V / 64;
                    float4 v[NV];
#pragma unroll
                    for (int j = 0; j < NV; ++j) {
                        const int idx = lane + 64 * j, t = idx / V, c4 = idx % V;
                        v[j] = t < nc ? ld4((pp ? Qf + (s0 + c0 + t) * Rs : sw + (int64_t)(c0 + t) * Rs) + r0 + 4 * c4)
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
                    }
#pragma unroll
                    for (int j = 0; j < NV; ++j) {
                        const int idx = lane + 64 * j, t = idx / V, c4 = idx % V;
                        *(float4*)&sq[t][4 * c4] = v[j];
                    }
                }
                __syncthreads();
                if (MODE == 2) {
                    // stored values: batches of NB steps; the next batch's loads are issued before the
                    // current batch's products.  Small batches keep the kernel at 168 registers, three
                    // waves per SIMD, which hides more of the memory latency than deeper batches do.
                    for (int kb = 0; kb < nc; kb += 4 * NB) {
                        rec_t v[NB], vn[NB];
#pragma unroll
                        for (int i = 0; i < NB; ++i) v[i] = va[i];
#pragma unroll
                        for (int i = 0; i < NB; ++i)
                            if (kb + 4 * NB + 4 * i < nc) vn[i] = __builtin_nontemporal_load(opq + 64 * (NB + i));
#pragma unroll
                        for (int i = 0; i < NB; ++i) {
                            if (kb + 4 * i >= nc) break;
                            const int s = kb + 4 * i + lk;
                            float b[NCT];
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct)
                                b[ct] = sqd[s * RT + 16 * ((RT == 32) ? (ct ^ (s & 1)) : ct) + lr];
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
                                for (int rt = 0; rt < NRT; ++rt) {
                                    Ce[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[i][2 * rt], b[ct], Ce[rt][ct], 0, 0, 0);
                                    Cn[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[i][2 * rt + 1], b[ct], Cn[rt][ct], 0, 0, 0);
                                }
                        }
                        opq += 64 * min(NB, (nc - kb + 3) / 4);
#pragma unroll
                        for (int i = 0; i < NB; ++i) va[i] = vn[i];
                    }
                    continue;
                }
                for (int ks = 0; ks < nc; ks += 4) {
                    const int s = ks + lk;
                    float ae[NRT], an[NRT];
                    {
                        const double ssx = sx[s], ssy = sy[s], ssz = sz[s];
#pragma unroll
                        for (int rt = 0; rt < NRT; ++rt) {
                            const float dx = (float)(ssx - tx[rt]), dy = (float)(ssy - ty[rt]),
                                        dz = (float)(ssz - tz[rt]);
                            const float c2 = chord2_f32(dx, dy, dz);
                            const float f = (pp && c2 < cutL) ? 0.0f : val32_c2(c2) * fic[rt];
                            ae[rt] = (fx[rt] * dy - fy[rt] * dx) * f;
                            an[rt] = (dz * fa[rt] - fz[rt] * fmaf(fx[rt], dx, fy[rt] * dy)) * f;
                        }
                        if (MODE == 1) {
                            rec_t r;
#pragma unroll
                            for (int rt = 0; rt < NRT; ++rt) { r[2 * rt] = ae[rt]; r[2 * rt + 1] = an[rt]; }
                            __builtin_nontemporal_store(r, opq);
                            opq += 64;
                        }
                    }
                    if (SUM) {
                        float b[NCT];
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) b[ct] = sq[s][16 * ct + lr];
#pragma unroll
                        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) {
                                Ce[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(ae[rt], b[ct], Ce[rt][ct], 0, 0, 0);
                                Cn[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(an[rt], b[ct], Cn[rt][ct], 0, 0, 0);
                            }
                    }
                }
            }
            // one source panel done: promote its FP32 sums into the FP64 accumulators
            if (SUM) {
#pragma unroll
                for (int a = 0; a < NRT; ++a)
#pragma unroll
                    for (int b = 0; b < NCT; ++b)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            De[a][b][i] += (double)Ce[a][b][i];
                            Dn[a][b][i] += (double)Cn[a][b][i];
                        }
            }
        }
    }
    if (!SUM) return;
    // L2P: [targets x m] basis products (FP32 products of the FP64 1-D bases) times an [m x 3 RT]
    // expansion (FP32 MFMA, m padded to a multiple of 4 with zero basis values), for the parent's
    // expansion on the parent's grid (the interpolant the parent pushes down is a degree-d
    // polynomial, which the leaf's grid would reproduce exactly) and for the leaf's own expansion
    // when the leaf is a cluster target itself; each (target, field) tangential vector is projected
    // on the target's east/north basis and added to the FP64 near sums.
    const int Pa = parent[L];
    const int lsp = (loc && Pa >= 0) ? lslot[Pa] : -1, lso = (loc && own[L]) ? lslot[L] : -1;
    float* sf = (MODE == 2) ? sqd : &sq[0][0];  // [2][32][12] floats: the row buffer is free now
    // one interpolant (expansion slot ls on panel B's grid), projected and added to the FP64 sums
    auto l2p = [&](const int ls, const int B) {
        __syncthreads();
        {  // lanes 0..31: xi basis of target t0 + lane on B's grid; lanes 32..63: eta basis
            const int d = lane >> 5, tt = lane & 31;
            int t = t0 + tt;
            if (t >= cnt) t = t0;  // padding rows repeat a valid target; their results are not written
            const int p = perm[st + t];
            const double* b = bounds + 6 * (int64_t)B;
            const double x = d ? eta[p] : xi[p];
            switch (nd) {
                case 5: l2p_basis<5>(sf, b, x, d, tt, cheb); break;
                case 6: l2p_basis<6>(sf, b, x, d, tt, cheb); break;
                case 7: l2p_basis<7>(sf, b, x, d, tt, cheb); break;
                case 8: l2p_basis<8>(sf, b, x, d, tt, cheb); break;
                case 9: l2p_basis<9>(sf, b, x, d, tt, cheb); break;
                case 10: l2p_basis<10>(sf, b, x, d, tt, cheb); break;
                default: l2p_basis<11>(sf, b, x, d, tt, cheb); break;
            }
        }
        __syncthreads();
        const float* lp = loc + (int64_t)ls * 3 * m * Rs + r0;
        f4 V[3][NRT][NCT];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) V[c][rt][ct] = (f4){0.f, 0.f, 0.f, 0.f};
        // groups of LG steps: every expansion load of a group is issued before its products, so an
        // item waits for one round trip per group instead of one per step (two steps for two row
        // tiles keep the kernel at 168 registers); steps past m multiply zeros, which leaves the sums
        // unchanged
        constexpr int LG = (NRT == 2) ? 2 : 4;
        for (int kg = 0; kg < (m + 3) / 4; kg += LG) {
            float a[LG][NRT], bv[LG][3][NCT];
#pragma unroll
            for (int u = 0; u < LG; ++u) {
                const int k = 4 * (kg + u) + lk;
                const bool kv = k < m;
                const int i = k % nd, j = k / nd;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct)
                        bv[u][c][ct] = kv ? lp[((int64_t)c * m + k) * Rs + 16 * ct + lr] : 0.f;
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt)
                    a[u][rt] = kv ? sf[(lr + 16 * rt) * 12 + i] * sf[(32 + lr + 16 * rt) * 12 + j] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < LG; ++u)
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct)
                            V[c][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][rt], bv[u][c][ct], V[c][rt][ct], 0, 0, 0);
        }
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = t0 + 16 * rt + 4 * lk + i;
                if (t >= cnt) continue;
                const int64_t k = st + t;
                const Tgt T = make_tgt(px[k], py[k], pz[k]);
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) {
                    double ge, gn;
                    project_en(T, V[0][rt][ct][i], V[1][rt][ct][i], V[2][rt][ct][i], ge, gn);
                    De[rt][ct][i] += ge;
                    Dn[rt][ct][i] += gn;
                }
            }
    };
    if ((lsp >= 0 || lso >= 0) && lane < nd) cheb[lane] = bli_cheb(nd - 1, lane);  // read after l2p's barrier
    if (lsp >= 0) l2p(lsp, Pa);
    if (lso >= 0) l2p(lso, L);
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = t0 + 16 * rt + 4 * lk + i, r = 16 * ct + lr;
                if (t < cnt && t < t0 + 16 * NRT && r0 + r < R) {
                    Ge[(st + t) * R + r0 + r] = De[rt][ct][i];
                    Gn[(st + t) * R + r0 + r] = Dn[rt][ct][i];
                }
            }
}

// ------------------------------------------------------------------------------ 5b. close pairs

// Close-pair list: for every target point k (tree order) of a near work item, the PP sources on its
// leaf's root path whose FP32 chord squared (chord2_f32 of the same FP64 differences k_near rounds)
// is below the leaf's cut2, in k_near's walk order, the self pair left out.  Two items per 64-thread
// block, one lane per target.  FILL = false writes the count of target k to out[k]; FILL = true
// writes its sources to out[off[k] ...].
template <bool FILL>
__global__ __launch_bounds__(64) void k_close_list(int nx, const int2* __restrict__ work, const int* __restrict__ parent,
                                                   const int* __restrict__ pt_off, const int* __restrict__ pt_ent,
                                                   const int64_t* __restrict__ start, const int* __restrict__ count,
                                                   const double* __restrict__ px, const double* __restrict__ py,
                                                   const double* __restrict__ pz, const float* __restrict__ cut2,
                                                   const int* __restrict__ off, int* __restrict__ out) {
    const int item = 2 * blockIdx.x + (threadIdx.x >> 5);
    if (item >= nx) return;
    const int2 wk = work[item];
    const int L = wk.x, t = wk.y + (threadIdx.x & 31);
    if (t >= count[L]) return;
    const int64_t k = start[L] + t;
    const double tx = px[k], ty = py[k], tz = pz[k];
    const float cut = cut2[L];
    int c = 0, pos = FILL ? off[k] : 0;
    for (int A = L; A >= 0; A = parent[A])
        for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
            if ((pt_ent[e] & 3) != 0) continue;
            const int S = pt_ent[e] >> 2;
            const int64_t s0 = start[S], s1 = s0 + count[S];
            for (int64_t s = s0; s < s1; ++s) {
                const float c2 = chord2_f32((float)(px[s] - tx), (float)(py[s] - ty), (float)(pz[s] - tz));
                if (s != k && c2 < cut) {
                    if (FILL) out[pos++] = (int)s;
                    else ++c;
                }
            }
        }
    if (!FILL) out[k] = c;
}

// FP64 kernel factors of the close pairs, once per plan: one thread per target point k (tree order)
// writes (east, north) of the cancellation-free kernel for each of its close sources.
__global__ __launch_bounds__(64) void k_close_vals(int64_t n, const int* __restrict__ off, const int* __restrict__ src,
                                                   const double* __restrict__ px, const double* __restrict__ py,
                                                   const double* __restrict__ pz, double2* __restrict__ val) {
    const int64_t k = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (k >= n) return;
    const int e0 = off[k], e1 = off[k + 1];
    if (e0 == e1) return;
    const Tgt T = make_tgt(px[k], py[k], pz[k]);
    for (int e = e0; e < e1; ++e) {
        const int s = src[e];
        double ge, gn;
        gfunc_stable(T, px[s], py[s], pz[s], ge, gn);
        val[e] = make_double2(ge, gn);
    }
}

// Close pairs in FP64: G[k][r] += sum over k's close pairs of (east, north) * q[s][r], the FP64 kernel
// factors times the source's packed field row (area * ssh rounded to FP32 once, as every other
// near-field product uses it), summed in FP64 and added to the near kernel's sums Ge, Gn [n][R]
// (runs after k_near on the same stream).  Four fields per lane, RT / 4 lanes per target (their
// 16-byte row pieces together read the row coalesced), 256 / (RT / 4) targets per block.
template <int RT>
__global__ __launch_bounds__(256) void k_close(int64_t n, int R, int Rs, int r0, const int* __restrict__ off,
                                               const int* __restrict__ src, const double2* __restrict__ val,
                                               const float* __restrict__ Qf, double* __restrict__ Ge,
                                               double* __restrict__ Gn) {
    constexpr int LPT = RT / 4;  // lanes per target
    const int64_t k = (int64_t)blockIdx.x * (256 / LPT) + threadIdx.x / LPT;
    const int f = 4 * (threadIdx.x % LPT);
    if (k >= n) return;
    const int e0 = off[k], e1 = off[k + 1];
    if (e0 == e1) return;
    double ae[4] = {0.0, 0.0, 0.0, 0.0}, an[4] = {0.0, 0.0, 0.0, 0.0};
    // groups of U pairs: every load of a group is issued before its products (the pairs' row loads
    // depend on their source ids), so a target waits for two round trips per group, not per pair;
    // pairs past e1 repeat the target's last pair with zero factors
    constexpr int U = 4;
    for (int e = e0; e < e1; e += U) {
        int s[U];
        double2 v[U];
        float4 q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int eu = min(e + u, e1 - 1);
            s[u] = src[eu];
            v[u] = val[eu];
            if (e + u >= e1) v[u] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) q[u] = *(const float4*)(Qf + (int64_t)s[u] * Rs + r0 + f);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ae[0] = fma(v[u].x, (double)q[u].x, ae[0]);
            ae[1] = fma(v[u].x, (double)q[u].y, ae[1]);
            ae[2] = fma(v[u].x, (double)q[u].z, ae[2]);
            ae[3] = fma(v[u].x, (double)q[u].w, ae[3]);
            an[0] = fma(v[u].y, (double)q[u].x, an[0]);
            an[1] = fma(v[u].y, (double)q[u].y, an[1]);
            an[2] = fma(v[u].y, (double)q[u].z, an[2]);
            an[3] = fma(v[u].y, (double)q[u].w, an[3]);
        }
    }
    double* ge_row = Ge + k * R + r0 + f;
    double* gn_row = Gn + k * R + r0 + f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (r0 + f + i < R) {
            ge_row[i] += ae[i];
            gn_row[i] += an[i];
        }
}

// ------------------------------------------------------------------------------ 6. out

// The reverse transpose: a block gathers the tree-order rows of OB consecutive caller points (32
// lanes read one row segment) into LDS and writes every field's outputs coalesced along points.
constexpr int OB = 64;  // 34 KB of LDS per block: four blocks per CU
__global__ __launch_bounds__(256) void k_out_mix(int64_t n, int R, const int* __restrict__ iperm,
                                                const double* __restrict__ Ge, const double* __restrict__ Gn,
                                                double* __restrict__ gx, double* __restrict__ gy, int accumulate) {
    __shared__ double te[OB][33], tn[OB][33];
    __shared__ int64_t src[OB];
    const int64_t p0 = (int64_t)blockIdx.x * OB;
    const int np = (int)min((int64_t)OB, n - p0);
    const int t = threadIdx.x;
    if (t < np) src[t] = (int64_t)iperm[p0 + t] * R;
    for (int r0 = 0; r0 < R; r0 += 32) {
        const int nr = min(32, R - r0);
        __syncthreads();
        const int c = t & 31;
        for (int pt = t >> 5; pt < np; pt += 256 / 32)
            if (c < nr) {
                te[pt][c] = Ge[src[pt] + r0 + c];
                tn[pt][c] = Gn[src[pt] + r0 + c];
            }
        __syncthreads();
        const int pt = t % OB;
        if (pt < np)
            for (int cc = t / OB; cc < nr; cc += 256 / OB) {
                double* ox = gx + (int64_t)(r0 + cc) * n + p0 + pt;
                double* oy = gy + (int64_t)(r0 + cc) * n + p0 + pt;
                if (accumulate) { *ox += te[pt][cc]; *oy += tn[pt][cc]; }
                else { *ox = te[pt][cc]; *oy = tn[pt][cc]; }
            }
    }
}

// ------------------------------------------------------------------------------ driver

template <int ND, int RT>
static void run_mixed(sal_geom g, const Plan* pl, const MixLists* ml, sal_work w, const sal_eval_opts& o, int R,
                      hipStream_t s) {
    constexpr int M = ND * ND;
    constexpr int MP = ((M + 15) / 16) * 16;
    const int Rs = row_stride(R, RT);
    if (ml->nU > 0) {
        {
            ProfScope ps("p2m_mix", s);
            for (int r0 = 0; r0 < R; r0 += RT)
                k_p2m_up<ND, RT><<<ml->nU, 128, 0, s>>>(R, Rs, r0, ml->u_panels.p, ml->p2m_off.p, ml->p2m_src.p,
                                                        g->d_start.p, g->d_count.p, g->d_bounds.p, g->perm.p, g->xi.p,
                                                        g->eta.p, w->Qf.p, w->wpart.p);
        }
        ProfScope ps("m2m_mix", s);
        for (size_t L = 0; L + 1 < ml->m2m_lo.size(); ++L) {
            const int a = ml->m2m_lo[L], b = ml->m2m_lo[L + 1];
            if (b > a)
                k_m2m<ND><<<b - a, M2M_THREADS, 0, s>>>(R, ml->m2m_panels.p + a, g->d_child1.p, ml->uslot.p, g->d_bounds.p,
                                                w->wpart.p);
        }
        SAL_HIP(hipGetLastError());
    }
    if (pl->Q > 0) {
        ProfScope ps("w32_mix", s);
        k_w32<<<pl->Q, 256, 0, s>>>(M, R, Rs, pl->slot_panel.p, ml->uslot.p, w->wpart.p, w->W32.p);
    }
    // the FP32 far field (FP64 chunk sums) evaluated on the fly runs in tiles of 16 fields: its
    // accumulators would halve the occupancy at 32; streaming a stored operator it takes the whole
    // RHS tile, so the operator is read once (the tiles cover the whole row stride, so every
    // expansion column the near kernel reads is written)
    const bool far32 = (o.m2l == SAL_M2L_FP32_MFMA), far3 = (o.m2l == SAL_M2L_BF16X6),
               ft16 = (far32 || far3) && !o.store;
    for (int r0 = 0; r0 < (ft16 ? Rs : R); r0 += ft16 ? 16 : RT) {
        if (pl->n_xt > 0) {
            ProfScope ps("far_mix", s);
#define SAL_FAR(KIND, FT, OP)                                                                                 \
    k_far_mfma<KIND, FT, MP><<<pl->n_xt, 256, 0, s>>>(M, R, Rs, r0, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p,  \
                                                      pl->slot.p, ml->lslot.p, g->d_start.p, g->d_count.p,      \
                                                      g->d_centre.p, g->px.p, g->py.p, g->pz.p, w->Qf.p,        \
                                                      pl->prox.p, w->W32.p, ml->far_off.p, OP, w->loc.p)
            if (far32 && o.store) SAL_FAR(5, RT, ml->far_op32.p);
            else if (far32) SAL_FAR(1, 16, nullptr);
            else if (far3 && o.store) SAL_FAR(8, RT, ml->far_op3.p);
            else if (far3) SAL_FAR(6, 16, nullptr);
            else if (o.store) SAL_FAR(2, RT, ml->far_op.p);
            else SAL_FAR(0, RT, nullptr);
#undef SAL_FAR
        }
        SAL_HIP(hipGetLastError());
    }
    {
        ProfScope ps("l2l_mix", s);
        for (size_t L = 0; L + 1 < ml->l2l_lo.size(); ++L) {
            const int a = ml->l2l_lo[L], b = ml->l2l_lo[L + 1];
            if (b > a)
                k_l2l<ND><<<b - a, std::min(256, (3 * R + 63) / 64 * 64), 0, s>>>(R, Rs, ml->l2l_panels.p + a, g->d_parent.p, ml->lslot.p, ml->own.p,
                                                g->d_bounds.p, w->loc.p);
        }
        SAL_HIP(hipGetLastError());
    }
    for (int r0 = 0; r0 < R; r0 += RT) {
        {
            ProfScope ps("near_mix", s);
            const int nf = ml->n_full, nh = ml->n_near - ml->n_full;
#define SAL_NEAR(MODE, NRT, N, OFF)                                                                             \
    if (N > 0)                                                                                                  \
        k_near<RT, MODE, NRT><<<xcd_grid(N), 64, 0, s>>>(                                                       \
            N, M, R, Rs, r0, ml->near_work.p + OFF, g->d_parent.p, pl->pt_off.p, pl->pt_ent.p,   \
            pl->slot.p, g->d_start.p, g->d_count.p, g->px.p, g->py.p, g->pz.p, w->Qf.p, pl->prox.p, w->W32.p,    \
            o.store ? ml->near_off.p + OFF : nullptr, ml->near_op.p, w->Ge.p, w->Gn.p, ND, ml->lslot.p,         \
            g->d_bounds.p, g->perm.p, g->xi.p, g->eta.p, ml->nloc > 0 ? w->loc.p : nullptr, ml->own.p, ml->cut2.p)
            if (o.store) {
                SAL_NEAR(2, 2, nf, 0);
                SAL_NEAR(2, 1, nh, nf);
            } else {
                SAL_NEAR(0, 2, nf, 0);
                SAL_NEAR(0, 1, nh, nf);
            }
#undef SAL_NEAR
        }
        if (ml->n_close > 0) {
            ProfScope ps("close_mix", s);
            k_close<RT><<<grid_for(g->n, 256 / (RT / 4)), 256, 0, s>>>(g->n, R, Rs, r0, ml->close_off.p, ml->close_src.p,
                                                                        ml->close_val.p, w->Qf.p, w->Ge.p, w->Gn.p);
        }
        SAL_HIP(hipGetLastError());
    }
}

template <int ND>
static void run_nd(sal_geom g, const Plan* pl, const MixLists* ml, sal_work w, const sal_eval_opts& o, int R, int rt,
                   hipStream_t s) {
    if (rt == 16) run_mixed<ND, 16>(g, pl, ml, w, o, R, s);
    else run_mixed<ND, 32>(g, pl, ml, w, o, R, s);
}

// Solid angle of the cube-face panel [xi0, xi1] x [eta0, eta1] (gnomonic angles): the face point
// (1, a, b) / |(1, a, b)| has area element da db / (1 + a^2 + b^2)^(3/2), whose antiderivative is
// F(a, b) = atan(a b / sqrt(1 + a^2 + b^2)), a = tan xi, b = tan eta (a whole face: 4 F(1, 1) = 2 pi / 3).
static double panel_solid_angle(double xi0, double xi1, double eta0, double eta1) {
    auto F = [](double xi, double eta) {
        const double a = tan(xi), b = tan(eta);
        return atan(a * b / sqrt(1.0 + a * a + b * b));
    };
    return F(xi1, eta1) - F(xi0, eta1) - F(xi1, eta0) + F(xi0, eta0);
}

// The close-pair list of a mixed plan (k_close_list): counts per target point, their prefix sum,
// then the sources.  Built once per plan on the handle's stream.
static void build_close(sal_geom g, const Plan* pl, MixLists* ml) {
    hipStream_t s = g->stream;
    const int64_t n = g->n;
    const unsigned blocks = (unsigned)((ml->n_near + 1) / 2);
    DBuf<int> cnt(n);
    SAL_HIP(hipMemsetAsync(cnt.p, 0, n * sizeof(int), s));
    k_close_list<false><<<blocks, 64, 0, s>>>(ml->n_near, ml->near_work.p, g->d_parent.p, pl->pt_off.p, pl->pt_ent.p,
                                              g->d_start.p, g->d_count.p, g->px.p, g->py.p, g->pz.p, ml->cut2.p,
                                              nullptr, cnt.p);
    SAL_HIP(hipGetLastError());
    ml->close_off.alloc(n + 1);
    const int64_t total = exclusive_scan_i32(cnt.p, ml->close_off.p, n, s);
    SAL_REQUIRE(total < (int64_t)INT32_MAX, SAL_ERR_LIMIT, "close-pair list longer than 2^31 - 1 entries: lower near64");
    const int t32 = (int)total;
    h2d(ml->close_off.p + n, &t32, 1, s);
    ml->close_src.alloc(std::max<int64_t>(total, 1));
    if (total > 0)
        k_close_list<true><<<blocks, 64, 0, s>>>(ml->n_near, ml->near_work.p, g->d_parent.p, pl->pt_off.p,
                                                 pl->pt_ent.p, g->d_start.p, g->d_count.p, g->px.p, g->py.p, g->pz.p,
                                                 ml->cut2.p, ml->close_off.p, ml->close_src.p);
    ml->close_val.alloc(std::max<int64_t>(total, 1));
    if (total > 0)
        k_close_vals<<<grid_for(n, 64), 64, 0, s>>>(n, ml->close_off.p, ml->close_src.p, g->px.p, g->py.p, g->pz.p,
                                                    ml->close_val.p);
    SAL_HIP(hipGetLastError());
    SAL_HIP(hipStreamSynchronize(s));
    ml->n_close = total;
}

// Work lists of the mixed plan: near work items, local-expansion panels (every panel with a cluster
// target on or above it), the top-down L2L order, the leaves that interpolate, the close pairs.
static MixLists* build_lists(sal_geom g, const Plan* pl) {
    auto ml = std::make_unique<MixLists>();
    const int P = g->P;
    std::vector<int> xt_off(P + 1), pt_off(P + 1), pt_ent(std::max<int64_t>(pl->pt_entries, 1)),
        xt_ent(std::max<int64_t>(pl->xt_entries, 1));
    d2h(xt_off.data(), pl->xt_off.p, P + 1, g->stream);
    d2h(pt_off.data(), pl->pt_off.p, P + 1, g->stream);
    d2h(pt_ent.data(), pl->pt_ent.p, pl->pt_entries, g->stream);
    d2h(xt_ent.data(), pl->xt_ent.p, pl->xt_entries, g->stream);
    SAL_HIP(hipStreamSynchronize(g->stream));
    // upward pass
    std::vector<char> is_src(P, 0);
    for (int64_t e = 0; e < pl->pt_entries; ++e)
        if ((pt_ent[e] & 3) == 1) is_src[pt_ent[e] >> 2] = 1;
    for (int64_t e = 0; e < pl->xt_entries; ++e)
        if ((xt_ent[e] & 3) == 3) is_src[xt_ent[e] >> 2] = 1;
    std::vector<int> uslot(P, -1), u_panels;
    for (int A = 0; A < P; ++A)
        if (g->count[A] > 0 && (g->count[A] > g->list_thresh || is_src[A])) {
            uslot[A] = (int)u_panels.size();
            u_panels.push_back(A);
        }
    std::vector<int> p2m_off(1, 0), p2m_src, m2m, m2m_lo(1, 0);
    for (int A : u_panels) {
        if (g->is_leaf[A]) p2m_src.push_back(A);
        else
            for (int k = 0; k < 4; ++k) {
                const int C = g->child1[A] + k;
                if (g->count[C] > 0 && uslot[C] < 0) p2m_src.push_back(C);
            }
        p2m_off.push_back((int)p2m_src.size());
    }
    for (int L = (int)g->level_lo.size() - 2; L >= 0; --L) {
        for (int A = g->level_lo[L]; A < g->level_lo[L + 1]; ++A) {
            if (uslot[A] < 0 || g->is_leaf[A]) continue;
            bool any = false;
            for (int k = 0; k < 4; ++k) any = any || uslot[g->child1[A] + k] >= 0;
            if (any) m2m.push_back(A);
        }
        m2m_lo.push_back((int)m2m.size());
    }
    // Local expansions: a panel "has" one when it or an ancestor is a cluster target.  Slots go to
    // the non-leaf panels that have one (the L2L push-down fills them) and to the leaves that are
    // cluster targets themselves (their own far-field values); a leaf evaluates its parent's
    // expansion on the parent's grid directly (exact for the degree-d polynomial), so leaves are not
    // pushed down to.
    std::vector<unsigned char> own(P, 0), has(P, 0);
    std::vector<int> lslot(P, -1);
    int nloc = 0;
    for (int A = 0; A < P; ++A) {  // ids are breadth first: parents before children
        own[A] = xt_off[A + 1] > xt_off[A];
        has[A] = own[A] || (g->parent[A] >= 0 && has[g->parent[A]]);
        if (has[A] && (!g->is_leaf[A] || own[A])) lslot[A] = nloc++;
    }
    std::vector<int> l2l, l2l_lo(1, 0), l2p;
    for (int L = 0; L + 1 < (int)g->level_lo.size(); ++L) {
        for (int A = g->level_lo[L]; A < g->level_lo[L + 1]; ++A)
            if (!g->is_leaf[A] && has[A] && g->parent[A] >= 0 && has[g->parent[A]]) l2l.push_back(A);
        l2l_lo.push_back((int)l2l.size());
    }
    std::vector<int2> work, half;  // items of 17..32 targets, items of at most 16
    for (int L = 0; L < P; ++L) {
        if (!g->is_leaf[L] || g->count[L] == 0) continue;
        for (int t = 0; t < g->count[L]; t += 32) (g->count[L] - t > 16 ? work : half).push_back(make_int2(L, t));
        if (has[L]) l2p.push_back(L);
    }
    // Longest chunks first: each launch's items, in chunks of XCD_CHUNK neighbouring leaves (the
    // unit the launch deals to the XCDs), ordered by the chunk's stored steps, largest first, so
    // the launch ends on short chunks instead of waiting for a long one.  Items keep their
    // neighbours, and no item's result depends on the order.
    auto longest_first = [&](std::vector<int2>& v) {
        const int nc = ((int)v.size() + XCD_CHUNK - 1) / XCD_CHUNK;
        std::vector<int64_t> steps(nc, 0);
        for (int i = 0; i < (int)v.size(); ++i)
            for (int A = v[i].x; A >= 0; A = g->parent[A])
                for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
                    const int ns = (pt_ent[e] & 3) == 0 ? g->count[pt_ent[e] >> 2] : pl->m;
                    steps[i / XCD_CHUNK] += (ns + 3) / 4;
                }
        std::vector<int> order(nc);
        for (int c = 0; c < nc; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return steps[a] > steps[b]; });
        std::vector<int2> out;
        out.reserve(v.size());
        for (int c : order)
            for (int i = c * XCD_CHUNK; i < std::min((int)v.size(), (c + 1) * XCD_CHUNK); ++i) out.push_back(v[i]);
        v.swap(out);
    };
    longest_first(work);
    longest_first(half);
    ml->n_full = (int)work.size();
    work.insert(work.end(), half.begin(), half.end());
    ml->nloc = nloc;
    ml->n_near = (int)work.size();
    ml->n_l2p = (int)l2p.size();
    ml->l2l_lo = l2l_lo;
    auto up = [&](auto& d, const auto& h) {
        d.alloc(std::max<size_t>(h.size(), 1));
        h2d(d.p, h.data(), h.size(), g->stream);
    };
    up(ml->near_work, work);
    up(ml->lslot, lslot);
    up(ml->l2l_panels, l2l);
    up(ml->own, own);
    up(ml->u_panels, u_panels);
    up(ml->uslot, uslot);
    up(ml->p2m_off, p2m_off);
    up(ml->p2m_src, p2m_src);
    up(ml->m2m_panels, m2m);
    ml->m2m_lo = m2m_lo;
    ml->nU = (int)u_panels.size();
    // close pairs: cut2 = (near64 h_L)^2 per leaf, h_L^2 = the leaf's solid angle over its points
    std::vector<float> cut2(P, -1.0f);
    if (pl->near64 > 0)
        for (int L = 0; L < P; ++L)
            if (g->is_leaf[L] && g->count[L] > 0) {
                const double* b = &g->bounds[6 * (size_t)L];
                const double k2 = (double)pl->near64 * pl->near64;
                cut2[L] = (float)(k2 * panel_solid_angle(b[0], b[2], b[3], b[5]) / g->count[L]);
            }
    up(ml->cut2, cut2);
    if (pl->near64 > 0 && ml->n_near > 0) build_close(g, pl, ml.get());
    SAL_HIP(hipStreamSynchronize(g->stream));
    return ml.release();
}

static MixLists* mix_lists(sal_geom g, Plan* pl) {
    std::lock_guard<std::mutex> lk(g->mu);
    if (!pl->mixed) {
        pl->mixed = build_lists(g, pl);
        pl->mixed_free = lists_free;
        pl->mixed_bytes = lists_bytes;
    }
    return (MixLists*)pl->mixed;
}

// Stored operator of a mixed plan: the evaluation kernels' own traversal in its storing variant
// (k_near MODE 1, k_far_mfma KIND 3) writes every kernel value it would use, so later evaluations
// read them (MODE 2 / KIND 2) with the same arithmetic result.  Sizes: 16 B per 4 sources per
// target pair row of a near work item (8 B per padded near pair), 4 B per far matrix entry of a
// proxy row (the padded rows of the last 16-row tile are not stored).
// Step offsets of the stored operator: noff per near work item (steps of 4 sources), foff per
// proxy-target panel (steps of 32 sources, every chunk of 96 rounded up to whole steps).
static void store_offsets(sal_geom g, const Plan* pl, const MixLists* ml, std::vector<int64_t>& noff,
                          std::vector<int64_t>& foff) {
    const int M = pl->m;
    hipStream_t s = g->stream;
    const int P = g->P;
    std::vector<int> pt_off(P + 1), xt_off(P + 1), pt_ent(std::max<int64_t>(pl->pt_entries, 1)),
        xt_ent(std::max<int64_t>(pl->xt_entries, 1)), xt_panels(std::max(pl->n_xt, 1));
    std::vector<int2> work(std::max(ml->n_near, 1));
    d2h(pt_off.data(), pl->pt_off.p, P + 1, s);
    d2h(xt_off.data(), pl->xt_off.p, P + 1, s);
    d2h(pt_ent.data(), pl->pt_ent.p, pl->pt_entries, s);
    d2h(xt_ent.data(), pl->xt_ent.p, pl->xt_entries, s);
    d2h(xt_panels.data(), pl->xt_panels.p, pl->n_xt, s);
    d2h(work.data(), ml->near_work.p, ml->n_near, s);
    SAL_HIP(hipStreamSynchronize(s));
    noff.assign(ml->n_near + 1, 0);
    foff.assign(pl->n_xt + 1, 0);
    for (int i = 0; i < ml->n_near; ++i) {  // floats: 64 lanes x 2 x (rows per lane) per step
        int64_t steps = 0;
        for (int A = work[i].x; A >= 0; A = g->parent[A])
            for (int e = pt_off[A]; e < pt_off[A + 1]; ++e) {
                const int ns = (pt_ent[e] & 3) == 0 ? g->count[pt_ent[e] >> 2] : M;
                steps += (ns + 3) / 4;
            }
        noff[i + 1] = noff[i] + steps * 64 * (i < ml->n_full ? 4 : 2);
    }
    for (int i = 0; i < pl->n_xt; ++i) {
        const int T = xt_panels[i];
        int64_t steps = 0;
        for (int e = xt_off[T]; e < xt_off[T + 1]; ++e) {
            const int ns = (xt_ent[e] & 3) == 2 ? g->count[xt_ent[e] >> 2] : M;
            for (int c0 = 0; c0 < ns; c0 += FAR_KC) steps += (std::min(FAR_KC, ns - c0) + 31) / 32;
        }
        foff[i + 1] = foff[i] + steps;
    }
}

// 16-byte units of a far step: full row tiles of 128, the last one 8 per proxy row it holds
static int64_t far_units(int m) {
    const int nt = (m + 15) / 16;
    return (int64_t)(nt - 1) * 128 + 8 * (m - 16 * (nt - 1));
}

// bytes of the stored operator: 1 KB per near step of a 32-target item (512 B for 16), 16 B per far
// unit per far step
static void store_sizes(const Plan* pl, const std::vector<int64_t>& noff, const std::vector<int64_t>& foff,
                        int64_t& near_b, int64_t& far_b) {
    near_b = noff.back() * (int64_t)sizeof(float);
    far_b = foff.back() * far_units(pl->m) * (int64_t)sizeof(bf8);
}

// The near part is built for every stored evaluation, the far part for the far-field kind the
// evaluation uses (split bf16, FP32, or three-part bf16); the far parts share the step offsets.
template <int ND>
static void build_store(sal_geom g, const Plan* pl, MixLists* ml, bool far, bool far32, bool far3) {
    constexpr int M = ND * ND, MP = ((M + 15) / 16) * 16, NRT = MP / 16;
    hipStream_t s = g->stream;
    std::vector<int64_t> noff, foff;
    store_offsets(g, pl, ml, noff, foff);
    const int nf = ml->n_full, nh = ml->n_near - ml->n_full;
    if (!ml->stored_near) {
    ml->near_off.alloc(noff.size());
    h2d(ml->near_off.p, noff.data(), noff.size(), s);
    ml->near_op.alloc((size_t)noff.back());
    if (nf > 0)
        k_near<16, 1, 2><<<nf, 64, 0, s>>>(0, M, 1, 16, 0, ml->near_work.p, g->d_parent.p, pl->pt_off.p, pl->pt_ent.p,
                                           pl->slot.p, g->d_start.p, g->d_count.p, g->px.p, g->py.p, g->pz.p, nullptr,
                                           pl->prox.p, nullptr, ml->near_off.p, ml->near_op.p, nullptr, nullptr, 0,
                                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ml->cut2.p);
    if (nh > 0)
        k_near<16, 1, 1><<<nh, 64, 0, s>>>(0, M, 1, 16, 0, ml->near_work.p + nf, g->d_parent.p, pl->pt_off.p,
                                           pl->pt_ent.p, pl->slot.p, g->d_start.p, g->d_count.p, g->px.p, g->py.p,
                                           g->pz.p, nullptr, pl->prox.p, nullptr, ml->near_off.p + nf, ml->near_op.p,
                                           nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                           ml->cut2.p);
    SAL_HIP(hipGetLastError());
    }
    if ((far && !ml->stored_far) || (far32 && !ml->stored_far32) || (far3 && !ml->stored_far3)) {
        if (!ml->far_off.p) {
            ml->far_off.alloc(foff.size());
            h2d(ml->far_off.p, foff.data(), foff.size(), s);
        }
    }
    if (far32 && !ml->stored_far32) {
        ml->far_op32.alloc((size_t)foff.back() * far_units(pl->m));
        if (pl->n_xt > 0)
            k_far_mfma<4, 16, MP><<<pl->n_xt, 256, 0, s>>>(M, 1, 16, 0, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p,
                                                           pl->slot.p, ml->lslot.p, g->d_start.p, g->d_count.p,
                                                           g->d_centre.p, g->px.p, g->py.p, g->pz.p, nullptr, pl->prox.p,
                                                           nullptr, ml->far_off.p, ml->far_op32.p, nullptr);
        SAL_HIP(hipGetLastError());
    }
    if (far3 && !ml->stored_far3) {
        ml->far_op3.alloc((size_t)foff.back() * far_units(pl->m) * 3 / 2);
        if (pl->n_xt > 0)
            k_far_mfma<7, 16, MP><<<pl->n_xt, 256, 0, s>>>(M, 1, 16, 0, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p,
                                                           pl->slot.p, ml->lslot.p, g->d_start.p, g->d_count.p,
                                                           g->d_centre.p, g->px.p, g->py.p, g->pz.p, nullptr, pl->prox.p,
                                                           nullptr, ml->far_off.p, ml->far_op3.p, nullptr);
        SAL_HIP(hipGetLastError());
    }
    if (far && !ml->stored_far) {
        ml->far_op.alloc((size_t)foff.back() * NRT * 64 * 2);
        if (pl->n_xt > 0)
            k_far_mfma<3, 16, MP><<<pl->n_xt, 256, 0, s>>>(M, 1, 16, 0, pl->xt_panels.p, pl->xt_off.p, pl->xt_ent.p,
                                                           pl->slot.p, ml->lslot.p, g->d_start.p, g->d_count.p,
                                                           g->d_centre.p, g->px.p, g->py.p, g->pz.p, nullptr, pl->prox.p,
                                                           nullptr, ml->far_off.p, ml->far_op.p, nullptr);
        SAL_HIP(hipGetLastError());
    }
    SAL_HIP(hipStreamSynchronize(s));
    ml->stored_near = true;
    if (far) ml->stored_far = true;
    if (far32) ml->stored_far32 = true;
    if (far3) ml->stored_far3 = true;
}

// the far operand of m2l: split bf16, FP32 or three-part bf16
static void ensure_store(sal_geom g, const Plan* pl, MixLists* ml, int m2l) {
    const bool far = (m2l == SAL_M2L_BF16X3), far32 = (m2l == SAL_M2L_FP32_MFMA), far3 = (m2l == SAL_M2L_BF16X6);
    std::lock_guard<std::mutex> lk(g->mu);
    if (ml->stored_near && (ml->stored_far || !far) && (ml->stored_far32 || !far32) && (ml->stored_far3 || !far3))
        return;
    switch (pl->degree) {
        case 4: build_store<5>(g, pl, ml, far, far32, far3); break;
        case 5: build_store<6>(g, pl, ml, far, far32, far3); break;
        case 6: build_store<7>(g, pl, ml, far, far32, far3); break;
        case 7: build_store<8>(g, pl, ml, far, far32, far3); break;
        case 8: build_store<9>(g, pl, ml, far, far32, far3); break;
        case 9: build_store<10>(g, pl, ml, far, far32, far3); break;
        default: build_store<11>(g, pl, ml, far, far32, far3); break;
    }
}

// RHS tile for R fields (16 or 32 fields per kernel pass)
static int check_opts(const sal_eval_opts& o, int R) {
    SAL_REQUIRE(o.mode == SAL_MODE_ACCURATE, SAL_ERR_ARG, "mixed precision runs in SAL_MODE_ACCURATE");
    SAL_REQUIRE(o.degree >= 4 && o.degree <= 10, SAL_ERR_ARG, "mixed precision supports degree 4..10");
    SAL_REQUIRE(o.m2l == SAL_M2L_BF16X3 || o.m2l == SAL_M2L_FP32_MFMA || o.m2l == SAL_M2L_BF16X6, SAL_ERR_ARG,
                "mixed precision: m2l must be SAL_M2L_BF16X3, SAL_M2L_FP32_MFMA or SAL_M2L_BF16X6");
    SAL_REQUIRE(R >= 1, SAL_ERR_ARG, "R must be >= 1");
    SAL_REQUIRE(o.near64 >= 0 && o.near64 <= 16, SAL_ERR_ARG, "near64 must be in [0, 16]");
    const int rt = o.rhs_tile > 0 ? o.rhs_tile : (R > 16 ? 32 : 16);
    SAL_REQUIRE(rt == 16 || rt == 32, SAL_ERR_ARG, "rhs_tile must be 16 or 32 in mixed precision");
    return rt;
}

// Element counts of the six workspace buffers for R fields.
static void work_sizes(sal_geom g, const Plan* pl, const MixLists* ml, int R, int rt, size_t c[6]) {
    const size_t n = (size_t)g->n, m = (size_t)pl->m, Rs = (size_t)row_stride(R, rt);
    c[0] = n * Rs;                                       // Qf
    c[1] = (size_t)std::max(pl->Q, 1) * m * Rs;          // W32
    c[2] = (size_t)std::max(ml->nloc, 1) * 3 * m * Rs;   // loc
    c[3] = (size_t)std::max(ml->nU, 1) * m * R;          // upward-pass expansions
    c[4] = c[5] = n * R;                                 // Ge, Gn
}

static int64_t work_need(sal_geom g, const Plan* pl, const MixLists* ml, int R, int rt) {
    size_t c[6];
    work_sizes(g, pl, ml, R, rt, c);
    return (int64_t)(4 * (c[0] + c[1] + c[2]) + 8 * (c[3] + c[4] + c[5]));
}

static void work_ensure(sal_work w, sal_geom g, const Plan* pl, const MixLists* ml, int R, int rt) {
    size_t c[6];
    work_sizes(g, pl, ml, R, rt, c);
    if (c[0] <= w->Qf.n && c[1] <= w->W32.n && c[2] <= w->loc.n && c[3] <= w->wpart.n && c[4] <= w->Ge.n &&
        c[5] <= w->Gn.n)
        return;
    w->drop_graphs();  // their kernel arguments name the buffers about to move
    w->Qf.ensure(c[0]);
    w->W32.ensure(c[1]);
    w->loc.ensure(c[2]);
    w->wpart.ensure(c[3]);
    w->Ge.ensure(c[4]);
    w->Gn.ensure(c[5]);
}

static void launch_all(sal_geom g, const Plan* pl, const MixLists* ml, sal_work w, const sal_eval_opts& o,
                       const double* area, const double* ssh, int R, int rt, double* gx, double* gy, int accumulate,
                       hipStream_t s) {
    const int64_t n = g->n;
    {
        ProfScope ps("qprep", s);
        k_qprep<<<grid_for(n, QB), QB, 0, s>>>(n, R, row_stride(R, rt), g->iperm.p, area, ssh, w->Qf.p);
    }
    switch (o.degree) {
        case 4: run_nd<5>(g, pl, ml, w, o, R, rt, s); break;
        case 5: run_nd<6>(g, pl, ml, w, o, R, rt, s); break;
        case 6: run_nd<7>(g, pl, ml, w, o, R, rt, s); break;
        case 7: run_nd<8>(g, pl, ml, w, o, R, rt, s); break;
        case 8: run_nd<9>(g, pl, ml, w, o, R, rt, s); break;
        case 9: run_nd<10>(g, pl, ml, w, o, R, rt, s); break;
        default: run_nd<11>(g, pl, ml, w, o, R, rt, s); break;
    }
    {
        ProfScope ps("out_mix", s);
        k_out_mix<<<grid_for(n, OB), 256, 0, s>>>(n, R, g->iperm.p, w->Ge.p, w->Gn.p, gx, gy, accumulate);
    }
    SAL_HIP(hipGetLastError());
}

static constexpr size_t MAX_GRAPHS = 8;  // captured sequences kept per workspace

void eval_mixed_work(sal_geom g, const sal_eval_opts& o, sal_work w, const double* area, const double* ssh, int R,
                     double* gx, double* gy, int accumulate, hipStream_t s) {
    const int rt = check_opts(o, R);
    Plan* pl = get_plan(g, o.degree, PLAN_MIXED, o.near64);
    MixLists* ml = mix_lists(g, pl);
    if (o.store) ensure_store(g, pl, ml, o.m2l);
    work_ensure(w, g, pl, ml, R, rt);
    if (!o.graph) {
        launch_all(g, pl, ml, w, o, area, ssh, R, rt, gx, gy, accumulate, s);
        return;
    }
    const void* key[4] = {area, ssh, gx, gy};
    const int acc = accumulate ? 1 : 0;
    for (auto& gr : w->graphs)
        if (gr.plan_serial == pl->serial && std::equal(key, key + 4, gr.ptr) && gr.R == R && gr.rt == rt &&
            gr.m2l == o.m2l && gr.acc == acc && gr.store == o.store) {
            SAL_HIP(hipGraphLaunch(gr.exec, s));
            return;
        }
    // capture on the workspace's own stream (the caller's may be the null stream), replay on the caller's
    if (!w->cap) SAL_HIP(hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    SAL_HIP(hipStreamBeginCapture(w->cap, hipStreamCaptureModeThreadLocal));
    set_capturing(true);
    try {
        launch_all(g, pl, ml, w, o, area, ssh, R, rt, gx, gy, accumulate, w->cap);
    } catch (...) {
        set_capturing(false);
        (void)hipStreamEndCapture(w->cap, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    set_capturing(false);
    SAL_HIP(hipStreamEndCapture(w->cap, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SAL_HIP(e);
    if (w->graphs.size() >= MAX_GRAPHS) {
        (void)hipGraphExecDestroy(w->graphs.front().exec);
        w->graphs.erase(w->graphs.begin());
    }
    sal_work_s::Graph gr{pl->serial, {key[0], key[1], key[2], key[3]}, R, rt, o.m2l, acc, o.store, exec};
    w->graphs.push_back(gr);
    SAL_HIP(hipGraphLaunch(exec, s));
}

static void work_free(void* p) { delete (sal_work)p; }

void eval_mixed(sal_geom g, const sal_eval_opts& o, const double* area, const double* ssh, int R, double* gx,
                double* gy, int accumulate, hipStream_t s) {
    check_opts(o, R);
    Plan* pl = get_plan(g, o.degree, PLAN_MIXED, o.near64);
    {
        std::lock_guard<std::mutex> lk(g->mu);
        if (!pl->work) {
            pl->work = new sal_work_s();
            pl->work_free = work_free;
        }
    }
    eval_mixed_work(g, o, (sal_work)pl->work, area, ssh, R, gx, gy, accumulate, s);
}

int64_t mixed_work_need(sal_geom g, const sal_eval_opts& o, int R) {
    const int rt = check_opts(o, R);
    Plan* pl = get_plan(g, o.degree, PLAN_MIXED, o.near64);
    MixLists* ml = mix_lists(g, pl);
    if (o.store) ensure_store(g, pl, ml, o.m2l);
    return work_need(g, pl, ml, R, rt);
}

int64_t work_bytes_after(sal_work w, sal_geom g, const sal_eval_opts& o, int R) {
    const int rt = check_opts(o, R);
    Plan* pl = get_plan(g, o.degree, PLAN_MIXED, o.near64);
    size_t c[6];
    work_sizes(g, pl, mix_lists(g, pl), R, rt, c);
    const size_t have[6] = {w->Qf.n, w->W32.n, w->loc.n, w->wpart.n, w->Ge.n, w->Gn.n};
    int64_t b = 0;
    for (int i = 0; i < 6; ++i) b += (int64_t)std::max(c[i], have[i]) * (i < 3 ? 4 : 8);
    return b;
}

sal_work work_new() { return new sal_work_s(); }
void work_delete(sal_work w) { delete w; }
int64_t work_held(sal_work w) { return w->bytes(); }

void work_release(sal_work w) {
    w->drop_graphs();
    w->Qf.release();
    w->W32.release();
    w->loc.release();
    w->wpart.release();
    w->Ge.release();
    w->Gn.release();
}

}  // namespace sal

// ------------------------------------------------------------------------------ C ABI

using namespace sal;

extern "C" {

int sal_work_create(sal_work* out) {
    return guard([&] {
        SAL_REQUIRE(out, SAL_ERR_ARG, "null pointer");
        *out = new sal_work_s();
    });
}

int sal_work_destroy(sal_work w) {
    return guard([&] { delete w; });
}

int sal_work_device_bytes(sal_work w, int64_t* bytes) {
    return guard([&] {
        SAL_REQUIRE(w && bytes, SAL_ERR_ARG, "null pointer");
        *bytes = w->bytes();
    });
}

int sal_plan_prepare(sal_geom g, const sal_eval_opts* o, int R, int64_t* work_bytes) {
    return guard([&] {
        SAL_REQUIRE(g && o, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->precision == SAL_PREC_MIXED, SAL_ERR_ARG, "sal_plan_prepare: mixed precision only");
        const int64_t need = mixed_work_need(g, *o, std::max(R, 1));
        if (work_bytes) *work_bytes = need;
    });
}

int sal_plan_stats(sal_geom g, const sal_eval_opts* o, int64_t st[13]) {
    return guard([&] {
        SAL_REQUIRE(g && o && st, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->precision == SAL_PREC_MIXED, SAL_ERR_ARG, "sal_plan_stats: mixed precision only");
        check_opts(*o, 1);
        Plan* pl = get_plan(g, o->degree, PLAN_MIXED, o->near64);
        MixLists* ml = mix_lists(g, pl);
        st[0] = pl->m;
        st[1] = pl->Q;
        st[2] = pl->n_chunks;
        st[3] = pl->n_xt;
        st[4] = pl->n_leaves;
        st[5] = pl->pt_entries;
        st[6] = pl->xt_entries;
        st[7] = ml->nloc;
        st[8] = ml->n_near;
        st[9] = ml->n_l2p;
        std::vector<int64_t> noff, foff;
        store_offsets(g, pl, ml, noff, foff);
        store_sizes(pl, noff, foff, st[10], st[11]);
        if (o->m2l == SAL_M2L_BF16X6) st[11] = st[11] * 3 / 2;  // three bf16 parts: 6 B per entry
        st[12] = ml->n_close;
    });
}

int sal_plan_close_export(sal_geom g, const sal_eval_opts* o, int* target, int* source, float* cut2) {
    return guard([&] {
        SAL_REQUIRE(g && o, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->precision == SAL_PREC_MIXED, SAL_ERR_ARG, "sal_plan_close_export: mixed precision only");
        check_opts(*o, 1);
        Plan* pl = get_plan(g, o->degree, PLAN_MIXED, o->near64);
        MixLists* ml = mix_lists(g, pl);
        hipStream_t s = g->stream;
        if (cut2) d2h(cut2, ml->cut2.p, g->P, s);
        if ((target || source) && ml->n_close > 0) {
            const int64_t n = g->n;
            std::vector<int> off(n + 1), src(ml->n_close), perm(n);
            d2h(off.data(), ml->close_off.p, n + 1, s);
            d2h(src.data(), ml->close_src.p, ml->n_close, s);
            d2h(perm.data(), g->perm.p, n, s);
            SAL_HIP(hipStreamSynchronize(s));
            for (int64_t k = 0; k < n; ++k)
                for (int e = off[k]; e < off[k + 1]; ++e) {
                    if (target) target[e] = perm[k];
                    if (source) source[e] = perm[src[e]];
                }
        }
        SAL_HIP(hipStreamSynchronize(s));
    });
}

int sal_plan_release(sal_geom g) {
    return guard([&] {
        SAL_REQUIRE(g, SAL_ERR_ARG, "null pointer");
        SAL_HIP(hipDeviceSynchronize());
        std::lock_guard<std::mutex> lk(g->mu);
        invalidate_plans(g);
    });
}

int sal_eval_work(sal_geom g, const sal_eval_opts* o, sal_work w, const double* d_area, const double* d_ssh, int R,
                  double* d_gx, double* d_gy, int accumulate, void* stream) {
    return guard([&] {
        SAL_REQUIRE(g && o && w && d_area && d_ssh && d_gx && d_gy, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->precision == SAL_PREC_MIXED, SAL_ERR_ARG, "sal_eval_work: mixed precision only");
        eval_mixed_work(g, *o, w, d_area, d_ssh, R, d_gx, d_gy, accumulate, (hipStream_t)stream);
    });
}

}  // extern "C"
