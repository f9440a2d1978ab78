// Evaluation plan: the interaction list regrouped for owner-computes evaluation, the proxy grids
// and the work buffers of one (degree, kind).
#pragma once

#include "common.h"

namespace sal {

constexpr int P2M_CHUNK = 4096;  // points per P2M block

enum PlanKind { PLAN_FP64_REF = 0, PLAN_FP64_ACC = 1, PLAN_MIXED = 2 };

struct Plan {
    uint64_t serial = 0;  // unique per plan built in this process
    int degree = 0, kind = 0, nd = 0, m = 0;
    int near64 = 0;      // PLAN_MIXED: close-pair cutoff factor of its work lists (sal_eval_opts.near64)
    int Q = 0;           // panels with proxies (more than cluster_thresh points)
    int n_chunks = 0;    // P2M chunks
    int n_xt = 0;        // proxy-target panels (targets of CP / CC entries)
    int n_leaves = 0;    // non-empty leaves
    int64_t pt_entries = 0, xt_entries = 0;
    DBuf<int> pt_off, pt_ent;   // PP / PC entries by target panel: (source << 2) | type
    DBuf<int> xt_off, xt_ent;   // CP / CC entries by target panel
    DBuf<int> slot, slot_panel; // panel -> proxy slot (-1), slot -> panel
    DBuf<int> xt_panels, leaves;
    DBuf<int> ck_slot, ck_first, ck_off;  // P2M chunks: panel, first point; chunk range per slot
    DBuf<double> prox;          // proxy points, [slot][k][3]
    DBuf<double2> uv;           // face angles (xi, eta) of every point, in tree order
    // FP64 reference plans: PP entries summed once per pair of mirrored entries (k_pair64).  pp_off
    // per pt_ent position: offset of the entry's partial sums (one per point of its target panel) in
    // pp_part, -1 where the leaf kernel sums the entry itself; pair_task {A, B, pos of A<-B, pos of
    // B<-A or -1}, A == B for a panel's entry with itself.
    int n_pair_tasks = 0;
    DBuf<int4> pair_task;
    DBuf<int64_t> pp_off;
    DBuf<double2> pp_part;
    // and the well-separated mirrors (k_farpair64): a CC entry with its CC mirror, a CP entry with its
    // PC mirror.  far_task {A | kind << 30, B, pos of A's entry, pos of B's entry}: kind 0 CC (both
    // xt_ent positions), kind 1 A's points <- B's proxies (pt_ent position) with B's proxies <- A's
    // points (xt_ent position).  PC partial sums go to pp_part (pp_off), proxy-target partial sums
    // (one double2 per proxy) to xp_part at xp_off per xt_ent position.
    int n_far_tasks = 0;
    DBuf<int4> far_task;
    DBuf<int64_t> xp_off;
    DBuf<double2> xp_part;
    // FP64 work buffers
    DBuf<double> q, W, wpart, phi;
    // mixed-precision work lists and the workspace of sal_eval_dev calls (evalmix.hip)
    void* mixed = nullptr;
    void (*mixed_free)(void*) = nullptr;
    void* work = nullptr;
    void (*work_free)(void*) = nullptr;
    int64_t (*mixed_bytes)(const void*) = nullptr;
    ~Plan() {
        if (work && work_free) work_free(work);
        if (mixed && mixed_free) mixed_free(mixed);
    }
};

Plan* get_plan(sal_geom g, int degree, int kind, int near64 = 0);
int64_t plan_device_bytes(const Plan* p);
Plan* build_plan(sal_geom g, int degree, int kind, const std::vector<int>& it, const std::vector<int>& is,
                 const std::vector<int>& ty);
void eval_fp64_plan(sal_geom g, Plan* pl, int mode, const double* area, const double* ssh, int R, double* gx,
                    double* gy, int accumulate, hipStream_t s);
void build_proxies(sal_geom g, Plan* pl, hipStream_t s);
void eval_fp64(sal_geom g, int degree, int mode, const double* area, const double* ssh, int R, double* gx, double* gy,
               int accumulate, hipStream_t s);
void eval_mixed(sal_geom g, const sal_eval_opts& o, const double* area, const double* ssh, int R, double* gx,
                double* gy, int accumulate, hipStream_t s);
void eval_mixed_work(sal_geom g, const sal_eval_opts& o, sal_work w, const double* area, const double* ssh, int R,
                     double* gx, double* gy, int accumulate, hipStream_t s);
int64_t mixed_work_need(sal_geom g, const sal_eval_opts& o, int R);
// bytes workspace w holds after growing for R fields of (g, o); release frees all its buffers
int64_t work_bytes_after(sal_work w, sal_geom g, const sal_eval_opts& o, int R);
void work_release(sal_work w);
sal_work work_new();
void work_delete(sal_work w);
int64_t work_held(sal_work w);

}  // namespace sal
