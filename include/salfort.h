/* salfort: MI355X-native SAL-gradient treecode, C ABI.
 *
 * This is the boundary the Fortran drop-in modules (tree_module, fast_sum_module,
 * mom6_fast_sum_module in sal-fort_amd/fortran/) bind through ISO_C_BINDING, and the one the
 * Python host side and the tests load with ctypes.  Plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns an int status: SAL_OK (0) or a SAL_ERR_* code; sal_last_error()
 *     gives the message of the last failure on the calling thread.  No exception crosses.
 *   - Panel and point ids are 0-based, -1 means none (the Fortran shim adds 1).
 *   - "host" pointers are ordinary host memory; "_dev" entry points take device pointers and a
 *     hipStream_t passed as void* (NULL = the null stream) and are asynchronous unless noted.
 *   - Per-panel bounds: 6 doubles per panel, min_xi, mid_xi, max_xi, min_eta, mid_eta, max_eta.
 *   - Outputs of the sums are accumulated into (+=) when `accumulate` is non-zero, as the
 *     reference's evaluators do (fast_sum_module.f90:69-70), else overwritten.
 *   - A geometry handle is not re-entrant: one thread and one stream at a time per handle.
 */
#ifndef SALFORT_H
#define SALFORT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SAL_OK = 0,
    SAL_ERR_ARG = 1,    /* bad argument (null pointer, size, degree, mode) */
    SAL_ERR_HIP = 2,    /* a HIP runtime call failed */
    SAL_ERR_NOMEM = 3,  /* device or host allocation failed, or the memory budget would be exceeded */
    SAL_ERR_STATE = 4,  /* call order: e.g. evaluation before an interaction list exists */
    SAL_ERR_LIMIT = 5,  /* an internal limit was hit (e.g. tree depth) */
    SAL_ERR_IO = 6      /* a file could not be read or written, or is not a valid geometry cache */
};

/* evaluation modes */
enum {
    SAL_MODE_REFERENCE = 0, /* the reference's semantics: kernel expression as written (no FMA
                               contraction), self pair included, cluster targets interpolate the
                               projected east/north components (fast_sum_module.f90:133-288) */
    SAL_MODE_ACCURATE = 1   /* cancellation-free kernel form, self pair skipped, cluster targets
                               interpolate the tangential Cartesian vector and project per target */
};

/* arithmetic */
enum {
    SAL_PREC_FP64 = 0,
    SAL_PREC_MIXED = 1      /* SAL_MODE_ACCURATE only: FP64 coordinate differences, FP32 kernel values,
                               FP32 MFMA near-field products summed per source panel in FP32 and across
                               panels in FP64, close near pairs in FP64 (near64 below), far-field
                               products on MFMA (below), FP64 outputs */
};

/* far-field (proxy-target) product variants in SAL_PREC_MIXED */
enum {
    SAL_M2L_BF16X3 = 1,     /* split-bf16 hi/lo, three MFMA products, FP32 accumulation */
    SAL_M2L_FP32_MFMA = 2,  /* FP32 MFMA, FP32 per-chunk sums added in FP64 */
    SAL_M2L_BF16X6 = 3      /* split bf16 in three parts (24 bits, as FP32), six MFMA products, FP32
                               per-chunk sums added in FP64: the FP32 variant's precision on bf16 MFMA */
};

typedef struct sal_geom_s* sal_geom;

const char* sal_last_error(void);
int sal_version(int* major, int* minor);

/* -------------------------------------------------------------------------------- tree
 * Replaces tree_traverse (tree_module.f90:59-246).  Builds the cubed-sphere quadtree on the GPU:
 * panel ids, bounds, point membership and the BFS numbering are bit-identical to the reference.
 * Panel centres and radii use the host's libm (the reference's own functions), and point-to-child
 * decisions that fall within a few ulp of a panel midline are re-decided with the host's libm. */
int sal_geom_create(const double* x, const double* y, const double* z, int64_t n, int cluster_thresh,
                    sal_geom* out);
int sal_geom_create_dev(const double* d_x, const double* d_y, const double* d_z, int64_t n,
                        int cluster_thresh, void* stream, sal_geom* out); /* synchronous */
int sal_geom_destroy(sal_geom g);
int sal_geom_counts(sal_geom g, int* n_panels, int* max_level, int64_t* n_inside_total);
/* any output pointer may be NULL; points_inside is each panel's points ascending, panels in id order */
int sal_geom_export(sal_geom g, int* level, int* is_leaf, int* parent, int* child4, int* face,
                    double* bounds6, double* radius, int* count, int64_t* offsets, int* points_inside);
/* the handle's point count and the cluster_thresh its tree was built with */
int sal_geom_size(sal_geom g, int64_t* n_points, int* cluster_thresh);
/* how many point decisions and MAC decisions the host re-decided (ties) */
int sal_geom_ties(sal_geom g, int64_t* point_ties, int64_t* pair_ties);
/* test hook: widen the tie window to `ulps` units in the last place (default 16) */
int sal_set_tie_ulps(double ulps);
/* device bytes held by the handle */
int sal_geom_device_bytes(sal_geom g, int64_t* bytes);

/* On-disk geometry cache (no counterpart in the reference, whose drivers rebuild the tree and the
 * lists on every run, code.f90:72-82, mom6_code.f90:177-220).  sal_geom_save writes the handle's
 * points, tree (panels, face angles, per-level panel of every point, DFS layout), DTT list and STT
 * lists to `path` (written to path.tmp, then renamed).  sal_geom_load creates a handle from such a
 * file: its exports, lists and sums are identical to the saved handle's.  Evaluation plans and
 * stored operators are rebuilt on first use.  A damaged, truncated or foreign file fails with
 * SAL_ERR_IO and creates no handle. */
int sal_geom_save(sal_geom g, const char* path);
int sal_geom_load(const char* path, sal_geom* out);

/* Replaces assign_points_to_panels (tree_module.f90:249-273): points_panels is n_own x nlev with the
 * level index fastest (the Fortran points_panels(nlev, n_own)), -1 past the leaf. */
int sal_assign_points(sal_geom g, const double* x, const double* y, const double* z, int64_t n_own, int nlev,
                      int* points_panels);

/* -------------------------------------------------------------------------------- lists
 * Replaces DUAL_TREE_TRAVERSAL (tree_module.f90:277-398): the interaction list, in the reference's
 * order, built on the GPU generation by generation.  type: 0 PP, 1 PC, 2 CP, 3 CC.
 * cluster_thresh is DUAL_TREE_TRAVERSAL's param_cluster_thresh (<= 0: the tree's). */
int sal_dtt(sal_geom g, double theta, int cluster_thresh, int64_t* n_interactions);
int sal_dtt_counts(sal_geom g, int64_t counts4[4]);
int sal_dtt_export(sal_geom g, int* target, int* source, int* type);
/* a list built elsewhere (e.g. handed over by a Fortran caller) replaces the handle's list */
int sal_dtt_import(sal_geom g, const int* target, const int* source, const int* type, int64_t n);

/* Replaces SINGLE_TREE_TRAVERSAL (tree_module.f90:401-487): per-target lists for the targets
 * x,y,z (host), pc then pp, each in target-major order exactly as the reference splits them. */
int sal_stt(sal_geom g, double theta, const double* x, const double* y, const double* z, int64_t n_own,
            int64_t* n_pc, int64_t* n_pp);
int sal_stt_export(sal_geom g, int* pc_target, int* pc_source, int* pp_target, int* pp_source);
/* sizes of the handle's STT lists: targets, pc entries, pp entries (0 before sal_stt) */
int sal_stt_counts(sal_geom g, int64_t* n_targets, int64_t* n_pc, int64_t* n_pp);

/* -------------------------------------------------------------------------------- sums
 * Replaces fast_sum (fast_sum_module.f90:291-332) for the handle's DTT list: host arrays of the
 * grid's n points in the caller's order.  FP64, SAL_MODE_REFERENCE, one field. */
int sal_fast_sum(sal_geom g, const double* area, const double* ssh, int degree, double* gx, double* gy,
                 int accumulate);

/* One interaction as fast_sum dispatches it (pp_interact / pc_interact / cp_interact / cc_interact,
 * fast_sum_module.f90:33-288): kind 0..3 between panels target and source of the handle's tree, FP64
 * in the reference's semantics.  area, ssh, gx, gy are the grid's n points in the caller's order;
 * only the source panel's area and ssh are read and only the target panel's gx, gy written (added to,
 * or, with accumulate = 0, written into zeroed arrays).  The cost is the entry's own (its points and
 * proxies), as in the reference; device scratch is kept with the handle. */
int sal_interact(sal_geom g, int kind, int target, int source, int degree, const double* area, const double* ssh,
                 double* gx, double* gy, int accumulate);

/* In SAL_PREC_MIXED every coordinate difference (near pairs and kernel-matrix entries alike) is
 * formed in FP64 before it is rounded to FP32, so the chord keeps its relative accuracy at any
 * distance.  The FP32 kernel value still carries a relative error of about 1e-7, and the closest
 * pairs of a target dominate the sum of squares of its near-field terms, so their rounding is
 * most of the near field's error (the cancelling sum of a smooth field's gradient amplifies it).
 * Close pairs are therefore evaluated in FP64: a point-to-point (PP) pair whose FP32 chord squared
 * is below near64^2 * h_L^2, with h_L^2 = (solid angle of the target's leaf) / (points in the
 * leaf), the leaf's mean point spacing squared, gets an FP64 kernel value, FP64 products and FP64
 * sums; every other pair stays in the FP32 stream.  The chordal cutoff is near64 * h_L (default
 * near64 = 2: about 12 pairs per target on a near-uniform grid). */
typedef struct {
    int degree;     /* BLI degree d (proxies per panel (d+1)^2): 0..16 FP64, 4..10 mixed */
    int mode;       /* SAL_MODE_* */
    int precision;  /* SAL_PREC_* */
    int m2l;        /* SAL_M2L_* (SAL_PREC_MIXED only) */
    int rhs_tile;   /* fields per kernel pass in SAL_PREC_MIXED: 16 or 32 (0 = automatic) */
    int graph;      /* SAL_PREC_MIXED: capture the launch sequence as a HIP graph on the first call
                       with these buffers and replay it on later calls (1), or launch directly (0) */
    int store;      /* SAL_PREC_MIXED: keep the evaluated kernel values in device memory (the near-field
                       values of every work item, the far-field matrices of every proxy-target panel in
                       the form m2l uses: split bf16 or FP32), built once per geometry, degree and far
                       form, and stream them on every later call (1); or evaluate them on every call (0).
                       Same results either way; the stored form costs about 8 B per near pair and 4 B per
                       far matrix entry (sal_plan_stats) */
    int near64;     /* SAL_PREC_MIXED: the close-pair cutoff factor above, 0..16 (0: no FP64 pairs;
                       default 2).  Plans, close-pair lists and stored operators are kept per value */
} sal_eval_opts;

int sal_eval_opts_default(sal_eval_opts* o);

/* R fields on the device: area[n] and ssh[R*n] (field-major), outputs gx, gy [R*n], in the caller's
 * point order.  Enqueued on `stream`.  Uses the handle's own work buffers: calls on one handle must
 * not overlap (use sal_eval_work with one workspace per concurrent caller). */
int sal_eval_dev(sal_geom g, const sal_eval_opts* o, const double* d_area, const double* d_ssh, int R,
                 double* d_gx, double* d_gy, int accumulate, void* stream);

/* -------------------------------------------------------------------------------- workspaces
 * A workspace holds the device buffers of one mixed-precision evaluation (packed fields, proxy
 * weights, local expansions, FP64 output sums) and its captured HIP graphs.  Several workspaces
 * evaluate one geometry concurrently on different streams; one workspace runs one call at a time.
 * No counterpart in the reference (its sums allocate per interaction). */
typedef struct sal_work_s* sal_work;
int sal_work_create(sal_work* out);
int sal_work_destroy(sal_work w);
int sal_work_device_bytes(sal_work w, int64_t* bytes);
/* Builds the mixed plan and work lists of (g, o) if absent (not concurrently with evaluations of g)
 * and returns the device bytes a workspace needs for R fields. */
int sal_plan_prepare(sal_geom g, const sal_eval_opts* o, int R, int64_t* work_bytes);
/* Sizes of the mixed plan of (g, o): st = { proxies per panel m, cluster panels Q, P2M chunks,
 * proxy-target panels, non-empty leaves, PP+PC entries, CP+CC entries, local-expansion panels,
 * near work items (32 targets each), interpolating leaves, stored near-operator bytes, stored
 * far-operator bytes, close (FP64) pairs } (the sizes the stored operator has, or would have:
 * nothing is built here) */
int sal_plan_stats(sal_geom g, const sal_eval_opts* o, int64_t st[13]);
/* The close-pair list of the mixed plan of (g, o) (the pairs summed in FP64, near64 above), in the
 * caller's point order: pair e is (target[e], source[e]), targets in tree order, each target's sources
 * in the near kernel's walk order; st[12] of sal_plan_stats pairs.  cut2 (may be NULL) receives each
 * panel's FP32 chord-squared cutoff ((near64 h_L)^2 for non-empty leaves, negative otherwise). */
int sal_plan_close_export(sal_geom g, const sal_eval_opts* o, int* target, int* source, float* cut2);
/* Frees every cached evaluation plan of g (work lists, stored operators, the handle's own work
 * buffers); they are rebuilt on the next evaluation.  Not concurrently with evaluations of g. */
int sal_plan_release(sal_geom g);
int sal_eval_work(sal_geom g, const sal_eval_opts* o, sal_work w, const double* d_area, const double* d_ssh, int R,
                  double* d_gx, double* d_gy, int accumulate, void* stream);

/* -------------------------------------------------------------------------------- serving
 * A request loop over HBM-resident geometries (no counterpart in the reference, whose drivers
 * evaluate one field per run).  Requests carry R fields on one geometry as host arrays in the
 * caller's point order (ssh [R*n] in, gx, gy [R*n] out; pinned memory copies fastest) and stay
 * owned by the caller until sal_server_wait reports them done.  `slots` evaluations run at once,
 * each on its own stream and workspace; an idle slot takes the oldest request and merges the queued
 * requests of the same geometry behind it, up to max_fields fields, if the device memory it needs
 * fits: the held bytes of every slot must stay within the budget (hbm_budget, or the free memory at
 * creation minus hbm_reserve when 0) and a slot's growth within the free memory minus hbm_reserve.
 * A request that can never fit fails with SAL_ERR_LIMIT. */
typedef struct sal_server_s* sal_server;
typedef struct {
    int slots;            /* concurrent evaluations, 1..8 */
    int max_fields;       /* fields per request and per merged evaluation */
    int batching;         /* merge queued requests of one geometry (1) or run them one by one (0) */
    int graph;            /* evaluate through captured HIP graphs */
    int device;           /* HIP device */
    int64_t hbm_budget;   /* device bytes for request buffers (0: free memory at creation - reserve) */
    int64_t hbm_reserve;  /* device bytes left free */
} sal_server_opts;

int sal_server_opts_default(sal_server_opts* o);
int sal_server_create(const sal_server_opts* o, sal_server* out);
/* Registers geometry g (with its DTT list) and its cell areas (host, n): the areas go to HBM and the
 * mixed plan of *o is built now.  g must outlive the server. */
int sal_server_add_geometry(sal_server s, sal_geom g, const double* area, const sal_eval_opts* o, int* gid);
int sal_server_submit(sal_server s, int gid, int R, const double* ssh, double* gx, double* gy, int64_t* req);
/* Waits up to timeout_s seconds (< 0: no limit); *done = 1 once finished.  A failed request returns its
 * error status. */
int sal_server_wait(sal_server s, int64_t req, double timeout_s, int* done);
/* Seconds since server creation at submit, admission and completion; the evaluation batch that ran
 * the request and that batch's field count. */
int sal_server_request_info(sal_server s, int64_t req, double t[3], int64_t* batch, int* batch_fields);
/* st = { submitted, completed, rejected, batches, requests that ran merged, peak busy slots,
 *        peak held device bytes, budget } */
int sal_server_stats(sal_server s, int64_t st[8]);
/* Finishes queued work, then stops the workers and frees the server's buffers. */
int sal_server_destroy(sal_server s);

/* Replaces proxy_source_compute (mom6_fast_sum_module.f90:33-74): weights of every panel,
 * (d+1)^2 per panel in the reference's xi-major layout.  Host arrays. */
int sal_mom6_proxy_source(sal_geom g, const double* area, const double* ssh, int degree, double* w);
/* Replaces pc_compute (mom6_fast_sum_module.f90:153-174) for the handle's STT lists: w is the caller's
 * proxy weights ((d+1)^2 per panel, xi-major); transposed != 0 reproduces the reference's eta-major
 * read of them (:96-104).  gx, gy have n_own entries (the STT targets).  Host arrays. */
int sal_mom6_pc_compute(sal_geom g, const double* w, int degree, int transposed, double* gx, double* gy,
                        int accumulate);
/* Replaces pp_compute (mom6_fast_sum_module.f90:177-192): direct sums over the PP entries. */
int sal_mom6_pp_compute(sal_geom g, const double* area, const double* ssh, double* gx, double* gy, int accumulate);
/* One target point (x, y, z) against one source panel, as pc_interact (kind 1: the panel's proxies
 * with weights w[(d+1)^2], read eta-major when transposed != 0; mom6_fast_sum_module.f90:76-107)
 * and pp_interact (kind 0: the panel's points with ssh * area of the grid's arrays in the caller's
 * order, only the panel's own entries read; :109-150) do it.  Adds to *gx, *gy. */
int sal_mom6_interact(sal_geom g, int kind, double x, double y, double z, int source, int degree, int transposed,
                      const double* w, const double* area, const double* ssh, double* gx, double* gy);
/* Replaces pp_interact (mom6_fast_sum_module.f90:109-150) with the sources given as points, the way
 * a caller at several ranks holds them (its own points and halo copies, relabeled_points_inside): one
 * target (x, y, z) against ns source points (sx, sy, sz) with their ssh and area (host arrays),
 * summed as the reference sums them (ssh * area * kernel, in the given order).  Adds to *gx, *gy. */
int sal_mom6_pp_points(sal_geom g, double x, double y, double z, int ns, const double* sx, const double* sy,
                       const double* sz, const double* ssh, const double* area, double* gx, double* gy);
/* proxy_source_compute + pc_compute + pp_compute in one call. */
int sal_mom6_sums(sal_geom g, const double* area, const double* ssh, int degree, int transposed,
                  double* gx, double* gy, int accumulate);

/* -------------------------------------------------------------------------------- oracle sums
 * FP64 O(N^2) direct sum at nt target indices (device arrays of the grid in the caller's order).
 * mode 0: reference expression, self pair included; 1: reference expression, self skipped;
 * 2: cancellation-free form, self skipped.  R fields (ssh field-major); outputs [R*nt]. */
int sal_direct_sum_dev(const double* d_x, const double* d_y, const double* d_z, const double* d_area,
                       const double* d_ssh, int64_t n, int R, const int* d_targets, int64_t nt, int mode,
                       double* d_gx, double* d_gy, void* stream);

/* sal_direct_sum_dev without any allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  Same arguments and results as sal_direct_sum_dev (the same kernels, bit for bit), with its
 * partial sums in a static 64 MiB device buffer that each GPU the library runs on holds until the
 * process ends: 2 doubles per target per 16384 points, at most 8388608 doubles (for example 20,000
 * targets of 2.3M points), else SAL_ERR_LIMIT before any launch.  Asynchronous on `stream`; targets
 * must be point indices in [0, n).  Every call on a GPU shares that GPU's buffer: calls that overlap
 * on one GPU (different streams or threads) race.  The first call on each device looks up the
 * buffer's address there (not inside a graph capture). */
int sal_direct_sum_dev_noalloc(const double* d_x, const double* d_y, const double* d_z, const double* d_area,
                               const double* d_ssh, int64_t n, int R, const int* d_targets, int64_t nt, int mode,
                               double* d_gx, double* d_gy, void* stream);

/* The gradient as a tangential Cartesian 3-vector per point: T = east e_lon + north e_lat, with the
 * unit vectors the reference's kernel projects on (fast_sum_module.f90:25-27).  The reference returns
 * the two components (sal_longrad, sal_latgrad); this gives the 3-vector form of the same result.
 * Device arrays: points x, y, z [n] and gx, gy [R*n] (field-major) in the same point order; g3 is
 * [R][n][3].  Zero at the poles, where both components are zero (fast_sum_module.f90:19). */
int sal_gradient_xyz_dev(const double* d_x, const double* d_y, const double* d_z, int64_t n, int R,
                         const double* d_gx, const double* d_gy, double* d_g3, void* stream);

/* -------------------------------------------------------------------------------- measurement
 * Kernel timing: while enabled, every evaluation kernel launch is bracketed by HIP events on the
 * stream it runs on; durations are summed per kernel name. */
int sal_prof_enable(int on);
int sal_prof_reset(void);
int sal_prof_count(int* n);
int sal_prof_get(int idx, char* name, int len, double* total_ms, int64_t* launches);
/* Pair evaluations of the handle's list at degree d (m = (d+1)^2): K4 = { sum_PP c_t c_s,
 * sum_PC c_t m, sum_CP m c_s, sum_CC m^2 } */
int sal_work_counts(sal_geom g, int degree, int64_t K4[4]);

/* -------------------------------------------------------------------------------- inputs
 * Seeded synthetic inputs (host, deterministic, OpenMP).  Unit vectors on the sphere. */
int sal_gen_fibonacci(int64_t n, double* x, double* y, double* z);
/* rejection sampling, density 1 + 15 exp(-d^2 / (2 (15 deg)^2)), d the angle to (40N, 60W);
 * areas proportional to 1 / density, normalised to total_area */
int sal_gen_mpas(int64_t n, uint64_t seed, double total_area, double* x, double* y, double* z, double* area);
/* 0.25-degree lat-lon ocean cell centres (nlon x nlat, lat lat0..lat1) with a seeded smooth land
 * mask at land_frac; returns the ocean count in *n_out (call with x == NULL to size) */
int sal_gen_mom6(int nlon, int nlat, double lat0, double lat1, uint64_t seed, double land_frac, double radius,
                 int64_t* n_out, double* x, double* y, double* z, double* area);
/* real spherical-harmonic field, degree <= lmax, coefficients ~ N(0, (amp / (l + 1))^2), plus white
 * noise N(0, noise^2); R fields with seeds seed0 .. seed0 + R - 1, field-major output [R*n] */
int sal_gen_ssh(int64_t n, const double* x, const double* y, const double* z, int lmax, double amp, double noise,
                uint64_t seed0, int R, double* ssh);

#ifdef __cplusplus
}
#endif
#endif
