/* Test infrastructure: CPU restatement of the reference SAL-gradient treecode (cygnari/sal-fort).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
 * only as the checker.  The product (libsalfort) never links or calls it.
 *
 * Conventions: all ids are 0-based, -1 means "none" (the reference is 1-based with -1 for none).
 * Per-panel bounds are stored 6 per panel: min_xi, mid_xi, max_xi, min_eta, mid_eta, max_eta.
 * 2-D BLI arrays are column-major (xi index fastest) exactly like the reference's vals(i, j).
 *
 * Pinning (what each function is checked against, see tests/test_oracle.py):
 *   or_face/or_xieta/or_xyz       compiled reference (golden), round-trip and norm invariants
 *   or_bli_*                      closed forms (Chebyshev-2 nodes, weights), polynomial exactness,
 *                                 partition of unity, Kronecker property, compiled reference
 *   or_gfunc                      compiled reference bit-for-bit; closed form p2 identity
 *   or_tree_*, or_assign          compiled reference bit-for-bit; partition / BFS / leaf invariants
 *   or_dtt, or_stt                compiled reference bit-for-bit; coverage-exactly-once, MAC
 *   or_fast_sum, or_interact      compiled reference (<=1e-14); all-PP limit equals the direct sum
 *   or_proxy_source, or_mom6_sums compiled reference (transposed proxy layout included)
 *   or_direct_sum                 brute force on tiny inputs in numpy (tests), reference gfunc
 */
#ifndef SAL_ORACLE_H
#define SAL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cube_sphere_module.f90 */
int or_face(double x, double y, double z);
void or_xieta(double x, double y, double z, int face, double* xi, double* eta); /* face<0: from xyz; faces 1..6 */
void or_xyz(double xi, double eta, int face, double* x, double* y, double* z);

/* bli_module.f90 */
void or_bli_coeffs(int degree, double* c);
void or_bli_points(double a, double b, int degree, double* p);
void or_interp_vals(double xi, double eta, double a, double b, double c, double d, int degree, double* v);

/* fast_sum_module.f90:9 (identical copy mom6_fast_sum_module.f90:9) */
void or_gfunc(double tx, double ty, double tz, double sx, double sy, double sz, double* gx, double* gy);

/* context: one tree plus its lists, like the reference drivers' module-level arrays */
typedef struct or_ctx or_ctx;
or_ctx* or_ctx_new(void);
void or_ctx_free(or_ctx* c);

/* tree_traverse (tree_module.f90:59) */
int or_tree_build(or_ctx* c, int n, const double* x, const double* y, const double* z, int thresh);
int or_tree_n_panels(const or_ctx* c);
int64_t or_tree_total_inside(const or_ctx* c);
void or_tree_export(const or_ctx* c, int* level, int* is_leaf, int* parent, int* child, int* face,
                    double* bounds, double* radius, int* count, int64_t* offsets, int* points);

/* assign_points_to_panels (tree_module.f90:249): pp is nlev x n, level fastest; -1 pre-filled here */
void or_assign(const or_ctx* c, int n, const double* x, const double* y, const double* z, int nlev, int* pp);

/* DUAL_TREE_TRAVERSAL (tree_module.f90:277) */
int64_t or_dtt(or_ctx* c, double theta, int thresh);
void or_dtt_export(const or_ctx* c, int* it, int* is, int* ty);

/* fast_sum (fast_sum_module.f90:291): accumulates into gx, gy */
void or_fast_sum(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                 const double* area, const double* ssh, int degree, int rank, int nproc,
                 double* gx, double* gy);
/* one interaction as fast_sum dispatches it: kind 0 PP, 1 PC, 2 CP, 3 CC (0-based panel ids) */
void or_interact(const or_ctx* c, int kind, int it, int is, const double* x, const double* y,
                 const double* z, const double* area, const double* ssh, int degree, double* gx, double* gy);

/* SINGLE_TREE_TRAVERSAL (tree_module.f90:401) */
int or_stt(or_ctx* c, double theta, int thresh, int n, const double* x, const double* y, const double* z,
           int64_t* npc, int64_t* npp);
void or_stt_export(const or_ctx* c, int* pc_t, int* pc_s, int* pp_t, int* pp_s);

/* proxy_source_compute (mom6_fast_sum_module.f90:33): w has (d+1)^2 * n_panels entries */
void or_proxy_source(const or_ctx* c, int n, int nlev, const int* pp, const double* x, const double* y,
                     const double* z, const double* ssh, const double* area, int degree, double* w);
/* pc_compute then pp_compute (mom6_fast_sum_module.f90:153, :177) at one rank */
void or_mom6_sums(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                  const double* area, const double* ssh, int degree, const double* w, double* gx, double* gy);
void or_mom6_sums_sel(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                      const double* area, const double* ssh, int degree, const double* w,
                      const unsigned char* sel, double* gx, double* gy);

/* accuracy mode (not in the reference): stable kernel, no self pair, cluster targets interpolate the
 * Cartesian vector sum and project per target; accumulates into gx, gy */
void or_fast_sum_cart(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                      const double* area, const double* ssh, int degree, double* gx, double* gy);

/* FP64 direct sum at targets tgt[0..nt).  mode 0: reference expression, self pair included;
 * mode 1: reference expression, self pair skipped; mode 2: cancellation-free form of the same
 * kernel (chord from coordinate differences), self pair skipped.  Overwrites gx, gy (nt each). */
void or_direct_sum(int n, const double* x, const double* y, const double* z, const double* area,
                   const double* ssh, int nt, const int* tgt, int mode, double* gx, double* gy);

#ifdef __cplusplus
}
#endif
#endif
