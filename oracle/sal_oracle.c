/* Test infrastructure: CPU restatement of the reference SAL-gradient treecode.  See sal_oracle.h
 * for the loading rule and for what pins each function.
 *
 * Floating point follows the reference's Fortran evaluation order exactly (operators of equal
 * precedence left to right, no contraction: built with -ffp-contract=off on the default x86-64
 * target, like the reference built with amdflang -O2) and calls the same libm entry points
 * (atan, tan, acos, cos, sqrt), so the tree, the lists and every kernel value are bit-identical to
 * the compiled reference on the same machine.
 */
#include "sal_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_REF (4.0 * atan(1.0)) /* pi = 4.D0*DATAN(1.D0), tree_module.f90:70, bli_module.f90:35 */

static void* xmalloc(size_t n) {
    void* p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "sal_oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static void* xrealloc(void* p, size_t n) {
    void* q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "sal_oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return q;
}

/* ---------------------------------------------------------------- cube_sphere_module.f90 */

/* face_from_xyz, cube_sphere_module.f90:7-35 */
int or_face(double x, double y, double z) {
    double ax = fabs(x), ay = fabs(y), az = fabs(z);
    if (ax >= ay && ax >= az) return x >= 0 ? 1 : 3;
    if (ay >= ax && ay >= az) return y >= 0 ? 2 : 4;
    return z >= 0 ? 5 : 6;
}

/* xieta_from_xyz, cube_sphere_module.f90:38-61 */
void or_xieta(double x, double y, double z, int face, double* xi, double* eta) {
    if (face < 1) face = or_face(x, y, z);
    switch (face) {
        case 1: *xi = atan(y / x); *eta = atan(z / x); break;
        case 2: *xi = atan(-x / y); *eta = atan(z / y); break;
        case 3: *xi = atan(y / x); *eta = atan(-z / x); break;
        case 4: *xi = atan(-x / y); *eta = atan(-z / y); break;
        case 5: *xi = atan(y / z); *eta = atan(-x / z); break;
        case 6: *xi = atan(-y / z); *eta = atan(-x / z); break;
        default: break; /* reference leaves xi, eta undefined */
    }
}

/* xyz_from_xieta, cube_sphere_module.f90:64-95 */
void or_xyz(double xi, double eta, int face, double* x, double* y, double* z) {
    double ax = tan(xi), ay = tan(eta);
    double r = sqrt(1.0 + ax * ax + ay * ay);
    switch (face) {
        case 1: *x = 1.0 / r; *y = ax * *x; *z = ay * *x; break;
        case 2: *y = 1.0 / r; *x = -ax * *y; *z = ay * *y; break;
        case 3: *x = -1.0 / r; *y = ax * *x; *z = -ay * *x; break;
        case 4: *y = -1.0 / r; *x = -ax * *y; *z = -ay * *y; break;
        case 5: *z = 1.0 / r; *x = -ay * *z; *y = ax * *z; break;
        default: *z = -1.0 / r; *x = -ay * *z; *y = -ax * *z; break;
    }
}

/* ---------------------------------------------------------------- bli_module.f90 */

/* bli_coeffs, bli_module.f90:6-26 */
void or_bli_coeffs(int degree, double* c) {
    if (degree == 0) {
        c[0] = 1.0;
        return;
    }
    for (int i = 1; i <= degree + 1; ++i) {
        double sgn = ((i + 1) % 2 == 0) ? 1.0 : -1.0; /* (-1)**(i+1) */
        if (i == 1) c[i - 1] = 0.5;
        else if (i == degree + 1) c[i - 1] = 0.5 * sgn;
        else c[i - 1] = sgn;
    }
}

/* bli_interp_points_shift, bli_module.f90:28-48 (nodes descend: i=1 sits at max_x) */
void or_bli_points(double a, double b, int degree, double* p) {
    double pi = PI_REF;
    double range = 0.5 * (b - a), shift = 0.5 * (b + a);
    if (degree == 0) {
        p[0] = shift;
        return;
    }
    for (int i = 1; i <= degree + 1; ++i) p[i - 1] = cos(pi / (double)degree * (double)(i - 1)) * range + shift;
}

/* 1-D factor of interp_vals_bli, bli_module.f90:71-105 */
static void bli_1d(double x, const double* nodes, const double* coef, int degree, double* f) {
    int n = degree + 1, found = 0;
    for (int i = 0; i < n; ++i) f[i] = 0.0;
    for (int i = 0; i < n; ++i)
        if (fabs(x - nodes[i]) < 1.0e-16) {
            found = 1;
            f[i] = 1.0;
        }
    if (!found) {
        double den = 0.0;
        for (int i = 0; i < n; ++i) {
            double v = coef[i] / (x - nodes[i]);
            f[i] = v;
            den = den + v;
        }
        for (int i = 0; i < n; ++i) f[i] = f[i] / den;
    }
}

/* interp_vals_bli, bli_module.f90:50-113: v[i + j*(d+1)] = l_xi(i) * l_eta(j) */
void or_interp_vals(double xi, double eta, double a, double b, double c, double d, int degree, double* v) {
    int n = degree + 1;
    double nx[64], ny[64], cf[64], fx[64], fy[64];
    or_bli_points(a, b, degree, nx);
    or_bli_points(c, d, degree, ny);
    or_bli_coeffs(degree, cf);
    /* the reference scans both dimensions in one loop (:71-80); the flags are independent */
    bli_1d(xi, nx, cf, degree, fx);
    bli_1d(eta, ny, cf, degree, fy);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) v[i + j * n] = fx[i] * fy[j];
}

/* ---------------------------------------------------------------- kernel */

/* sal_grad_gfunc, fast_sum_module.f90:9-31 */
void or_gfunc(double tx, double ty, double tz, double sx, double sy, double sz, double* gx, double* gy) {
    const double cons = -2.343256857908601e-9;
    *gx = 0.0;
    *gy = 0.0;
    if (fabs(tz - 1.0) > 1e-15 && fabs(tz + 1.0) > 1e-15) {
        double g = tx * sx + ty * sy + tz * sz;
        g = g < 1.0 ? g : 1.0; /* MAX(MIN(., 1), -1) */
        g = g > -1.0 ? g : -1.0;
        double mp = 2.0 - 2.0 * g;
        double sqp = sqrt(mp);
        double p1 = (1.0 - 6.21196) / (sqp * mp + 1e-16);
        double p2 = (2.7 + 6.0) * (2.0 * g + sqp) / (2.0 * (g * g - 1.0) + 1e-16);
        double val = (p1 + p2) * cons;
        double x32 = tz * tz;
        double mp2 = sqrt(1.0 - x32);
        *gy = (sz * (1.0 - x32) - tz * (tx * sx + ty * sy)) / mp2 * val;
        *gx = (tx * sy - ty * sx) / mp2 * val;
    }
}

/* cancellation-free restatement of the same kernel: the chord c from coordinate differences,
 * 2g + c = 2 - c^2 + c and 2(g^2 - 1) = -c^2 (4 - c^2) / 2 for unit t, s; (s - g t) is
 * replaced by its tangential part d = s - t.  Regularisers kept. */
static void gfunc_stable(double tx, double ty, double tz, double sx, double sy, double sz, double* gx, double* gy) {
    const double cons = -2.343256857908601e-9;
    *gx = 0.0;
    *gy = 0.0;
    if (fabs(tz - 1.0) > 1e-15 && fabs(tz + 1.0) > 1e-15) {
        double dx = sx - tx, dy = sy - ty, dz = sz - tz;
        double c2 = dx * dx + dy * dy + dz * dz;
        double c = sqrt(c2);
        double p1 = (1.0 - 6.21196) / (c * c2 + 1e-16);
        double p2 = (2.7 + 6.0) * (2.0 - c2 + c) / (-0.5 * c2 * (4.0 - c2) + 1e-16);
        double val = (p1 + p2) * cons;
        double a = 1.0 - tz * tz;
        double cp = sqrt(a);
        *gy = (dz * a - tz * (tx * dx + ty * dy)) / cp * val;
        *gx = (tx * dy - ty * dx) / cp * val;
    }
}

/* ---------------------------------------------------------------- tree */

typedef struct {
    int level, is_leaf, parent, child[4], face, count;
    double b[6]; /* min_xi, mid_xi, max_xi, min_eta, mid_eta, max_eta */
    double radius;
    int* pts;
} or_panel;

struct or_ctx {
    int np;
    or_panel* p;
    int64_t nint;
    int *it, *is, *ty;
    int64_t npc, npp;
    int *pc_t, *pc_s, *pp_t, *pp_s;
};

or_ctx* or_ctx_new(void) {
    or_ctx* c = (or_ctx*)xmalloc(sizeof(or_ctx));
    memset(c, 0, sizeof(*c));
    return c;
}

static void free_tree(or_ctx* c) {
    for (int i = 0; i < c->np; ++i) free(c->p[i].pts);
    free(c->p);
    c->p = NULL;
    c->np = 0;
}
static void free_dtt(or_ctx* c) {
    free(c->it); free(c->is); free(c->ty);
    c->it = c->is = c->ty = NULL;
    c->nint = 0;
}
static void free_stt(or_ctx* c) {
    free(c->pc_t); free(c->pc_s); free(c->pp_t); free(c->pp_s);
    c->pc_t = c->pc_s = c->pp_t = c->pp_s = NULL;
    c->npc = c->npp = 0;
}

void or_ctx_free(or_ctx* c) {
    if (!c) return;
    free_tree(c);
    free_dtt(c);
    free_stt(c);
    free(c);
}

/* tree_traverse, tree_module.f90:59-246 */
int or_tree_build(or_ctx* c, int n, const double* x, const double* y, const double* z, int thresh) {
    free_tree(c);
    free_dtt(c);
    free_stt(c);
    double pi = PI_REF;
    int cap = 64, np = 6;
    or_panel* P = (or_panel*)xmalloc(sizeof(or_panel) * cap);
    for (int i = 0; i < 6; ++i) { /* :76-88 */
        or_panel* q = &P[i];
        memset(q, 0, sizeof(*q));
        q->level = 0;
        q->is_leaf = 1;
        q->parent = -1;
        for (int k = 0; k < 4; ++k) q->child[k] = -1;
        q->face = i + 1;
        q->b[0] = -pi / 4.0; q->b[1] = 0.0; q->b[2] = pi / 4.0;
        q->b[3] = -pi / 4.0; q->b[4] = 0.0; q->b[5] = pi / 4.0;
        q->pts = (int*)xmalloc(sizeof(int) * (size_t)n);
        q->count = 0;
    }
    for (int i = 0; i < n; ++i) { /* :93-101 */
        int f = or_face(x[i], y[i], z[i]);
        P[f - 1].pts[P[f - 1].count++] = i;
    }
    for (int i = 0; i < np; ++i) { /* :126-220, FIFO over panels in id order */
        or_panel* par = &P[i];
        int cnt = par->count;
        if (cnt >= thresh && par->is_leaf) {
            if (np + 4 > cap) {
                cap *= 2;
                P = (or_panel*)xrealloc(P, sizeof(or_panel) * cap);
                par = &P[i];
            }
            par->is_leaf = 0;
            double mnx = par->b[0], mdx = par->b[1], mxx = par->b[2];
            double mne = par->b[3], mde = par->b[4], mxe = par->b[5];
            for (int j = 0; j < 4; ++j) {
                or_panel* q = &P[np + j];
                memset(q, 0, sizeof(*q));
                q->parent = i;
                q->level = par->level + 1;
                q->face = par->face;
                q->is_leaf = 1;
                for (int k = 0; k < 4; ++k) q->child[k] = -1;
                q->pts = (int*)xmalloc(sizeof(int) * (size_t)cnt);
            }
            /* :150-168 */
            double* b1 = P[np + 0].b; b1[0] = mnx; b1[3] = mne; b1[2] = mdx; b1[5] = mde;
            b1[1] = 0.5 * (mnx + mdx); b1[4] = 0.5 * (mne + mde);
            double* b2 = P[np + 1].b; b2[0] = mdx; b2[3] = mne; b2[2] = mxx; b2[5] = mde;
            b2[1] = 0.5 * (mdx + mxx); b2[4] = 0.5 * (mne + mde);
            double* b3 = P[np + 2].b; b3[0] = mdx; b3[3] = mde; b3[2] = mxx; b3[5] = mxe;
            b3[1] = 0.5 * (mdx + mxx); b3[4] = 0.5 * (mde + mxe);
            double* b4 = P[np + 3].b; b4[0] = mnx; b4[3] = mde; b4[2] = mdx; b4[5] = mxe;
            b4[1] = 0.5 * (mnx + mdx); b4[4] = 0.5 * (mde + mxe);
            for (int k = 0; k < 4; ++k) par->child[k] = np + k;
            for (int j = 0; j < cnt; ++j) { /* :175-197 stable partition */
                int idx = par->pts[j];
                double xi, eta;
                or_xieta(x[idx], y[idx], z[idx], par->face, &xi, &eta);
                int w;
                if (xi < mdx) w = (eta < mde) ? 0 : 3;
                else w = (eta < mde) ? 1 : 2;
                or_panel* q = &P[np + w];
                q->pts[q->count++] = idx;
            }
            np += 4;
        }
    }
    for (int i = 0; i < np; ++i) { /* :223-244 radius: max centre-to-corner arc */
        or_panel* q = &P[i];
        double x1, x2, x3, y1, y2, y3, d[4];
        or_xyz(q->b[1], q->b[4], q->face, &x1, &x2, &x3);
        const double cx[4] = {q->b[0], q->b[0], q->b[2], q->b[2]};
        const double ce[4] = {q->b[3], q->b[5], q->b[5], q->b[3]};
        for (int k = 0; k < 4; ++k) {
            or_xyz(cx[k], ce[k], q->face, &y1, &y2, &y3);
            double dot = x1 * y1 + x2 * y2 + x3 * y3;
            dot = dot < 1.0 ? dot : 1.0;
            dot = dot > -1.0 ? dot : -1.0;
            d[k] = acos(dot);
        }
        double r = d[0];
        for (int k = 1; k < 4; ++k) r = d[k] > r ? d[k] : r;
        q->radius = r;
    }
    c->np = np;
    c->p = P;
    return np;
}

int or_tree_n_panels(const or_ctx* c) { return c->np; }

int64_t or_tree_total_inside(const or_ctx* c) {
    int64_t t = 0;
    for (int i = 0; i < c->np; ++i) t += c->p[i].count;
    return t;
}

void or_tree_export(const or_ctx* c, int* level, int* is_leaf, int* parent, int* child, int* face,
                    double* bounds, double* radius, int* count, int64_t* offsets, int* points) {
    int64_t off = 0;
    for (int i = 0; i < c->np; ++i) {
        const or_panel* q = &c->p[i];
        if (level) level[i] = q->level;
        if (is_leaf) is_leaf[i] = q->is_leaf;
        if (parent) parent[i] = q->parent;
        if (child) for (int k = 0; k < 4; ++k) child[4 * i + k] = q->child[k];
        if (face) face[i] = q->face;
        if (bounds) for (int k = 0; k < 6; ++k) bounds[6 * i + k] = q->b[k];
        if (radius) radius[i] = q->radius;
        if (count) count[i] = q->count;
        if (offsets) offsets[i] = off;
        if (points) memcpy(points + off, q->pts, sizeof(int) * (size_t)q->count);
        off += q->count;
    }
    if (offsets) offsets[c->np] = off;
}

/* contains_point, tree_module.f90:39-56 */
static int contains_point(const or_panel* q, double x, double y, double z) {
    int f = or_face(x, y, z);
    if (f != q->face) return 0;
    double xi, eta;
    or_xieta(x, y, z, f, &xi, &eta);
    return xi >= q->b[0] && xi < q->b[2] && eta >= q->b[3] && eta < q->b[5];
}

/* assign_points_to_panels, tree_module.f90:249-273.  Where the reference would probe past the
 * last panel (a point on a face's max edge, which no panel contains) the walk stops instead. */
void or_assign(const or_ctx* c, int n, const double* x, const double* y, const double* z, int nlev, int* pp) {
    for (int i = 0; i < n; ++i) {
        for (int l = 0; l < nlev; ++l) pp[(size_t)i * nlev + l] = -1;
        int level = 0, j = 0;
        while (j != -1 && j < c->np) {
            if (contains_point(&c->p[j], x[i], y[i], z[i])) {
                if (level < nlev) pp[(size_t)i * nlev + level] = j;
                level++;
                j = c->p[j].child[0];
            } else {
                j++;
            }
        }
    }
}

/* ---------------------------------------------------------------- interaction lists */

typedef struct {
    int* a;
    int64_t n, cap;
} ivec;
static void ivec_push(ivec* v, int x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->a = (int*)xrealloc(v->a, sizeof(int) * (size_t)v->cap);
    }
    v->a[v->n++] = x;
}

static double centre_dist(const or_panel* t, const or_panel* s) {
    double x1t, x2t, x3t, x1s, x2s, x3s;
    or_xyz(t->b[1], t->b[4], t->face, &x1t, &x2t, &x3t);
    or_xyz(s->b[1], s->b[4], s->face, &x1s, &x2s, &x3s);
    double dot = x1t * x1s + x2t * x2s + x3t * x3s;
    dot = dot > -1.0 ? dot : -1.0; /* MIN(MAX(., -1), 1), :311 */
    dot = dot < 1.0 ? dot : 1.0;
    return acos(dot);
}

/* DUAL_TREE_TRAVERSAL, tree_module.f90:277-398.  The queue grows on demand (the reference's is
 * fixed at 64 x panels, :288-290, and unchecked). */
int64_t or_dtt(or_ctx* c, double theta, int thresh) {
    free_dtt(c);
    ivec qt = {0}, qs = {0}, ot = {0}, os = {0}, oy = {0};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            ivec_push(&qt, i);
            ivec_push(&qs, j);
        }
    for (int64_t cur = 0; cur < qt.n; ++cur) {
        int it = qt.a[cur], is = qs.a[cur];
        const or_panel* T = &c->p[it];
        const or_panel* S = &c->p[is];
        int ct = T->count, cs = S->count;
        if (ct > 0 && cs > 0) {
            double dist = centre_dist(T, S);
            double sep = (T->radius + S->radius) / dist;
            if (dist > 0 && sep < theta) {
                int ty = 0;
                if (ct > thresh) ty += 2;
                if (cs > thresh) ty += 1;
                ivec_push(&ot, it); ivec_push(&os, is); ivec_push(&oy, ty);
            } else if (ct < thresh && cs < thresh) {
                ivec_push(&ot, it); ivec_push(&os, is); ivec_push(&oy, 0);
            } else if (T->is_leaf && S->is_leaf) {
                ivec_push(&ot, it); ivec_push(&os, is); ivec_push(&oy, 0);
            } else if (T->is_leaf || (!S->is_leaf && ct < cs)) { /* refine source, :340-350, :375-385 */
                for (int k = 0; k < 4; ++k) {
                    ivec_push(&qt, it);
                    ivec_push(&qs, S->child[k]);
                }
            } else { /* refine target, :351-374 */
                for (int k = 0; k < 4; ++k) {
                    ivec_push(&qt, T->child[k]);
                    ivec_push(&qs, is);
                }
            }
        }
    }
    free(qt.a);
    free(qs.a);
    c->nint = ot.n;
    c->it = ot.a;
    c->is = os.a;
    c->ty = oy.a;
    return c->nint;
}

void or_dtt_export(const or_ctx* c, int* it, int* is, int* ty) {
    memcpy(it, c->it, sizeof(int) * (size_t)c->nint);
    memcpy(is, c->is, sizeof(int) * (size_t)c->nint);
    memcpy(ty, c->ty, sizeof(int) * (size_t)c->nint);
}

/* SINGLE_TREE_TRAVERSAL, tree_module.f90:401-487 (param_cluster_thresh is unused there too) */
int or_stt(or_ctx* c, double theta, int thresh, int n, const double* x, const double* y, const double* z,
           int64_t* npc, int64_t* npp) {
    free_stt(c);
    ivec pct = {0}, pcs = {0}, ppt = {0}, pps = {0};
    int* q = (int*)xmalloc(sizeof(int) * (size_t)(c->np + 8));
    for (int i = 0; i < n; ++i) {
        int qn = 6;
        for (int j = 0; j < 6; ++j) q[j] = j;
        double x1t = x[i], x2t = y[i], x3t = z[i];
        for (int cur = 0; cur < qn; ++cur) {
            int is = q[cur];
            const or_panel* S = &c->p[is];
            if (S->count > 0) {
                double x1s, x2s, x3s;
                or_xyz(S->b[1], S->b[4], S->face, &x1s, &x2s, &x3s);
                double dot = x1t * x1s + x2t * x2s + x3t * x3s;
                dot = dot > -1.0 ? dot : -1.0;
                dot = dot < 1.0 ? dot : 1.0;
                double dist = acos(dot);
                double sep = S->radius / dist;
                if (dist > 0 && sep < theta) {
                    ivec_push(&pct, i); ivec_push(&pcs, is);
                } else if (S->is_leaf) {
                    ivec_push(&ppt, i); ivec_push(&pps, is);
                } else {
                    for (int k = 0; k < 4; ++k) q[qn++] = S->child[k];
                }
            }
        }
    }
    free(q);
    c->npc = pct.n; c->pc_t = pct.a; c->pc_s = pcs.a;
    c->npp = ppt.n; c->pp_t = ppt.a; c->pp_s = pps.a;
    *npc = c->npc;
    *npp = c->npp;
    return 0;
}

void or_stt_export(const or_ctx* c, int* pc_t, int* pc_s, int* pp_t, int* pp_s) {
    memcpy(pc_t, c->pc_t, sizeof(int) * (size_t)c->npc);
    memcpy(pc_s, c->pc_s, sizeof(int) * (size_t)c->npc);
    memcpy(pp_t, c->pp_t, sizeof(int) * (size_t)c->npp);
    memcpy(pp_s, c->pp_s, sizeof(int) * (size_t)c->npp);
}

/* ---------------------------------------------------------------- MPAS evaluator */

/* pp_interact, fast_sum_module.f90:33-73 */
static void pp_interact(const or_panel* T, const or_panel* S, const double* x, const double* y, const double* z,
                        const double* area, const double* ssh, double* gx, double* gy) {
    for (int i = 0; i < T->count; ++i) {
        int t = T->pts[i];
        for (int j = 0; j < S->count; ++j) {
            int s = S->pts[j];
            double sx, sy;
            or_gfunc(x[t], y[t], z[t], x[s], y[s], z[s], &sx, &sy);
            gx[t] = gx[t] + sx * area[s] * ssh[s];
            gy[t] = gy[t] + sy * area[s] * ssh[s];
        }
    }
}

/* source-side anterpolation shared by pc_interact (:99-111) and cc_interact (:229-241) */
static void anterp(const or_panel* S, const double* x, const double* y, const double* z, const double* area,
                   const double* ssh, int degree, double* w) {
    int n = degree + 1;
    double bv[64 * 64];
    for (int k = 0; k < n * n; ++k) w[k] = 0.0;
    for (int i = 0; i < S->count; ++i) {
        int s = S->pts[i];
        double xi, eta;
        or_xieta(x[s], y[s], z[s], S->face, &xi, &eta);
        or_interp_vals(xi, eta, S->b[0], S->b[2], S->b[3], S->b[5], degree, bv);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) w[j + k * n] = w[j + k * n] + bv[j + k * n] * area[s] * ssh[s];
    }
}

/* target-side interpolation shared by cp_interact (:186-199) and cc_interact (:274-287) */
static void interp_out(const or_panel* T, const double* x, const double* y, const double* z, int degree,
                       const double* px, const double* py, double* gx, double* gy) {
    int n = degree + 1;
    double bv[64 * 64];
    for (int i = 0; i < T->count; ++i) {
        int t = T->pts[i];
        double xi, eta;
        or_xieta(x[t], y[t], z[t], T->face, &xi, &eta);
        or_interp_vals(xi, eta, T->b[0], T->b[2], T->b[3], T->b[5], degree, bv);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) {
                gx[t] = gx[t] + bv[j + k * n] * px[j + k * n];
                gy[t] = gy[t] + bv[j + k * n] * py[j + k * n];
            }
    }
}

/* pc_interact, fast_sum_module.f90:75-131 */
static void pc_interact(const or_panel* T, const or_panel* S, const double* x, const double* y, const double* z,
                        const double* area, const double* ssh, int degree, double* gx, double* gy) {
    int n = degree + 1;
    double cxi[64], ceta[64], w[64 * 64];
    or_bli_points(S->b[0], S->b[2], degree, cxi);
    or_bli_points(S->b[3], S->b[5], degree, ceta);
    anterp(S, x, y, z, area, ssh, degree, w);
    for (int i = 0; i < T->count; ++i) {
        int t = T->pts[i];
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) {
                double cx, cy, cz, sx, sy;
                or_xyz(cxi[j], ceta[k], S->face, &cx, &cy, &cz);
                or_gfunc(x[t], y[t], z[t], cx, cy, cz, &sx, &sy);
                gx[t] = gx[t] + sx * w[j + k * n];
                gy[t] = gy[t] + sy * w[j + k * n];
            }
    }
}

/* cp_interact, fast_sum_module.f90:133-201 */
static void cp_interact(const or_panel* T, const or_panel* S, const double* x, const double* y, const double* z,
                        const double* area, const double* ssh, int degree, double* gx, double* gy) {
    int n = degree + 1;
    double cxi[64], ceta[64], px[64 * 64], py[64 * 64];
    or_bli_points(T->b[0], T->b[2], degree, cxi);
    or_bli_points(T->b[3], T->b[5], degree, ceta);
    for (int k = 0; k < n * n; ++k) px[k] = py[k] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double cx, cy, cz;
            or_xyz(cxi[i], ceta[j], T->face, &cx, &cy, &cz);
            for (int k = 0; k < S->count; ++k) {
                int s = S->pts[k];
                double sx, sy;
                or_gfunc(cx, cy, cz, x[s], y[s], z[s], &sx, &sy);
                px[i + j * n] = px[i + j * n] + sx * area[s] * ssh[s];
                py[i + j * n] = py[i + j * n] + sy * area[s] * ssh[s];
            }
        }
    interp_out(T, x, y, z, degree, px, py, gx, gy);
}

/* cc_interact, fast_sum_module.f90:204-288 */
static void cc_interact(const or_panel* T, const or_panel* S, const double* x, const double* y, const double* z,
                        const double* area, const double* ssh, int degree, double* gx, double* gy) {
    int n = degree + 1;
    double sxi[64], seta[64], txi[64], teta[64], w[64 * 64], px[64 * 64], py[64 * 64];
    or_bli_points(S->b[0], S->b[2], degree, sxi);
    or_bli_points(S->b[3], S->b[5], degree, seta);
    anterp(S, x, y, z, area, ssh, degree, w);
    or_bli_points(T->b[0], T->b[2], degree, txi);
    or_bli_points(T->b[3], T->b[5], degree, teta);
    for (int k = 0; k < n * n; ++k) px[k] = py[k] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double ctx, cty, ctz;
            or_xyz(txi[i], teta[j], T->face, &ctx, &cty, &ctz);
            for (int l = 0; l < n; ++l)
                for (int k = 0; k < n; ++k) {
                    double csx, csy, csz, sx, sy;
                    or_xyz(sxi[k], seta[l], S->face, &csx, &csy, &csz);
                    or_gfunc(ctx, cty, ctz, csx, csy, csz, &sx, &sy);
                    px[i + j * n] = px[i + j * n] + sx * w[k + l * n];
                    py[i + j * n] = py[i + j * n] + sy * w[k + l * n];
                }
        }
    interp_out(T, x, y, z, degree, px, py, gx, gy);
}

void or_interact(const or_ctx* c, int kind, int it, int is, const double* x, const double* y, const double* z,
                 const double* area, const double* ssh, int degree, double* gx, double* gy) {
    const or_panel* T = &c->p[it];
    const or_panel* S = &c->p[is];
    switch (kind) {
        case 0: pp_interact(T, S, x, y, z, area, ssh, gx, gy); break;
        case 1: pc_interact(T, S, x, y, z, area, ssh, degree, gx, gy); break;
        case 2: cp_interact(T, S, x, y, z, area, ssh, degree, gx, gy); break;
        case 3: cc_interact(T, S, x, y, z, area, ssh, degree, gx, gy); break;
        default: break;
    }
}

/* fast_sum, fast_sum_module.f90:291-332: interaction i (1-based) belongs to rank MOD(i, nproc) */
void or_fast_sum(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                 const double* area, const double* ssh, int degree, int rank, int nproc, double* gx, double* gy) {
    for (int64_t i = 1; i <= c->nint; ++i) {
        if (i % nproc != rank) continue;
        or_interact(c, c->ty[i - 1], c->it[i - 1], c->is[i - 1], x, y, z, area, ssh, degree, gx, gy);
    }
}

/* ---------------------------------------------------------------- MOM6 evaluator */

/* proxy_source_compute, mom6_fast_sum_module.f90:33-74: w[(panel)*m + (k)*(d+1) + l] gets
 * basis(k, l) with k the xi index (xi-major) */
void or_proxy_source(const or_ctx* c, int n, int nlev, const int* pp, const double* x, const double* y,
                     const double* z, const double* ssh, const double* area, int degree, double* w) {
    int nd = degree + 1, m = nd * nd;
    double bv[64 * 64];
    for (int64_t k = 0; k < (int64_t)m * c->np; ++k) w[k] = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < nlev; ++j) {
            int it = pp[(size_t)i * nlev + j];
            if (it == -1) break;
            const or_panel* T = &c->p[it];
            double xi, eta;
            or_xieta(x[i], y[i], z[i], T->face, &xi, &eta);
            or_interp_vals(xi, eta, T->b[0], T->b[2], T->b[3], T->b[5], degree, bv);
            int64_t shift = (int64_t)it * m;
            int off = 0;
            for (int k = 0; k < nd; ++k)
                for (int l = 0; l < nd; ++l) {
                    w[shift + off] = w[shift + off] + bv[k + l * nd] * area[i] * ssh[i];
                    off++;
                }
        }
    }
}

/* pc_compute / pc_interact (mom6_fast_sum_module.f90:153-174, :76-107), then pp_compute /
 * pp_interact (:177-192, :109-150).  pc_interact reads the weights eta-major (offset
 * (j-1)(d+1)+i for node (xi_i, eta_j)), transposed against proxy_source_compute: kept as is. */
void or_mom6_sums_sel(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                      const double* area, const double* ssh, int degree, const double* w,
                      const unsigned char* sel, double* gx, double* gy) {
    int nd = degree + 1, m = nd * nd;
    for (int64_t e = 0; e < c->npc; ++e) {
        int it = c->pc_t[e], is = c->pc_s[e];
        if (sel && !sel[it]) continue;
        const or_panel* S = &c->p[is];
        double cxi[64], ceta[64];
        or_bli_points(S->b[0], S->b[2], degree, cxi);
        or_bli_points(S->b[3], S->b[5], degree, ceta);
        const double* ws = w + (int64_t)is * m;
        int off = 0;
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i < nd; ++i) {
                double cx, cy, cz, sx, sy;
                or_xyz(cxi[i], ceta[j], S->face, &cx, &cy, &cz);
                or_gfunc(x[it], y[it], z[it], cx, cy, cz, &sx, &sy);
                gx[it] = gx[it] + sx * ws[off];
                gy[it] = gy[it] + sy * ws[off];
                off++;
            }
    }
    for (int64_t e = 0; e < c->npp; ++e) {
        int it = c->pp_t[e], is = c->pp_s[e];
        if (sel && !sel[it]) continue;
        const or_panel* S = &c->p[is];
        for (int i = 0; i < S->count; ++i) {
            int s = S->pts[i]; /* relabeled_points_inside == points_inside at one rank */
            double sx, sy;
            or_gfunc(x[it], y[it], z[it], x[s], y[s], z[s], &sx, &sy);
            gx[it] = gx[it] + ssh[s] * area[s] * sx;
            gy[it] = gy[it] + ssh[s] * area[s] * sy;
        }
    }
}

/* every target; or_mom6_sums_sel: only the targets with sel[t] != 0 (the others keep their sums) */
void or_mom6_sums(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                  const double* area, const double* ssh, int degree, const double* w, double* gx, double* gy) {
    or_mom6_sums_sel(c, n, x, y, z, area, ssh, degree, w, 0, gx, gy);
}


/* ---------------------------------------------------------------- accuracy mode (not in the reference)
 *
 * Same tree, lists and proxy grids as fast_sum, three changes: (1) every pair uses the
 * cancellation-free kernel form, (2) the self pair is skipped, (3) cluster targets (CP, CC) collect
 * the tangential vector T(z) = sum val(z, s) q_s P_z (s - z) in Cartesian components at their
 * proxies.  T is smooth over a whole cube face, where the projected east/north components (which
 * carry 1/cos(lat)) are not, and it has no radial part to amplify interpolation error after the
 * projection; each target interpolates T and projects it on its own east/north basis.  For unit
 * vectors P_t s = P_t (s - g t), so direct terms are unchanged.  Pinned by convergence in the degree to the direct sum (mode 2) and by equality with
 * or_fast_sum's PP/PC terms up to the kernel form. */

static double val_stable(double tx, double ty, double tz, double sx, double sy, double sz) {
    const double cons = -2.343256857908601e-9;
    double dx = sx - tx, dy = sy - ty, dz = sz - tz;
    double c2 = dx * dx + dy * dy + dz * dz;
    double c = sqrt(c2);
    double p1 = (1.0 - 6.21196) / (c * c2 + 1e-16);
    double p2 = (2.7 + 6.0) * (2.0 - c2 + c) / (-0.5 * c2 * (4.0 - c2) + 1e-16);
    return (p1 + p2) * cons;
}

/* v += q * P_z (s - z): the tangential part at z of the direction to s */
static void tangential_add(double zx, double zy, double zz, double sx, double sy, double sz, double q, double* v) {
    double dx = sx - zx, dy = sy - zy, dz = sz - zz;
    double r = dx * zx + dy * zy + dz * zz;
    v[0] += q * (dx - r * zx);
    v[1] += q * (dy - r * zy);
    v[2] += q * (dz - r * zz);
}

static void project(double tx, double ty, double tz, double vx, double vy, double vz, double* ge, double* gn) {
    *ge = 0.0;
    *gn = 0.0;
    if (fabs(tz - 1.0) > 1e-15 && fabs(tz + 1.0) > 1e-15) {
        double a = 1.0 - tz * tz, cp = sqrt(a);
        *ge = (tx * vy - ty * vx) / cp;
        *gn = (vz * a - tz * (tx * vx + ty * vy)) / cp;
    }
}

void or_fast_sum_cart_mask(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                      const double* area, const double* ssh, int degree, int mask, double* gx, double* gy) {
    int nd = degree + 1, m = nd * nd;
    double* V = (double*)xmalloc(sizeof(double) * 3 * (size_t)m * (size_t)c->np);
    int* hasV = (int*)xmalloc(sizeof(int) * (size_t)c->np);
    memset(V, 0, sizeof(double) * 3 * (size_t)m * (size_t)c->np);
    memset(hasV, 0, sizeof(int) * (size_t)c->np);
    double w[64 * 64], sxi[64], seta[64], txi[64], teta[64], bv[64 * 64];
    for (int64_t e = 0; e < c->nint; ++e) {
        const or_panel* T = &c->p[c->it[e]];
        const or_panel* S = &c->p[c->is[e]];
        int ty = c->ty[e];
        if (!((mask >> ty) & 1)) continue;
        if (ty == 0) {
            for (int i = 0; i < T->count; ++i) {
                int t = T->pts[i];
                for (int j = 0; j < S->count; ++j) {
                    int s = S->pts[j];
                    if (s == t) continue;
                    double sx, sy;
                    gfunc_stable(x[t], y[t], z[t], x[s], y[s], z[s], &sx, &sy);
                    gx[t] += sx * area[s] * ssh[s];
                    gy[t] += sy * area[s] * ssh[s];
                }
            }
            continue;
        }
        if (ty == 1 || ty == 3) {
            or_bli_points(S->b[0], S->b[2], degree, sxi);
            or_bli_points(S->b[3], S->b[5], degree, seta);
            anterp(S, x, y, z, area, ssh, degree, w);
        }
        if (ty == 1) {
            for (int i = 0; i < T->count; ++i) {
                int t = T->pts[i];
                for (int k = 0; k < nd; ++k)
                    for (int j = 0; j < nd; ++j) {
                        double cx, cy, cz, sx, sy;
                        or_xyz(sxi[j], seta[k], S->face, &cx, &cy, &cz);
                        gfunc_stable(x[t], y[t], z[t], cx, cy, cz, &sx, &sy);
                        gx[t] += sx * w[j + k * nd];
                        gy[t] += sy * w[j + k * nd];
                    }
            }
            continue;
        }
        double* Vt = V + 3 * (size_t)m * (size_t)c->it[e];
        hasV[c->it[e]] = 1;
        or_bli_points(T->b[0], T->b[2], degree, txi);
        or_bli_points(T->b[3], T->b[5], degree, teta);
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i < nd; ++i) {
                double zx, zy, zz;
                or_xyz(txi[i], teta[j], T->face, &zx, &zy, &zz);
                double* v = Vt + 3 * (i + j * nd);
                if (ty == 2) {
                    for (int k = 0; k < S->count; ++k) {
                        int s = S->pts[k];
                        double q = val_stable(zx, zy, zz, x[s], y[s], z[s]) * area[s] * ssh[s];
                        tangential_add(zx, zy, zz, x[s], y[s], z[s], q, v);
                    }
                } else {
                    for (int l = 0; l < nd; ++l)
                        for (int k = 0; k < nd; ++k) {
                            double cx, cy, cz;
                            or_xyz(sxi[k], seta[l], S->face, &cx, &cy, &cz);
                            double q = val_stable(zx, zy, zz, cx, cy, cz) * w[k + l * nd];
                            tangential_add(zx, zy, zz, cx, cy, cz, q, v);
                        }
                }
            }
    }
    for (int p = 0; p < c->np; ++p) {
        if (!hasV[p]) continue;
        const or_panel* T = &c->p[p];
        const double* Vt = V + 3 * (size_t)m * (size_t)p;
        for (int i = 0; i < T->count; ++i) {
            int t = T->pts[i];
            double xi, eta, vx = 0, vy = 0, vz = 0, ge, gn;
            or_xieta(x[t], y[t], z[t], T->face, &xi, &eta);
            or_interp_vals(xi, eta, T->b[0], T->b[2], T->b[3], T->b[5], degree, bv);
            for (int k = 0; k < m; ++k) {
                vx += bv[k] * Vt[3 * k];
                vy += bv[k] * Vt[3 * k + 1];
                vz += bv[k] * Vt[3 * k + 2];
            }
            project(x[t], y[t], z[t], vx, vy, vz, &ge, &gn);
            gx[t] += ge;
            gy[t] += gn;
        }
    }
    free(V);
    free(hasV);
}

void or_fast_sum_cart(const or_ctx* c, int n, const double* x, const double* y, const double* z,
                      const double* area, const double* ssh, int degree, double* gx, double* gy) {
    or_fast_sum_cart_mask(c, n, x, y, z, area, ssh, degree, 15, gx, gy);
}

/* ---------------------------------------------------------------- direct sum */

void or_direct_sum(int n, const double* x, const double* y, const double* z, const double* area,
                   const double* ssh, int nt, const int* tgt, int mode, double* gx, double* gy) {
    for (int a = 0; a < nt; ++a) {
        int t = tgt[a];
        double ax = 0.0, ay = 0.0;
        for (int s = 0; s < n; ++s) {
            if (mode != 0 && s == t) continue;
            double sx, sy;
            if (mode == 2) gfunc_stable(x[t], y[t], z[t], x[s], y[s], z[s], &sx, &sy);
            else or_gfunc(x[t], y[t], z[t], x[s], y[s], z[s], &sx, &sy);
            ax = ax + sx * area[s] * ssh[s];
            ay = ay + sy * area[s] * ssh[s];
        }
        gx[a] = ax;
        gy[a] = ay;
    }
}
