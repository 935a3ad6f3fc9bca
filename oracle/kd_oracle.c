/*
 * kd_oracle.c -- CPU restatement of corruscant's k-d tree build, radius pair
 * count and RA/DEC conversion.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (corruscant_amd/) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function states the reference lines it follows.  Arithmetic that the
 * reference leaves to GCC's -O3 -march=native contraction is written with
 * explicit fma() so that no compiler flag can change the result:
 *   - squared distance   n = fma(d2,d2, fma(d1,d1, d0*d0)),  d_i = p_i - q_i
 *     (src/kdquery.c:39-47, contracted by GCC into vfmadd132sd)
 *   - variance           var = fma(t,t,var),  t = v - avg
 *     (src/kdbuild.c:378-385, contracted into vfmadd231sd)
 *
 * Node record layout (40 bytes) follows src/kdtree.h:13-25 as laid out by GCC
 * on x86-64: value[3] at 0, field byte at 24, weight byte at 25, child flags
 * and split dim in the 32-bit word at 32 (bit 0 has_lchild, bit 1 has_rchild,
 * bits 2..9 dim).  Padding bytes are zero here.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NDIM 3

typedef struct {
    double value[NDIM];
    uint8_t field;
    uint8_t weight;
    uint8_t pad0[6];
    uint32_t bits;
    uint32_t pad1;
} or_node_t;

typedef struct {
    double value[NDIM];
    int field;  /* already truncated to 8 bits */
    int weight; /* already truncated to 8 bits */
    int first;  /* lowest original index of the duplicate run */
} or_datum_t;

typedef struct {
    or_node_t *nodes; /* memsize records, slot 0 unused */
    int size;         /* number of unique points */
    int memsize;
    int num_fields;
    or_datum_t *unique; /* size records, (x,y,z) order */
    int *args[NDIM];    /* presorted argsort arrays (before the build) */
    int *perm;          /* heap index -> unique index, -1 for empty slots */
    int *inorder;       /* heap index -> in-order position (the segment median) */
} or_tree_t;

/* ------------------------------------------------------------------------ */
/* comparisons: IEEE ==/< semantics, so -0.0 == +0.0 (src/kdbuild.c:59-96)   */

typedef struct {
    const double *xyz;
} xyz_ctx_t;

/* lexicographic (x,y,z) with ==/<; ties broken by index so that the first of
 * each run of equal points is its lowest original index (src/kdbuild.c:98-152
 * keeps the left-run element first on ties, which gives the same first). */
static int cmp_xyz_idx(const void *pa, const void *pb, void *vctx)
{
    const double *xyz = ((xyz_ctx_t *)vctx)->xyz;
    int a = *(const int *)pa, b = *(const int *)pb;
    for (int d = 0; d < NDIM; d++) {
        double av = xyz[3 * (size_t)a + d], bv = xyz[3 * (size_t)b + d];
        if (av == bv) continue;
        return av < bv ? -1 : 1;
    }
    return (a > b) - (a < b);
}

typedef struct {
    const or_datum_t *u;
    int msd;
} super_ctx_t;

/* cyclic superkey (v[msd], v[msd+1], v[msd+2]) (src/kdbuild.c:59-80) */
static inline int superkey_less(const or_datum_t *u, int msd, int a, int b)
{
    for (int i = 0; i < NDIM; i++) {
        int d = (msd + i) % NDIM;
        double av = u[a].value[d], bv = u[b].value[d];
        if (av == bv) continue;
        return av < bv;
    }
    return 0; /* only reachable for equal points, excluded after dedup */
}

static int cmp_super(const void *pa, const void *pb, void *vctx)
{
    super_ctx_t *c = (super_ctx_t *)vctx;
    int a = *(const int *)pa, b = *(const int *)pb;
    if (superkey_less(c->u, c->msd, a, b)) return -1;
    if (superkey_less(c->u, c->msd, b, a)) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* build                                                                    */

typedef struct {
    or_tree_t *t;
    int *args[NDIM];
    int *scratch;
} build_ctx_t;

/* src/kdbuild.c:308-351: sequential FP64 mean per dim in args[j] order,
 * division by the segment size, sequential fma variance, strict > argmax. */
static int or_max_variance_dim(build_ctx_t *c, int l, int r)
{
    const or_datum_t *u = c->t->unique;
    double avg[NDIM] = {0, 0, 0}, var[NDIM] = {0, 0, 0};
    for (int j = 0; j < NDIM; j++) {
        const int *a = c->args[j];
        double s = 0.0;
        for (int i = l; i <= r; i++) s += u[a[i]].value[j];
        avg[j] = s;
    }
    for (int j = 0; j < NDIM; j++) avg[j] /= (double)(r - l + 1);
    for (int j = 0; j < NDIM; j++) {
        const int *a = c->args[j];
        double v = 0.0;
        for (int i = l; i <= r; i++) {
            double t = u[a[i]].value[j] - avg[j];
            v = fma(t, t, v);
        }
        var[j] = v;
    }
    int dim = 0;
    double best = var[0];
    for (int j = 1; j < NDIM; j++)
        if (var[j] > best) {
            best = var[j];
            dim = j;
        }
    return dim;
}

/* src/kdbuild.c:251-291: stable partition of args[d][l..r] (d != tdim) into
 * the points below / above the median in superkey tdim; median slot skipped. */
static void or_partition(build_ctx_t *c, int l, int r, int tdim)
{
    const or_datum_t *u = c->t->unique;
    int med = (l + r + 1) / 2;
    int mu = c->args[tdim][med];
    for (int d = 0; d < NDIM; d++) {
        if (d == tdim) continue;
        int *a = c->args[d];
        int lo = 0, hi = med - l + 1;
        for (int i = l; i <= r; i++) {
            int x = a[i];
            if (x == mu) continue;
            if (superkey_less(u, tdim, x, mu))
                c->scratch[lo++] = x;
            else
                c->scratch[hi++] = x;
        }
        /* the median slot keeps whatever it held; it is never read again */
        c->scratch[med - l] = a[med];
        memcpy(a + l, c->scratch, sizeof(int) * (size_t)(r - l + 1));
    }
}

static void set_node(or_tree_t *t, int ind, int uidx, int tdim, int hasl, int hasr, int inord)
{
    or_node_t *nd = &t->nodes[ind];
    const or_datum_t *d = &t->unique[uidx];
    memset(nd, 0, sizeof(*nd));
    for (int k = 0; k < NDIM; k++) nd->value[k] = d->value[k];
    nd->field = (uint8_t)d->field;
    nd->weight = (uint8_t)d->weight;
    nd->bits = (uint32_t)(hasl & 1) | ((uint32_t)(hasr & 1) << 1) | ((uint32_t)(tdim & 0xFF) << 2);
    t->perm[ind] = uidx;
    t->inorder[ind] = inord;
}

/* src/kdbuild.c:373-433 (spec: med = (l+r+1)/2; split dim by variance only
 * for segments of 4 or more; sizes 1, 2, 3 are base cases) */
static void or_build(build_ctx_t *c, int ind, int l, int r)
{
    int med = (l + r + 1) / 2;
    int tdim = 0;
    if (r - l > 2) tdim = or_max_variance_dim(c, l, r);
    int uidx = c->args[tdim][med];
    switch (r - l) {
    case 0:
        set_node(c->t, ind, uidx, tdim, 0, 0, med);
        return;
    case 1:
        set_node(c->t, ind, uidx, tdim, 1, 0, med);
        or_build(c, 2 * ind, l, l);
        return;
    case 2:
        set_node(c->t, ind, uidx, tdim, 1, 1, med);
        or_build(c, 2 * ind, l, l);
        or_build(c, 2 * ind + 1, r, r);
        return;
    default:
        break;
    }
    or_partition(c, l, r, tdim);
    set_node(c->t, ind, uidx, tdim, 1, 1, med);
    or_build(c, 2 * ind, l, med - 1);
    or_build(c, 2 * ind + 1, med + 1, r);
}

/* src/kdbuild.c:435-449: 2^(floor(log2 n) + 1) for n >= 1 */
static int or_pow2ceil(int n)
{
    int p = 1;
    while (n > 1) {
        n >>= 1;
        p <<= 1;
    }
    return p * 2;
}

void or_tree_free(or_tree_t *t)
{
    if (!t) return;
    free(t->nodes);
    free(t->unique);
    for (int d = 0; d < NDIM; d++) free(t->args[d]);
    free(t->perm);
    free(t->inorder);
    memset(t, 0, sizeof(*t));
}

/*
 * src/kdbuild.c:451-537 (tree_construct).  Returns 0 on success, 1 when a run
 * of 256 or more identical points overflows the 8-bit weight (the reference
 * prints "Invalid weights" and exits, src/kdbuild.c:501-511), 2 on a NaN
 * coordinate (unsupported by the reference's comparisons), 3 on n < 1.
 */
int or_tree_build(const double *xyz, const int *fields, int n, int num_fields, or_tree_t *t)
{
    memset(t, 0, sizeof(*t));
    if (n < 1) return 3;
    for (size_t i = 0; i < 3 * (size_t)n; i++)
        if (xyz[i] != xyz[i]) return 2;

    int *ord = malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) ord[i] = i;
    xyz_ctx_t xc = {xyz};
    qsort_r(ord, (size_t)n, sizeof(int), cmp_xyz_idx, &xc);

    or_datum_t *u = malloc(sizeof(or_datum_t) * (size_t)n);
    int nu = 0;
    for (int i = 0; i < n; i++) {
        int p = ord[i];
        const double *v = xyz + 3 * (size_t)p;
        if (nu > 0) {
            const double *w = u[nu - 1].value;
            if (v[0] == w[0] && v[1] == w[1] && v[2] == w[2]) {
                u[nu - 1].weight++;
                continue;
            }
        }
        u[nu].value[0] = v[0];
        u[nu].value[1] = v[1];
        u[nu].value[2] = v[2];
        u[nu].field = fields ? (fields[p] & 0xFF) : 0;
        u[nu].weight = 1;
        u[nu].first = p;
        nu++;
    }
    free(ord);
    for (int i = 0; i < nu; i++)
        if (u[i].weight > 255) {
            free(u);
            return 1;
        }

    t->size = nu;
    t->memsize = or_pow2ceil(nu);
    t->num_fields = num_fields;
    t->unique = u;
    t->nodes = calloc((size_t)t->memsize, sizeof(or_node_t));
    t->perm = malloc(sizeof(int) * (size_t)t->memsize);
    t->inorder = malloc(sizeof(int) * (size_t)t->memsize);
    for (int i = 0; i < t->memsize; i++) t->perm[i] = t->inorder[i] = -1;

    build_ctx_t c;
    c.t = t;
    for (int d = 0; d < NDIM; d++) {
        int *a = malloc(sizeof(int) * (size_t)nu);
        for (int i = 0; i < nu; i++) a[i] = i;
        if (d > 0) {
            super_ctx_t sc = {u, d};
            qsort_r(a, (size_t)nu, sizeof(int), cmp_super, &sc);
        }
        t->args[d] = a;
        c.args[d] = malloc(sizeof(int) * (size_t)nu);
        memcpy(c.args[d], a, sizeof(int) * (size_t)nu);
    }
    c.scratch = malloc(sizeof(int) * (size_t)nu);
    or_build(&c, 1, 0, nu - 1);
    for (int d = 0; d < NDIM; d++) free(c.args[d]);
    free(c.scratch);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* pair count                                                               */

static inline double or_norm2(const double *p, const double *q)
{
    double d0 = p[0] - q[0], d1 = p[1] - q[1], d2 = p[2] - q[2];
    return fma(d2, d2, fma(d1, d1, d0 * d0));
}

typedef struct {
    const or_node_t *nodes;
    int memsize;
    const double *q;
    const int *qf;
    int nq_fields;
    double r;
    int begin, end;
    long long *counts;
    int ncounts;
    int bad;
    long long visits;
} pc_job_t;

/* src/kdquery.c:62-113 for one query point, with an explicit stack */
static long long or_radius_one(const pc_job_t *j, int qi, long long *counts, int *bad)
{
    long long visits = 0;
    const double *q = j->q + 3 * (size_t)qi;
    int qf = j->qf ? j->qf[qi] : 0;
    double r = j->r, rsq = r * r;
    int stack[128];
    int sp = 0;
    stack[sp++] = 1;
    while (sp) {
        int pi = stack[--sp];
        const or_node_t *p = &j->nodes[pi];
        double n = or_norm2(p->value, q);
        visits++;
        int hasl = p->bits & 1, hasr = (p->bits >> 1) & 1;
        int idx = qf * j->nq_fields + p->field;
        int counted = 0;
        if (!hasl) {
            counted = 1;
        } else if (!hasr) {
            counted = 1;
            stack[sp++] = 2 * pi;
        } else {
            int k = (p->bits >> 2) & 0xFF;
            double val = q[k];
            double up = val + r, lo = val - r, s = p->value[k];
            if (up < s) {
                stack[sp++] = 2 * pi;
            } else if (lo > s) {
                stack[sp++] = 2 * pi + 1;
            } else {
                counted = 1;
                stack[sp++] = 2 * pi + 1;
                stack[sp++] = 2 * pi;
            }
        }
        if (counted && n < rsq) {
            if (idx < 0 || idx >= j->ncounts) {
                *bad = 1;
                continue;
            }
            counts[idx] += p->weight;
        }
    }
    return visits;
}

static void *pc_thread(void *arg)
{
    pc_job_t *j = (pc_job_t *)arg;
    for (int i = j->begin; i < j->end; i++) j->visits += or_radius_one(j, i, j->counts, &j->bad);
    return NULL;
}

/*
 * src/kdquery.c:165-249 (pair_count): cumulative weighted counts of (query
 * point, tree node) pairs with n < r*r, reached by the reference's pruning,
 * as a [qf][tf] matrix with row stride nq_fields.  The result is an integer
 * sum, so the thread split does not change it.  Returns 0, or 1 when a field
 * index falls outside the matrix (undefined behaviour in the reference).
 */
int or_pair_count_v(const or_tree_t *t, const double *q, const int *qf, int nq, int nq_fields,
                    double r, int nthreads, long long *out, long long *visits_out)
{
    int ncounts = nq_fields * t->num_fields;
    memset(out, 0, sizeof(long long) * (size_t)ncounts);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nq) nthreads = nq > 0 ? nq : 1;
    pthread_t th[256];
    if (nthreads > 256) nthreads = 256;
    pc_job_t jobs[256];
    long long *bufs = calloc((size_t)nthreads * (size_t)ncounts, sizeof(long long));
    for (int i = 0; i < nthreads; i++) {
        pc_job_t *j = &jobs[i];
        j->nodes = t->nodes;
        j->memsize = t->memsize;
        j->q = q;
        j->qf = qf;
        j->nq_fields = nq_fields;
        j->r = r;
        j->begin = (int)((long long)nq * i / nthreads);
        j->end = (int)((long long)nq * (i + 1) / nthreads);
        j->counts = bufs + (size_t)i * (size_t)ncounts;
        j->ncounts = ncounts;
        j->bad = 0;
        j->visits = 0;
        pthread_create(&th[i], NULL, pc_thread, j);
    }
    int bad = 0;
    long long visits = 0;
    for (int i = 0; i < nthreads; i++) {
        pthread_join(th[i], NULL);
        bad |= jobs[i].bad;
        visits += jobs[i].visits;
        for (int k = 0; k < ncounts; k++) out[k] += jobs[i].counts[k];
    }
    free(bufs);
    if (visits_out) *visits_out = visits;
    return bad;
}

/* the count alone (visits = distances evaluated by the traversal, optional) */
int or_pair_count(const or_tree_t *t, const double *q, const int *qf, int nq, int nq_fields,
                  double r, int nthreads, long long *out)
{
    return or_pair_count_v(t, q, qf, nq, nq_fields, r, nthreads, out, NULL);
}

/*
 * Independent O(N_q * N_nodes) formulation of the same count, for tiny
 * inputs: every node p is tested with the exact predicate n < r*r, and is
 * counted only if the reference's traversal reaches and counts it, which is
 * decided by walking p's ancestors (the reach predicate of src/kdquery.c:86-110):
 * a left child is reached iff !(fl(q_k - r) > s), a right child iff
 * !(fl(q_k + r) < s), a one-child node always passes to its left child, and a
 * two-child node counts itself only when both of its own tests pass.
 * With use_reach = 0 the pruning is ignored (the plain exact count).
 */
int or_pair_count_brute(const or_tree_t *t, const double *q, const int *qf, int nq, int nq_fields,
                        double r, int use_reach, long long *out)
{
    int ncounts = nq_fields * t->num_fields;
    memset(out, 0, sizeof(long long) * (size_t)ncounts);
    double rsq = r * r;
    int bad = 0;
    for (int qi = 0; qi < nq; qi++) {
        const double *qq = q + 3 * (size_t)qi;
        int f = qf ? qf[qi] : 0;
        for (int pi = 1; pi < t->memsize; pi++) {
            if (t->perm[pi] < 0) continue;
            const or_node_t *p = &t->nodes[pi];
            double n = or_norm2(p->value, qq);
            if (!(n < rsq)) continue;
            int ok = 1;
            if (use_reach) {
                if ((p->bits & 3) == 3) {
                    int k = (p->bits >> 2) & 0xFF;
                    double s = p->value[k];
                    if ((qq[k] + r) < s || (qq[k] - r) > s) ok = 0;
                }
                int c = pi;
                while (ok && c > 1) {
                    int a = c >> 1;
                    const or_node_t *an = &t->nodes[a];
                    if ((an->bits & 3) == 3) {
                        int k = (an->bits >> 2) & 0xFF;
                        double s = an->value[k];
                        if ((c & 1) == 0) {
                            if ((qq[k] - r) > s) ok = 0;
                        } else {
                            if ((qq[k] + r) < s) ok = 0;
                        }
                    }
                    c = a;
                }
            }
            if (!ok) continue;
            int idx = f * nq_fields + p->field;
            if (idx < 0 || idx >= ncounts) {
                bad = 1;
                continue;
            }
            out[idx] += p->weight;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* coordinates (src/coordmath.c:38-80)                                      */

#define OR_PI_OVER_180 (M_PI / 180.0)

void or_radec2cart64(const double *ra, const double *dec, double *xyz, int n)
{
    for (int i = 0; i < n; i++) {
        double sl, cl, so, co;
        sincos(dec[i] * OR_PI_OVER_180, &sl, &cl);
        sincos(ra[i] * OR_PI_OVER_180, &so, &co);
        xyz[3 * (size_t)i] = cl * co;
        xyz[3 * (size_t)i + 1] = cl * so;
        xyz[3 * (size_t)i + 2] = sl;
    }
}

void or_radecdist2cart64(const double *ra, const double *dec, const double *dist, double *xyz, int n)
{
    for (int i = 0; i < n; i++) {
        double sl, cl, so, co;
        sincos(dec[i] * OR_PI_OVER_180, &sl, &cl);
        sincos(ra[i] * OR_PI_OVER_180, &so, &co);
        xyz[3 * (size_t)i] = dist[i] * cl * co;
        xyz[3 * (size_t)i + 1] = dist[i] * cl * so;
        xyz[3 * (size_t)i + 2] = dist[i] * sl;
    }
}

void or_radec2sph64(const double *ra, const double *dec, double *theta, double *phi, int n)
{
    for (int i = 0; i < n; i++) {
        theta[i] = (90.0 - dec[i]) * OR_PI_OVER_180;
        phi[i] = ra[i] * OR_PI_OVER_180;
    }
}

/* small accessors for the ctypes wrapper */
int or_tree_size(const or_tree_t *t) { return t->size; }
int or_tree_memsize(const or_tree_t *t) { return t->memsize; }
size_t or_tree_struct_size(void) { return sizeof(or_tree_t); }
size_t or_node_struct_size(void) { return sizeof(or_node_t); }
