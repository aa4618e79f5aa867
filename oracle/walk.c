/*
 * Oracle (test infrastructure only): the reference's two-level software walk restated in C.
 *   pathtrace.wgsl:732-740 ray_aabb          (units.rs:965-976)
 *   pathtrace.wgsl:742-760 intersect_tri     (units.rs:979-1002)
 *   units.rs:1004-1006     Kernel::inv
 *   units.rs:1008-1015     proc_dir / proc_point (pathtrace.wgsl proc_dir)
 *   trace_units.wgsl:44-135 trace_units      (units.rs:1031-1122), every read bounds-checked, steps budgeted
 *   units.rs:1369-1388     Scene::brute
 * glam Vec3A semantics: dot = (x*x' + y*y') + z*z'; cross without FMA; lane min/max as SSE2.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

static inline float smin(float a, float b) { return a < b ? a : b; }
static inline float smax(float a, float b) { return a > b ? a : b; }
static inline float fmaxnum(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a > b ? a : b)); }
static inline float fminnum(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a < b ? a : b)); }
static inline uint32_t fbits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static inline float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const float *a, const float *b, float *r) {
    r[0] = a[1] * b[2] - b[1] * a[2];
    r[1] = a[2] * b[0] - b[2] * a[0];
    r[2] = a[0] * b[1] - b[0] * a[1];
}

float or_ray_aabb(const float *o, const float *inv, const float *bmin, const float *bmax, float tmax) {
    float t0[3], t1[3], lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        t0[k] = (bmin[k] - o[k]) * inv[k];
        t1[k] = (bmax[k] - o[k]) * inv[k];
        lo[k] = smin(t0[k], t1[k]);
        hi[k] = smax(t0[k], t1[k]);
    }
    float tn = fmaxnum(fmaxnum(lo[0], lo[1]), fmaxnum(lo[2], 0.0f));
    float tf = fminnum(fminnum(hi[0], hi[1]), fminnum(hi[2], tmax));
    return tn <= tf ? tn : OR_INF;
}

/* The EXACT walk's box test: the box widened by 2^-8 of its largest extent plus 2^-20 of its largest coordinate. */
static float ray_aabb_wide(const float *o, const float *inv, const float *bmin, const float *bmax, float tmax) {
    float ext = fmaxf(fmaxf(bmax[0] - bmin[0], bmax[1] - bmin[1]), bmax[2] - bmin[2]);
    float mag = 0.0f;
    for (int k = 0; k < 3; k++) mag = fmaxf(mag, fmaxf(fabsf(bmin[k]), fabsf(bmax[k])));
    float m = ext * 0.00390625f + mag * 9.5367431640625e-07f;
    float lo[3] = {bmin[0] - m, bmin[1] - m, bmin[2] - m}, hi[3] = {bmax[0] + m, bmax[1] + m, bmax[2] + m};
    return or_ray_aabb(o, inv, lo, hi, tmax);
}
static inline int exact(uint32_t variant) { return variant != OR_WALK_REF; }
static inline float box_test(const float *o, const float *inv, const float *bmin, const float *bmax, float tmax,
                             uint32_t variant) {
    return exact(variant) ? ray_aabb_wide(o, inv, bmin, bmax, tmax) : or_ray_aabb(o, inv, bmin, bmax, tmax);
}

void or_intersect_tri(const float *o, const float *d, const or_leaf *l, float tmax, float *tuv) {
    const float *v0 = l->r[0], *e1 = l->r[1], *e2 = l->r[2];
    tuv[0] = OR_INF;
    tuv[1] = tuv[2] = 0.0f;
    float pv[3];
    cross3(d, e2, pv);
    float det = dot3(e1, pv);
    if (fabsf(det) < 1e-12f) return;
    float inv = 1.0f / det;
    float tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    float u = dot3(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f || isnan(u)) return;
    float qv[3];
    cross3(tv, e1, qv);
    float v = dot3(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return;
    float t = dot3(e2, qv) * inv;
    if (t <= 0.0f || t >= tmax) return;
    tuv[0] = t;
    tuv[1] = u;
    tuv[2] = v;
}

/* The watertight test (Woop, Benthin and Wald, JCGT 2013): the vertices v0, v0 + e1, v0 + e2 translated to the
 * ray's origin and sheared so the ray runs along +z (kz its largest component, the first of equals; kx, ky the next
 * two, swapped when d[kz] < 0), the edge functions U, V, W of their projections (again in double where one rounds to
 * exactly 0), the same signs or a miss, t = T / det; u = V / det and v = W / det weigh v1 and v2. */
void or_intersect_tri_wt(const float *o, const float *d, const or_leaf *l, float tmax, float *tuv) {
    tuv[0] = OR_INF;
    tuv[1] = tuv[2] = 0.0f;
    float ax_ = fabsf(d[0]), ay_ = fabsf(d[1]), az_ = fabsf(d[2]);
    int kz = ax_ >= ay_ ? (ax_ >= az_ ? 0 : 2) : (ay_ >= az_ ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    if (d[kz] < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    float sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1.0f / d[kz];
    const float *v0 = l->r[0], *e1 = l->r[1], *e2 = l->r[2];
    float v1[3] = {v0[0] + e1[0], v0[1] + e1[1], v0[2] + e1[2]};
    float v2[3] = {v0[0] + e2[0], v0[1] + e2[1], v0[2] + e2[2]};
    float A[3], B[3], C[3];
    for (int k = 0; k < 3; k++) {
        A[k] = v0[k] - o[k];
        B[k] = v1[k] - o[k];
        C[k] = v2[k] - o[k];
    }
    float ax = A[kx] - sx * A[kz], ay = A[ky] - sy * A[kz];
    float bx = B[kx] - sx * B[kz], by = B[ky] - sy * B[kz];
    float cx = C[kx] - sx * C[kz], cy = C[ky] - sy * C[kz];
    float U = cx * by - cy * bx, V = ax * cy - ay * cx, W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)cx * by - (double)cy * bx);
        V = (float)((double)ax * cy - (double)ay * cx);
        W = (float)((double)bx * ay - (double)by * ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return;
    float det = U + V + W;
    if (det == 0.0f) return;
    float T = U * (sz * A[kz]) + V * (sz * B[kz]) + W * (sz * C[kz]);
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t > 0.0f) || t >= tmax) return;
    tuv[0] = t;
    tuv[1] = V * rcp;
    tuv[2] = W * rcp;
}

static inline void tri_test(uint32_t variant, const float *o, const float *d, const or_leaf *l, float tmax,
                            float *tuv) {
    if (variant == OR_WALK_WT)
        or_intersect_tri_wt(o, d, l, tmax, tuv);
    else
        or_intersect_tri(o, d, l, tmax, tuv);
}

void or_inv(const float *d, float *inv) {
    for (int k = 0; k < 3; k++) inv[k] = 1.0f / (fabsf(d[k]) < 1e-20f ? 1e-20f : d[k]);
}

static inline void proc_dir(const or_instance *r, const float *d, float *out) {
    for (int k = 0; k < 3; k++) out[k] = r->n[0][k] * d[0] + r->n[1][k] * d[1] + r->n[2][k] * d[2];
}

/* The ray placed in a unit's space, as the walk places it (trace_units.wgsl:92-97). */
static inline void unit_ray(const or_instance *u, const float *o, const float *d, float *ro, float *rd, float *rinv,
                            const float *inv_d) {
    float t[3] = {o[0] - u->m[0][3], o[1] - u->m[1][3], o[2] - u->m[2][3]};
    if ((fbits(u->n[1][3]) & OR_UNIT_SHIFTED) == 0) {
        proc_dir(u, t, ro);
        proc_dir(u, d, rd);
        or_inv(rd, rinv);
    } else {
        memcpy(ro, t, 12);
        memcpy(rd, d, 12);
        memcpy(rinv, inv_d, 12);
    }
}

static inline float clamp_tmax(float t) { return t < OR_INF ? t : OR_INF; }
static inline float cull_t(float t, uint32_t variant) { return exact(variant) ? t + t * OR_CULL_SLACK : t; }

/* Whether a candidate (t, inst, tri), already inside (0, tmax), replaces the hit. */
static inline int accept(const or_hit *h, float x, uint32_t inst, uint32_t tri, uint32_t variant) {
    if (x < h->t) return 1;
    if (exact(variant) && x == h->t && h->tri != OR_NO_TRI)
        return inst < h->inst || (inst == h->inst && tri < h->tri);
    return 0;
}

int or_trace(const or_scene *s, const float *o, const float *d, float tmax_in, uint32_t mode, uint32_t variant,
             uint64_t budget, or_hit *hit, or_reads *reads) {
    or_reads rd_local;
    if (!reads) reads = &rd_local;
    memset(reads, 0, sizeof(*reads));
    float tmax = clamp_tmax(tmax_in);
    or_hit h = {tmax, OR_NO_TRI, 0.0f, 0.0f, OR_NO_INST, 0};
    *hit = h;
    if (budget == 0) budget = 64 * (s->node_count + s->leaf_count * 3) + 1024;
    if (s->units == 0) return 0;
    float inv_d[3];
    or_inv(d, inv_d);
    uint32_t top = s->top;
#define NODE(i, dst)                                        \
    do {                                                    \
        if ((uint64_t)(i) >= s->node_count) return -1;      \
        if ((i) >= top) reads->top_nodes++;                 \
        else reads->mesh_nodes++;                           \
        dst = s->nodes[(i)];                                \
    } while (0)
    or_node whole;
    NODE(top, whole);
    if (box_test(o, inv_d, whole.bmin, whole.bmax, cull_t(h.t, variant), variant) == OR_INF) return 0;
    uint32_t stack[OR_STACK];
    uint32_t sp = 0, node = top, rec = OR_NO_INST, off = top, first = 0, base = 0;
    float ro[3], rdir[3], rinv[3];
    memcpy(ro, o, 12);
    memcpy(rdir, d, 12);
    memcpy(rinv, inv_d, 12);
    for (;;) {
        if (++reads->steps > budget) return -1;
        or_node n;
        NODE(node, n);
        if (n.count == 0) {
            uint32_t l = off + n.left_or_first, r = l + 1;
            or_node nl, nr;
            NODE(l, nl);
            NODE(r, nr);
            float tl = box_test(ro, rinv, nl.bmin, nl.bmax, cull_t(h.t, variant), variant);
            float tr = box_test(ro, rinv, nr.bmin, nr.bmax, cull_t(h.t, variant), variant);
            if (tl != OR_INF || tr != OR_INF) {
                uint32_t near = l, far = r;
                if (tr < tl) {
                    near = r;
                    far = l;
                    float tmp = tl;
                    tl = tr;
                    tr = tmp;
                }
                node = near;
                if (tr != OR_INF && sp < OR_STACK) {
                    stack[sp++] = far;
                    if (sp > reads->deepest) reads->deepest = sp;
                }
                continue;
            }
        } else if (rec == OR_NO_INST) {
            if (n.left_or_first >= s->instance_count) return -1;
            reads->records++;
            const or_instance *u = &s->instances[n.left_or_first];
            if (mode != OR_TRACE_SHADOW || (fbits(u->n[1][3]) & OR_UNIT_SHADOW) != 0) {
                rec = n.left_or_first;
                unit_ray(u, o, d, ro, rdir, rinv, inv_d);
                off = fbits(u->n[0][3]);
                first = u->info[0];
                base = sp;
                node = off;
                continue;
            }
        } else {
            for (uint32_t k = 0; k < n.count; k++) {
                if (++reads->steps > budget) return -1;
                uint32_t ti = first + n.left_or_first + k;
                if ((uint64_t)ti >= s->leaf_count) return -1;
                reads->mesh_verts += 3;
                float x[3];
                tri_test(variant, ro, rdir, &s->leaves[ti], exact(variant) ? tmax : h.t, x);
                if (x[0] < tmax && accept(&h, x[0], rec, ti, variant)) {
                    h.t = x[0];
                    h.tri = ti;
                    h.u = x[1];
                    h.v = x[2];
                    h.inst = rec;
                    h.part = fbits(s->leaves[ti].r[0][3]);
                    if (mode != OR_TRACE_CLOSEST) {
                        *hit = h;
                        return 0;
                    }
                }
            }
        }
        if (rec != OR_NO_INST && sp == base) {
            rec = OR_NO_INST;
            memcpy(ro, o, 12);
            memcpy(rdir, d, 12);
            memcpy(rinv, inv_d, 12);
            off = top;
        }
        if (sp == 0) break;
        sp--;
        node = stack[sp];
    }
#undef NODE
    *hit = h;
    return 0;
}

/* The triangles of a mesh tree from `root` (its leaves' ranges, relative to the unit's first), ascending. */
typedef struct {
    uint32_t first, count;
} Range;
static int cmp_range(const void *a, const void *b) {
    uint32_t x = ((const Range *)a)->first, y = ((const Range *)b)->first;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int or_brute(const or_scene *s, const float *o, const float *d, float tmax_in, uint32_t mode, uint32_t variant,
             or_hit *hit) {
    float tmax = clamp_tmax(tmax_in);
    or_hit h = {tmax, OR_NO_TRI, 0.0f, 0.0f, OR_NO_INST, 0};
    *hit = h;
    if (s->units == 0) return 0;
    /* The units: the top tree's leaves, by record (units.rs:1372 iterates the units by record). */
    uint64_t cap = s->node_count + 1;
    uint32_t *recs = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    uint32_t *st = (uint32_t *)malloc(sizeof(uint32_t) * (cap + 2));
    uint64_t nrec = 0, sp = 0;
    st[sp++] = s->top;
    while (sp) {
        uint32_t i = st[--sp];
        if (i >= s->node_count) goto fail;
        const or_node *n = &s->nodes[i];
        if (n->count > 0) {
            recs[nrec++] = n->left_or_first;
        } else {
            if (sp + 2 > cap + 2) goto fail;
            st[sp++] = s->top + n->left_or_first;
            st[sp++] = s->top + n->left_or_first + 1;
        }
    }
    /* ascending records */
    for (uint64_t a = 1; a < nrec; a++) {
        uint32_t x = recs[a];
        uint64_t b = a;
        while (b > 0 && recs[b - 1] > x) {
            recs[b] = recs[b - 1];
            b--;
        }
        recs[b] = x;
    }
    float inv_d[3];
    or_inv(d, inv_d);
    Range *ranges = (Range *)malloc(sizeof(Range) * cap);
    for (uint64_t q = 0; q < nrec; q++) {
        uint32_t rec = recs[q];
        if (rec >= s->instance_count) {
            free(ranges);
            goto fail;
        }
        const or_instance *u = &s->instances[rec];
        if (mode == OR_TRACE_SHADOW && (fbits(u->n[1][3]) & OR_UNIT_SHADOW) == 0) continue;
        float ro[3], rd[3], rinv[3];
        unit_ray(u, o, d, ro, rd, rinv, inv_d);
        uint32_t root = fbits(u->n[0][3]), first = u->info[0];
        uint64_t nr = 0;
        sp = 0;
        st[sp++] = root;
        while (sp) {
            uint32_t i = st[--sp];
            if (i >= s->node_count) {
                free(ranges);
                goto fail;
            }
            const or_node *n = &s->nodes[i];
            if (n->count > 0) {
                ranges[nr++] = (Range){n->left_or_first, n->count};
            } else {
                st[sp++] = root + n->left_or_first;
                st[sp++] = root + n->left_or_first + 1;
            }
            if (nr >= cap || sp + 2 > cap) {
                free(ranges);
                goto fail;
            }
        }
        qsort(ranges, nr, sizeof(Range), cmp_range);
        for (uint64_t g = 0; g < nr; g++) {
            for (uint32_t k = 0; k < ranges[g].count; k++) {
                uint32_t ti = first + ranges[g].first + k;
                if ((uint64_t)ti >= s->leaf_count) {
                    free(ranges);
                    goto fail;
                }
                float x[3];
                tri_test(variant, ro, rd, &s->leaves[ti], exact(variant) ? tmax : h.t, x);
                if (x[0] < tmax && accept(&h, x[0], rec, ti, variant)) {
                    h.t = x[0];
                    h.tri = ti;
                    h.u = x[1];
                    h.v = x[2];
                    h.inst = rec;
                    h.part = fbits(s->leaves[ti].r[0][3]);
                }
            }
        }
    }
    free(ranges);
    free(recs);
    free(st);
    *hit = h;
    return 0;
fail:
    free(recs);
    free(st);
    return -1;
}

int64_t or_trace_batch(const or_scene *s, const or_ray *rays, uint64_t n, uint32_t variant, int threads,
                       or_hit *hits, or_reads *total) {
    int64_t failed = 0;
    uint64_t tn = 0, mn = 0, mv = 0, rc = 0, st = 0, deep = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256) reduction(+ : failed, tn, mn, mv, rc, st) \
    reduction(max : deep)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++) {
        or_reads r;
        if (or_trace(s, rays[i].o, rays[i].d, rays[i].tmax, rays[i].mode, variant, 0, &hits[i], &r)) failed++;
        tn += r.top_nodes;
        mn += r.mesh_nodes;
        mv += r.mesh_verts;
        rc += r.records;
        st += r.steps;
        if (r.deepest > deep) deep = r.deepest;
    }
    (void)threads;
    if (total) {
        total->top_nodes = tn;
        total->mesh_nodes = mn;
        total->mesh_verts = mv;
        total->records = rc;
        total->steps = st;
        total->deepest = deep;
    }
    return failed;
}

int64_t or_brute_batch(const or_scene *s, const or_ray *rays, uint64_t n, uint32_t variant, int threads,
                       or_hit *hits) {
    int64_t failed = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16) reduction(+ : failed)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++) {
        if (or_brute(s, rays[i].o, rays[i].d, rays[i].tmax, rays[i].mode, variant, &hits[i])) failed++;
    }
    (void)threads;
    return failed;
}

/* ---- the FP64 brute force (or_brute64_batch) ------------------------------------------------------------------ */
static int brute64(const or_scene *s, const float *of, const float *df, float tmax_in, or_hit64 *hit) {
    or_hit64 h = {-1.0, 0.0, 0.0, OR_NO_TRI, OR_NO_INST};
    double tmax = tmax_in < OR_INF ? (double)tmax_in : (double)OR_INF;
    /* every record the top tree names */
    uint32_t cap = 64 + 2 * (uint32_t)s->units;
    uint32_t *st = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    uint32_t sp = 0;
    if (s->units) st[sp++] = s->top;
    while (sp) {
        uint32_t i = st[--sp];
        if (i >= s->node_count) {
            free(st);
            return -1;
        }
        const or_node *n = &s->nodes[i];
        if (n->count == 0) {
            if (sp + 2 > cap) {
                free(st);
                return -1;
            }
            st[sp++] = s->top + n->left_or_first;
            st[sp++] = s->top + n->left_or_first + 1;
            continue;
        }
        uint32_t rec = n->left_or_first;
        if (rec >= s->instance_count) {
            free(st);
            return -1;
        }
        const or_instance *u = &s->instances[rec];
        double o[3], d[3];
        double t3[3] = {(double)of[0] - u->m[0][3], (double)of[1] - u->m[1][3], (double)of[2] - u->m[2][3]};
        if ((fbits(u->n[1][3]) & OR_UNIT_SHIFTED) == 0) {
            for (int k = 0; k < 3; k++) {
                o[k] = (double)u->n[0][k] * t3[0] + (double)u->n[1][k] * t3[1] + (double)u->n[2][k] * t3[2];
                d[k] = (double)u->n[0][k] * df[0] + (double)u->n[1][k] * df[1] + (double)u->n[2][k] * df[2];
            }
        } else {
            for (int k = 0; k < 3; k++) o[k] = t3[k], d[k] = df[k];
        }
        /* the unit's triangles: its mesh tree's leaves */
        uint32_t root = fbits(u->n[0][3]), first = u->info[0];
        uint32_t ms[256]; /* a mesh tree is at most OR_MESH_DEPTH deep: at most one pending child a level */
        uint64_t msp = 0;
        ms[msp++] = root;
        while (msp) {
            uint32_t j = ms[--msp];
            if (j >= s->node_count) {
                free(st);
                return -1;
            }
            const or_node *m = &s->nodes[j];
            if (m->count == 0) {
                if (msp + 2 > 256) {
                    free(st);
                    return -1;
                }
                ms[msp++] = root + m->left_or_first;
                ms[msp++] = root + m->left_or_first + 1;
                continue;
            }
            for (uint32_t k = 0; k < m->count; k++) {
                uint32_t ti = first + m->left_or_first + k;
                if ((uint64_t)ti >= s->leaf_count) {
                    free(st);
                    return -1;
                }
                const or_leaf *l = &s->leaves[ti];
                double v0[3] = {l->r[0][0], l->r[0][1], l->r[0][2]}, e1[3] = {l->r[1][0], l->r[1][1], l->r[1][2]},
                       e2[3] = {l->r[2][0], l->r[2][1], l->r[2][2]};
                double pv[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0], d[0] * e2[1] - e2[0] * d[1]};
                double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
                if (det == 0.0) continue;
                double inv = 1.0 / det;
                double tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
                double uu = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
                if (!(uu >= 0.0 && uu <= 1.0)) continue;
                double qv[3] = {tv[1] * e1[2] - e1[1] * tv[2], tv[2] * e1[0] - e1[2] * tv[0],
                                tv[0] * e1[1] - e1[0] * tv[1]};
                double vv = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
                if (vv < 0.0 || uu + vv > 1.0) continue;
                double tt = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
                if (!(tt > 0.0) || tt >= tmax) continue;
                int better = h.tri == OR_NO_TRI || tt < h.t ||
                             (tt == h.t && (rec < h.inst || (rec == h.inst && ti < h.tri)));
                if (better) {
                    h.t = tt, h.u = uu, h.v = vv, h.tri = ti, h.inst = rec;
                }
            }
        }
    }
    free(st);
    *hit = h;
    return 0;
}

int64_t or_brute64_batch(const or_scene *s, const or_ray *rays, uint64_t n, int threads, or_hit64 *hits) {
    int64_t failed = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : failed)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++)
        if (brute64(s, rays[i].o, rays[i].d, rays[i].tmax, &hits[i])) failed++;
    (void)threads;
    return failed;
}
