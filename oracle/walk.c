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
static inline float box_test(const float *o, const float *inv, const float *bmin, const float *bmax, float tmax,
                             uint32_t variant) {
    return variant == OR_WALK_EXACT ? ray_aabb_wide(o, inv, bmin, bmax, tmax) : or_ray_aabb(o, inv, bmin, bmax, tmax);
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
static inline float cull_t(float t, uint32_t variant) { return variant == OR_WALK_EXACT ? t + t * OR_CULL_SLACK : t; }

/* Whether a candidate (t, inst, tri), already inside (0, tmax), replaces the hit. */
static inline int accept(const or_hit *h, float x, uint32_t inst, uint32_t tri, uint32_t variant) {
    if (x < h->t) return 1;
    if (variant == OR_WALK_EXACT && x == h->t && h->tri != OR_NO_TRI)
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
                or_intersect_tri(ro, rdir, &s->leaves[ti], variant == OR_WALK_EXACT ? tmax : h.t, x);
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
                or_intersect_tri(ro, rd, &s->leaves[ti], variant == OR_WALK_EXACT ? tmax : h.t, x);
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
