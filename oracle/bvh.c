/*
 * Oracle (test infrastructure only): the reference's tree builders restated in C.
 *   bvh.rs:17-37   Aabb (EMPTY, grow, merge, area)
 *   bvh.rs:81-106  build_tris
 *   bvh.rs:118-121 levels
 *   bvh.rs:130-239 build_from (binned SAH, 16 bins, leaves <= 4 or <= 16 by SAH, depth cap, median fallback)
 *   bvh.rs:246-282 check_tree
 *   units.rs:57-79 check_mesh / check_top, 84-118 build_median, 127-223 build_sah, 225-246 tri_boxes / mesh_tree,
 *   units.rs:288-297 UnitBox::new, 303-321 top_tree, 329-342 unit_record / translation_alone, 365-367 top_cost.
 * glam Vec3A (SSE2) lane semantics: min(a, b) = a < b ? a : b, max(a, b) = a > b ? a : b.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    float v[3];
} V3;

typedef struct {
    V3 min, max;
} Aabb;

static inline float smin(float a, float b) { return a < b ? a : b; } /* _mm_min_ps lane */
static inline float smax(float a, float b) { return a > b ? a : b; } /* _mm_max_ps lane */

static inline V3 vmin(V3 a, V3 b) {
    V3 r;
    for (int k = 0; k < 3; k++) r.v[k] = smin(a.v[k], b.v[k]);
    return r;
}
static inline V3 vmax(V3 a, V3 b) {
    V3 r;
    for (int k = 0; k < 3; k++) r.v[k] = smax(a.v[k], b.v[k]);
    return r;
}
static inline int vfinite(V3 a) { return isfinite(a.v[0]) && isfinite(a.v[1]) && isfinite(a.v[2]); }

static const Aabb AABB_EMPTY = {{{INFINITY, INFINITY, INFINITY}}, {{-INFINITY, -INFINITY, -INFINITY}}};

static inline void grow(Aabb *b, V3 p) {
    b->min = vmin(b->min, p);
    b->max = vmax(b->max, p);
}
static inline void merge(Aabb *b, const Aabb *o) {
    b->min = vmin(b->min, o->min);
    b->max = vmax(b->max, o->max);
}
/* bvh.rs:33-36 (and units.rs:129-132 `area`) */
static inline float area_of(V3 lo, V3 hi) {
    float e[3];
    for (int k = 0; k < 3; k++) e[k] = smax(hi.v[k] - lo.v[k], 0.0f);
    return 2.0f * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
}
static inline int aabb_finite(const Aabb *b) {
    return vfinite(b->min) && vfinite(b->max) && b->min.v[0] <= b->max.v[0] && b->min.v[1] <= b->max.v[1] &&
           b->min.v[2] <= b->max.v[2];
}

/* Rust `f as usize`: saturating, NaN -> 0. */
static inline uint64_t as_usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)f;
}

/* f32::total_cmp key. */
static inline int32_t total_key(float f) {
    int32_t b;
    memcpy(&b, &f, 4);
    b ^= (int32_t)(((uint32_t)(b >> 31)) >> 1);
    return b;
}

static inline uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
/* bvh.rs:119-121 */
static inline uint64_t levels(uint64_t count) {
    uint64_t p = next_pow2((count + 3) / 4);
    uint64_t tz = 0;
    while (!(p & 1)) {
        p >>= 1;
        tz++;
    }
    return tz;
}

static const or_node NODE_ZERO = {{0, 0, 0}, 0, {0, 0, 0}, 0};
static const or_node NODE_EMPTY = {{0, 0, 0}, 0, {0, 0, 0}, 1}; /* bvh.rs:63, units.rs:43 FILL */

typedef struct {
    or_node *v;
    uint64_t n, cap;
} NodeVec;

static int nv_push(NodeVec *nv, or_node x) {
    if (nv->n == nv->cap) {
        uint64_t c = nv->cap ? nv->cap * 2 : 64;
        or_node *p = (or_node *)realloc(nv->v, c * sizeof(or_node));
        if (!p) return -1;
        nv->v = p;
        nv->cap = c;
    }
    nv->v[nv->n++] = x;
    return 0;
}

static or_node mk(V3 lo, uint32_t lf, V3 hi, uint32_t count) {
    or_node n;
    memcpy(n.bmin, lo.v, 12);
    memcpy(n.bmax, hi.v, 12);
    n.left_or_first = lf;
    n.count = count;
    return n;
}

typedef struct {
    uint64_t ni, start, end, depth;
} Job;

/* Sort helper: items by (key on an axis, index); the oracle's stand-in for select_nth_unstable_by. */
static const V3 *g_keys;
static int g_axis;
static int by_key_then_index(const void *pa, const void *pb) {
    uint32_t a = *(const uint32_t *)pa, b = *(const uint32_t *)pb;
    int32_t ka = total_key(g_keys[a].v[g_axis]), kb = total_key(g_keys[b].v[g_axis]);
    if (ka != kb) return ka < kb ? -1 : 1;
    return a < b ? -1 : (a > b ? 1 : 0);
}
static void sort_by_axis(uint32_t *items, uint64_t n, const V3 *keys, int axis) {
    g_keys = keys;
    g_axis = axis;
    qsort(items, n, sizeof(uint32_t), by_key_then_index);
}

void or_tree_free(or_tree *t) {
    if (!t) return;
    free(t->nodes);
    free(t->order);
    t->nodes = NULL;
    t->order = NULL;
    t->node_count = t->order_count = 0;
}

/* bvh.rs:130-239 build_from */
static int build_from(const Aabb *boxes, const V3 *centroids, uint64_t total, uint64_t max_depth, or_tree *out) {
    enum { BINS = 16, MAX_LEAF = 4, SAH_LEAF = 16 };
    uint32_t *order = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
    uint64_t n = 0;
    for (uint64_t i = 0; i < total; i++)
        if (aabb_finite(&boxes[i])) order[n++] = (uint32_t)i;
    uint64_t m = n;
    for (uint64_t i = 0; i < total; i++)
        if (!aabb_finite(&boxes[i])) order[m++] = (uint32_t)i;
    NodeVec nv = {0};
    nv_push(&nv, NODE_ZERO);
    out->order = order;
    out->order_count = total;
    out->depth = 0;
    if (n == 0) {
        out->nodes = nv.v;
        out->node_count = nv.n;
        return 0;
    }
    if (levels(n) > max_depth) {
        free(nv.v);
        free(order);
        out->order = NULL;
        return -1;
    }
    uint64_t deepest = 0;
    Job *stack = (Job *)malloc(sizeof(Job) * (2 * max_depth + 8));
    uint64_t sp = 0;
    stack[sp++] = (Job){0, 0, n, 0};
    while (sp) {
        Job j = stack[--sp];
        if (j.depth > deepest) deepest = j.depth;
        Aabb bounds = AABB_EMPTY, cb = AABB_EMPTY;
        for (uint64_t q = j.start; q < j.end; q++) {
            merge(&bounds, &boxes[order[q]]);
            grow(&cb, centroids[order[q]]);
        }
        uint64_t count = j.end - j.start;
        if (count <= MAX_LEAF) {
            nv.v[j.ni] = mk(bounds.min, (uint32_t)j.start, bounds.max, (uint32_t)count);
            continue;
        }
        V3 ext;
        for (int k = 0; k < 3; k++) ext.v[k] = cb.max.v[k] - cb.min.v[k];
        float best_cost = INFINITY;
        int best_axis = 0;
        uint64_t best_split = 0;
        for (int axis = 0; axis < 3; axis++) {
            if (ext.v[axis] <= 1e-9f) continue;
            Aabb bin_box[BINS];
            uint64_t bin_cnt[BINS];
            for (int i = 0; i < BINS; i++) {
                bin_box[i] = AABB_EMPTY;
                bin_cnt[i] = 0;
            }
            float scale = (float)BINS / ext.v[axis];
            for (uint64_t q = j.start; q < j.end; q++) {
                uint32_t t = order[q];
                uint64_t b = as_usize((centroids[t].v[axis] - cb.min.v[axis]) * scale);
                if (b > BINS - 1) b = BINS - 1;
                bin_cnt[b]++;
                merge(&bin_box[b], &boxes[t]);
            }
            float left_area[BINS];
            uint64_t left_cnt[BINS];
            Aabb acc = AABB_EMPTY;
            uint64_t c = 0;
            for (int i = 0; i < BINS; i++) {
                merge(&acc, &bin_box[i]);
                c += bin_cnt[i];
                left_area[i] = area_of(acc.min, acc.max);
                left_cnt[i] = c;
            }
            acc = AABB_EMPTY;
            c = 0;
            for (int i = BINS - 1; i >= 1; i--) {
                merge(&acc, &bin_box[i]);
                c += bin_cnt[i];
                float cost = left_area[i - 1] * (float)left_cnt[i - 1] + area_of(acc.min, acc.max) * (float)c;
                uint64_t larger = left_cnt[i - 1] > c ? left_cnt[i - 1] : c;
                int fits = j.depth + 1 + levels(larger) <= max_depth;
                if (left_cnt[i - 1] > 0 && c > 0 && cost < best_cost && fits) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = (uint64_t)i;
                }
            }
        }
        float leaf_cost = area_of(bounds.min, bounds.max) * (float)count;
        int splittable = isfinite(best_cost);
        if (count <= SAH_LEAF && (!splittable || best_cost >= leaf_cost)) {
            nv.v[j.ni] = mk(bounds.min, (uint32_t)j.start, bounds.max, (uint32_t)count);
            continue;
        }
        uint64_t mid;
        if (splittable) {
            float scale = (float)BINS / ext.v[best_axis];
            uint32_t *slice = order + j.start;
            uint64_t i = 0;
            for (uint64_t q = 0; q < count; q++) {
                uint64_t b = as_usize((centroids[slice[q]].v[best_axis] - cb.min.v[best_axis]) * scale);
                if (b > BINS - 1) b = BINS - 1;
                if (b < best_split) {
                    uint32_t tmp = slice[i];
                    slice[i] = slice[q];
                    slice[q] = tmp;
                    i++;
                }
            }
            mid = j.start + i;
        } else if (ext.v[0] <= 1e-9f && ext.v[1] <= 1e-9f && ext.v[2] <= 1e-9f) {
            mid = j.start + count / 2;
        } else {
            /* Iterator::max_by with total_cmp: the last of equal maxima. */
            int axis = 0;
            for (int k = 1; k < 3; k++)
                if (!(total_key(ext.v[axis]) > total_key(ext.v[k]))) axis = k;
            uint64_t half = count / 2;
            sort_by_axis(order + j.start, count, centroids, axis);
            mid = j.start + half;
        }
        uint64_t left = nv.n;
        nv_push(&nv, NODE_ZERO);
        nv_push(&nv, NODE_ZERO);
        nv.v[j.ni] = mk(bounds.min, (uint32_t)left, bounds.max, 0);
        stack[sp++] = (Job){left + 1, mid, j.end, j.depth + 1};
        stack[sp++] = (Job){left, j.start, mid, j.depth + 1};
    }
    free(stack);
    out->nodes = nv.v;
    out->node_count = nv.n;
    out->depth = (uint32_t)deepest;
    return 0;
}

/* bvh.rs:81-106 */
int or_build_tris(const float *tris, uint64_t n, uint32_t max_depth, or_tree *out) {
    memset(out, 0, sizeof(*out));
    if (n == 0) {
        out->nodes = (or_node *)malloc(sizeof(or_node));
        out->nodes[0] = NODE_EMPTY;
        out->node_count = 1;
        out->order = (uint32_t *)malloc(sizeof(uint32_t));
        out->order_count = 0;
        return 0;
    }
    Aabb *boxes = (Aabb *)malloc(n * sizeof(Aabb));
    V3 *cent = (V3 *)malloc(n * sizeof(V3));
    for (uint64_t i = 0; i < n; i++) {
        const float *t = tris + 9 * i;
        V3 p[3];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) p[c].v[k] = t[3 * c + k];
        Aabb b = AABB_EMPTY;
        if (vfinite(p[0]) && vfinite(p[1]) && vfinite(p[2]))
            for (int c = 0; c < 3; c++) grow(&b, p[c]);
        boxes[i] = b;
        for (int k = 0; k < 3; k++) cent[i].v[k] = (p[0].v[k] + p[1].v[k] + p[2].v[k]) / 3.0f;
    }
    int rc = build_from(boxes, cent, n, max_depth, out);
    free(boxes);
    free(cent);
    if (rc) return rc;
    if (out->node_count == 1 && out->nodes[0].count == 0) out->nodes[0] = NODE_EMPTY;
    return 0;
}

/* bvh.rs:246-282 check_tree; the leaf rule is `count <= max_leaf && first + count <= items`. */
int or_check_tree(const or_node *nodes, uint64_t count, uint32_t max_depth, uint64_t items, uint32_t max_leaf,
                  char *msg, uint64_t msg_len) {
    if (count == 0) {
        snprintf(msg, msg_len, "a tree without nodes");
        return -1;
    }
    uint8_t *seen = (uint8_t *)calloc(count, 1);
    uint64_t *st = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (count + 2));
    uint64_t sp = 0, deepest = 0, reached = 0;
    st[sp++] = 0;
    st[sp++] = 0;
    int rc = 0;
    while (sp) {
        uint64_t depth = st[--sp];
        uint64_t i = st[--sp];
        if (seen[i]) {
            snprintf(msg, msg_len, "node %llu is reached twice", (unsigned long long)i);
            rc = -1;
            break;
        }
        seen[i] = 1;
        reached++;
        if (depth > deepest) deepest = depth;
        const or_node *n = &nodes[i];
        if (n->count > 0) {
            if (!(n->count <= max_leaf && (uint64_t)n->left_or_first + n->count <= items)) {
                snprintf(msg, msg_len, "leaf %llu: items %u..%llu of %llu", (unsigned long long)i, n->left_or_first,
                         (unsigned long long)n->left_or_first + n->count, (unsigned long long)items);
                rc = -1;
                break;
            }
            continue;
        }
        int fin = 1;
        for (int k = 0; k < 3; k++)
            fin &= isfinite(n->bmin[k]) && isfinite(n->bmax[k]) && n->bmin[k] <= n->bmax[k];
        if (!fin) {
            snprintf(msg, msg_len, "node %llu's box is not finite", (unsigned long long)i);
            rc = -1;
            break;
        }
        uint64_t l = n->left_or_first;
        if (l <= i || l + 1 >= count) {
            snprintf(msg, msg_len, "node %llu's children are not after it and inside the tree",
                     (unsigned long long)i);
            rc = -1;
            break;
        }
        if (depth + 1 > max_depth) {
            snprintf(msg, msg_len, "deeper than %u", max_depth);
            rc = -1;
            break;
        }
        st[sp++] = l;
        st[sp++] = depth + 1;
        st[sp++] = l + 1;
        st[sp++] = depth + 1;
    }
    if (rc == 0 && reached != count) {
        snprintf(msg, msg_len, "%llu of %llu nodes are not reached", (unsigned long long)(count - reached),
                 (unsigned long long)count);
        rc = -1;
    }
    free(seen);
    free(st);
    return rc ? -1 : (int)deepest;
}

/* units.rs:68-79 check_top */
int or_check_top(const or_node *nodes, uint64_t count, uint32_t records, char *msg, uint64_t msg_len) {
    int depth = or_check_tree(nodes, count, OR_TOP_DEPTH, UINT64_MAX, 1, msg, msg_len);
    if (depth < 0) return -1;
    uint8_t *seen = (uint8_t *)calloc(records ? records : 1, 1);
    for (uint64_t i = 0; i < count; i++) {
        if (nodes[i].count == 0) continue;
        uint32_t r = nodes[i].left_or_first;
        if (nodes[i].count != 1 || r >= records) {
            snprintf(msg, msg_len, "record %u (count %u) is not one unit", r, nodes[i].count);
            free(seen);
            return -1;
        }
        if (seen[r]) {
            snprintf(msg, msg_len, "record %u twice", r);
            free(seen);
            return -1;
        }
        seen[r] = 1;
    }
    free(seen);
    return depth;
}

static void read_boxes(const float *boxes, uint64_t n, V3 *lo, V3 *hi, V3 *centre) {
    for (uint64_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            lo[i].v[k] = boxes[6 * i + k];
            hi[i].v[k] = boxes[6 * i + 3 + k];
            centre[i].v[k] = (lo[i].v[k] + hi[i].v[k]) * 0.5f;
        }
    }
}

/* units.rs:84-118 build_median */
int or_build_median(const float *boxes, uint64_t n, uint32_t max_leaf, or_tree *out) {
    memset(out, 0, sizeof(*out));
    out->order = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    out->order_count = n;
    for (uint64_t i = 0; i < n; i++) out->order[i] = (uint32_t)i;
    NodeVec nv = {0};
    nv_push(&nv, NODE_EMPTY);
    if (n == 0) {
        out->nodes = nv.v;
        out->node_count = nv.n;
        return 0;
    }
    V3 *lo = (V3 *)malloc(n * sizeof(V3)), *hi = (V3 *)malloc(n * sizeof(V3)), *ce = (V3 *)malloc(n * sizeof(V3));
    read_boxes(boxes, n, lo, hi, ce);
    uint32_t *order = out->order;
    Job *stack = (Job *)malloc(sizeof(Job) * 160);
    uint64_t sp = 0;
    stack[sp++] = (Job){0, 0, n, 0};
    while (sp) {
        Job j = stack[--sp];
        V3 l = {{INFINITY, INFINITY, INFINITY}}, h = {{-INFINITY, -INFINITY, -INFINITY}}, cl = l, ch = h;
        for (uint64_t q = j.start; q < j.end; q++) {
            uint32_t k = order[q];
            l = vmin(l, lo[k]);
            h = vmax(h, hi[k]);
            cl = vmin(cl, ce[k]);
            ch = vmax(ch, ce[k]);
        }
        uint64_t count = j.end - j.start;
        if (count <= max_leaf) {
            nv.v[j.ni] = mk(l, (uint32_t)j.start, h, (uint32_t)count);
            continue;
        }
        float ex = ch.v[0] - cl.v[0], ey = ch.v[1] - cl.v[1], ez = ch.v[2] - cl.v[2];
        int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
        uint64_t mid = count / 2;
        sort_by_axis(order + j.start, count, ce, axis);
        uint64_t left = nv.n;
        nv_push(&nv, NODE_EMPTY);
        nv_push(&nv, NODE_EMPTY);
        nv.v[j.ni] = mk(l, (uint32_t)left, h, 0);
        stack[sp++] = (Job){left + 1, j.start + mid, j.end, 0};
        stack[sp++] = (Job){left, j.start, j.start + mid, 0};
    }
    free(stack);
    free(lo);
    free(hi);
    free(ce);
    out->nodes = nv.v;
    out->node_count = nv.n;
    return 0;
}

/* units.rs:140-223 build_sah */
int or_build_sah(const float *boxes, uint64_t n, uint32_t max_depth, or_tree *out) {
    enum { TOP_BINS = 32 };
    memset(out, 0, sizeof(*out));
    out->order = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    out->order_count = n;
    for (uint64_t i = 0; i < n; i++) out->order[i] = (uint32_t)i;
    NodeVec nv = {0};
    nv_push(&nv, NODE_EMPTY);
    if (n == 0) {
        out->nodes = nv.v;
        out->node_count = nv.n;
        return 0;
    }
    V3 *lo = (V3 *)malloc(n * sizeof(V3)), *hi = (V3 *)malloc(n * sizeof(V3)), *ce = (V3 *)malloc(n * sizeof(V3));
    read_boxes(boxes, n, lo, hi, ce);
    uint32_t *order = out->order;
    Job *stack = (Job *)malloc(sizeof(Job) * (2 * max_depth + 8));
    uint64_t sp = 0, deepest = 0;
    stack[sp++] = (Job){0, 0, n, 0};
    while (sp) {
        Job j = stack[--sp];
        if (j.depth > deepest) deepest = j.depth;
        V3 l = {{INFINITY, INFINITY, INFINITY}}, h = {{-INFINITY, -INFINITY, -INFINITY}}, cl = l, ch = h;
        for (uint64_t q = j.start; q < j.end; q++) {
            uint32_t k = order[q];
            l = vmin(l, lo[k]);
            h = vmax(h, hi[k]);
            cl = vmin(cl, ce[k]);
            ch = vmax(ch, ce[k]);
        }
        uint64_t count = j.end - j.start;
        if (count == 1) {
            nv.v[j.ni] = mk(l, (uint32_t)j.start, h, 1);
            continue;
        }
        uint64_t shift = max_depth - j.depth - 1;
        uint64_t room = shift >= 64 ? UINT64_MAX : (1ull << shift);
        V3 ext;
        for (int k = 0; k < 3; k++) ext.v[k] = ch.v[k] - cl.v[k];
        float best_cost = INFINITY;
        int best_axis = 0;
        uint64_t best_split = 0;
        for (int axis = 0; axis < 3; axis++) {
            if (!(ext.v[axis] > 0.0f)) continue;
            V3 blo[TOP_BINS], bhi[TOP_BINS];
            uint64_t bc[TOP_BINS];
            for (int i = 0; i < TOP_BINS; i++) {
                blo[i] = (V3){{INFINITY, INFINITY, INFINITY}};
                bhi[i] = (V3){{-INFINITY, -INFINITY, -INFINITY}};
                bc[i] = 0;
            }
            for (uint64_t q = j.start; q < j.end; q++) {
                uint32_t k = order[q];
                uint64_t b = as_usize((ce[k].v[axis] - cl.v[axis]) * ((float)TOP_BINS / ext.v[axis]));
                if (b > TOP_BINS - 1) b = TOP_BINS - 1;
                blo[b] = vmin(blo[b], lo[k]);
                bhi[b] = vmax(bhi[b], hi[k]);
                bc[b]++;
            }
            float right[TOP_BINS];
            memset(right, 0, sizeof(right));
            V3 al = {{INFINITY, INFINITY, INFINITY}}, ah = {{-INFINITY, -INFINITY, -INFINITY}};
            for (int i = TOP_BINS - 1; i >= 1; i--) {
                al = vmin(al, blo[i]);
                ah = vmax(ah, bhi[i]);
                right[i] = area_of(al, ah);
            }
            al = (V3){{INFINITY, INFINITY, INFINITY}};
            ah = (V3){{-INFINITY, -INFINITY, -INFINITY}};
            uint64_t nl = 0;
            for (int i = 1; i < TOP_BINS; i++) {
                al = vmin(al, blo[i - 1]);
                ah = vmax(ah, bhi[i - 1]);
                nl += bc[i - 1];
                uint64_t nr = count - nl;
                if (nl == 0 || nr == 0 || nl > room || nr > room) continue;
                float cost = area_of(al, ah) * (float)nl + right[i] * (float)nr;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = (uint64_t)i;
                }
            }
        }
        uint64_t mid;
        if (isfinite(best_cost)) {
            uint32_t *slice = order + j.start;
            uint64_t i = 0;
            for (uint64_t q = 0; q < count; q++) {
                uint64_t b =
                    as_usize((ce[slice[q]].v[best_axis] - cl.v[best_axis]) * ((float)TOP_BINS / ext.v[best_axis]));
                if (b > TOP_BINS - 1) b = TOP_BINS - 1;
                if (b < best_split) {
                    uint32_t tmp = slice[i];
                    slice[i] = slice[q];
                    slice[q] = tmp;
                    i++;
                }
            }
            mid = i;
        } else {
            int axis = (ext.v[0] >= ext.v[1] && ext.v[0] >= ext.v[2]) ? 0 : (ext.v[1] >= ext.v[2] ? 1 : 2);
            sort_by_axis(order + j.start, count, ce, axis);
            mid = count / 2;
        }
        uint64_t left = nv.n;
        nv_push(&nv, NODE_EMPTY);
        nv_push(&nv, NODE_EMPTY);
        nv.v[j.ni] = mk(l, (uint32_t)left, h, 0);
        stack[sp++] = (Job){left + 1, j.start + mid, j.end, j.depth + 1};
        stack[sp++] = (Job){left, j.start, j.start + mid, j.depth + 1};
    }
    free(stack);
    free(lo);
    free(hi);
    free(ce);
    out->nodes = nv.v;
    out->node_count = nv.n;
    out->depth = (uint32_t)deepest;
    return 0;
}

/* units.rs:231-246 mesh_tree (with tri_boxes, units.rs:225-227). */
int or_mesh_tree(const float *tris, uint64_t n, or_tree *out) {
    memset(out, 0, sizeof(*out));
    if (n == 0) return -1;
    for (uint64_t i = 0; i < 9 * n; i++)
        if (!isfinite(tris[i])) return -1;
    char msg[256];
    or_tree b;
    if (or_build_tris(tris, n, OR_BVH_STACK, &b)) return -1;
    int d = or_check_tree(b.nodes, b.node_count, OR_MESH_DEPTH, n, 16, msg, sizeof msg);
    if (d < 0) {
        or_tree_free(&b);
        float *boxes = (float *)malloc(n * 6 * sizeof(float));
        for (uint64_t i = 0; i < n; i++) {
            const float *t = tris + 9 * i;
            for (int k = 0; k < 3; k++) {
                boxes[6 * i + k] = smin(smin(t[k], t[3 + k]), t[6 + k]);
                boxes[6 * i + 3 + k] = smax(smax(t[k], t[3 + k]), t[6 + k]);
            }
        }
        or_build_median(boxes, n, 4, &b);
        free(boxes);
        d = or_check_tree(b.nodes, b.node_count, OR_MESH_DEPTH, n, 16, msg, sizeof msg);
        if (d < 0) {
            or_tree_free(&b);
            return -1;
        }
    }
    *out = b;
    out->depth = (uint32_t)d;
    V3 lo = {{INFINITY, INFINITY, INFINITY}}, hi = {{-INFINITY, -INFINITY, -INFINITY}};
    for (uint64_t i = 0; i < 3 * n; i++) {
        V3 p = {{tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]}};
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    memcpy(out->lo, lo.v, 12);
    memcpy(out->hi, hi.v, 12);
    return 0;
}

/* units.rs:288-297 UnitBox::new. world: glam Mat4, column-major (x_axis, y_axis, z_axis, w_axis). */
void or_unit_box(const float *w, const float *lo, const float *hi, float *out) {
    V3 a = {{INFINITY, INFINITY, INFINITY}}, b = {{-INFINITY, -INFINITY, -INFINITY}};
    for (int c = 0; c < 8; c++) {
        float px = (c & 1) ? hi[0] : lo[0], py = (c & 2) ? hi[1] : lo[1], pz = (c & 4) ? hi[2] : lo[2];
        V3 p;
        for (int k = 0; k < 3; k++) {
            float r = w[0 + k] * px;
            r = w[4 + k] * py + r;
            r = w[8 + k] * pz + r;
            r = w[12 + k] + r;
            p.v[k] = r;
        }
        a = vmin(a, p);
        b = vmax(b, p);
    }
    /* max_element (SSE2): max(max(x, z), y) */
    float e[3], m[3];
    for (int k = 0; k < 3; k++) {
        e[k] = b.v[k] - a.v[k];
        m[k] = smax(fabsf(a.v[k]), fabsf(b.v[k]));
    }
    float emax = smax(smax(e[0], e[2]), e[1]);
    float mmax = smax(smax(m[0], m[2]), m[1]);
    float pad = emax * 1e-4f + mmax * 1e-5f + 1e-6f;
    for (int k = 0; k < 3; k++) {
        out[k] = a.v[k] - pad;
        out[3 + k] = b.v[k] + pad;
    }
}

/* units.rs:303-321 top_tree */
int or_top_tree(const float *boxes, const uint32_t *records, uint64_t n, or_tree *out) {
    memset(out, 0, sizeof(*out));
    if (n == 0 || n > OR_MAX_UNITS) return -1;
    for (uint64_t i = 0; i < 6 * n; i++)
        if (!isfinite(boxes[i])) return -1;
    if (or_build_sah(boxes, n, OR_TOP_DEPTH - 1, out)) return -1;
    uint32_t maxrec = 0;
    for (uint64_t i = 0; i < n; i++)
        if (records[i] > maxrec) maxrec = records[i];
    for (uint64_t i = 0; i < out->node_count; i++)
        if (out->nodes[i].count > 0) out->nodes[i].left_or_first = records[out->order[out->nodes[i].left_or_first]];
    char msg[256];
    int d = or_check_top(out->nodes, out->node_count, maxrec + 1, msg, sizeof msg);
    if (d < 0) {
        or_tree_free(out);
        return -1;
    }
    out->depth = (uint32_t)d;
    return 0;
}

static inline uint32_t fbits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float bitsf(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* units.rs:329-342 unit_record, translation_alone */
void or_unit_record(or_instance *r, uint32_t root, int cast_shadow) {
    int alone = 1;
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) alone &= r->n[k][c] == (k == c ? 1.0f : 0.0f);
    r->n[0][3] = bitsf(root);
    r->n[1][3] = bitsf((fbits(r->n[1][3]) & OR_INST_BITS) | (cast_shadow ? OR_UNIT_SHADOW : 0u) |
                       (alone ? OR_UNIT_SHIFTED : 0u));
}

/* units.rs:365-367 top_cost */
double or_top_cost(const or_node *nodes, uint64_t count) {
    double s = 0.0;
    for (uint64_t i = 0; i < count; i++) {
        if (nodes[i].count != 0) continue;
        V3 lo = {{nodes[i].bmin[0], nodes[i].bmin[1], nodes[i].bmin[2]}};
        V3 hi = {{nodes[i].bmax[0], nodes[i].bmax[1], nodes[i].bmax[2]}};
        s += (double)area_of(lo, hi);
    }
    return s;
}
