/*
 * The oracle: a plain-C CPU restatement of ThreeRT's software path-trace pass, used ONLY as the checker by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg. Nothing in the threert_amd product links or
 * calls it.
 *
 * Every function names the reference file:line it restates (paths relative to the ThreeRT repository root).
 * Arithmetic follows the reference's operation order with no FMA contraction (built with -ffp-contract=off); where
 * the reference leans on glam's SSE2 Vec3A, the SSE2 lane semantics are restated (min/max return the second operand
 * when either is NaN; f32 -> usize casts saturate and send NaN to 0).
 *
 * Pinning (see tests/test_oracle_*.py): the builders against the reference's own test scenarios and invariants
 * (check_tree, depth caps, node counts, SAH-vs-median cost), the walk against brute force over every triangle
 * (closest t bit-equal, as units.rs:1405 holds it), the triangle and box tests against an FP64 Moller-Trumbore and
 * closed-form special cases, the sampler against the Sobol known answers and the (0,m,2)-net property.
 *
 * Parity unpinned: glam's Mat3::inverse().transpose() bits (records are computed by the caller, the same bytes go
 * to both sides); the order Rust's select_nth_unstable leaves items in (the oracle sorts by (key, index) instead;
 * only the depth-cap median fallbacks use it).
 */
#ifndef THREERT_ORACLE_H
#define THREERT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* bvh.rs:6-15 GpuNode: 32 bytes. count 0 = interior (children left, left+1). */
typedef struct {
    float bmin[3];
    uint32_t left_or_first;
    float bmax[3];
    uint32_t count;
} or_node;

/* indexed.rs:19-26 Leaf: (v0, section id), (e1, index in section), (e2, 0). 48 bytes. */
typedef struct {
    float r[3][4];
} or_leaf;

/* scene.rs:229-234 GpuInstance: 160 bytes. */
typedef struct {
    float m[3][4];
    float n[3][4];
    uint32_t info[4];
    float p[3][4];
} or_instance;

/* pathtrace.wgsl Hit(t, tri, u, v, inst, part): 24 bytes. */
typedef struct {
    float t;
    uint32_t tri;
    float u, v;
    uint32_t inst;
    uint32_t part;
} or_hit;

/* units.rs:922-927 Reads. */
typedef struct {
    uint64_t top_nodes, mesh_nodes, mesh_verts, records, steps;
    uint64_t deepest;
} or_reads;

/* A tree and the order it puts its items in (bvh.rs:39-47 Bvh, units.rs MeshTree). */
typedef struct {
    or_node *nodes;
    uint64_t node_count;
    uint32_t *order;
    uint64_t order_count;
    uint32_t depth;
    float lo[3], hi[3];
} or_tree;

/* The kernel's buffers (units.rs:882-903 Kernel). */
typedef struct {
    const or_node *nodes;
    uint64_t node_count;
    const or_leaf *leaves;
    uint64_t leaf_count;
    const or_instance *instances;
    uint64_t instance_count;
    uint32_t top;   /* P.motion.z */
    uint32_t units; /* P.motion.w */
} or_scene;

/* A ray of a batch: 32 bytes (origin, tmax, direction, mode). */
typedef struct {
    float o[3];
    float tmax;
    float d[3];
    uint32_t mode;
} or_ray;

#define OR_INF 3.0e38f
#define OR_NO_INST 0xffffffffu
#define OR_NO_TRI 0xffffffffu
#define OR_TRACE_CLOSEST 0u
#define OR_TRACE_SHADOW 1u
#define OR_TRACE_ANY 2u
#define OR_UNIT_SHADOW 1u
#define OR_UNIT_SHIFTED 2u
#define OR_INST_BITS (4u | 16u | 32u | 64u | 128u)

#define OR_BVH_STACK 48u   /* bvh.rs:56 */
#define OR_STACK 64u       /* units.rs:30 */
#define OR_TOP_DEPTH 19u   /* units.rs:33 */
#define OR_MESH_DEPTH 45u  /* units.rs:35 */
#define OR_MAX_UNITS (1u << 18) /* units.rs:37 */

/* Walk variants. REF: trace_units.wgsl as written (strict `<`, first found wins a tie, boxes culled at hit.t).
 * EXACT: the same walk, with ties broken to the lowest (instance, triangle), boxes culled at hit.t widened by
 * OR_CULL_SLACK, and every box widened before its slab test by 2^-8 of its largest extent plus 2^-20 of its largest
 * coordinate: a box's float slab test can miss a ray that the triangle test inside it accepts (an edge of a
 * triangle on a flat box's face), so only a box with a margin finds what brute force finds, whatever the tree. */
#define OR_WALK_REF 0u
#define OR_WALK_EXACT 1u
/* WT: EXACT with the watertight triangle test (Woop, Benthin and Wald 2013) in place of Moller-Trumbore. */
#define OR_WALK_WT 2u
#define OR_CULL_SLACK (1.0f / 16384.0f)

void or_tree_free(or_tree *t);

/* bvh.rs:81-106 build_tris (tris: n x 9 floats). Returns 0, or -1 when the items do not fit the depth. */
int or_build_tris(const float *tris, uint64_t n, uint32_t max_depth, or_tree *out);
/* bvh.rs:246-282 check_tree with the leaf rule of bvh.rs:235 (count <= 16, inside `items`). Returns the depth, or
 * -1 (msg filled). */
int or_check_tree(const or_node *nodes, uint64_t count, uint32_t max_depth, uint64_t items, uint32_t max_leaf,
                  char *msg, uint64_t msg_len);
/* units.rs:68-79 check_top: leaves of one record below `records`, each once. */
int or_check_top(const or_node *nodes, uint64_t count, uint32_t records, char *msg, uint64_t msg_len);
/* units.rs:84-118 build_median over boxes (n x 6: min xyz, max xyz). */
int or_build_median(const float *boxes, uint64_t n, uint32_t max_leaf, or_tree *out);
/* units.rs:140-223 build_sah over boxes. */
int or_build_sah(const float *boxes, uint64_t n, uint32_t max_depth, or_tree *out);
/* units.rs:231-246 mesh_tree. Returns 0, or -1 (no triangles / a vertex not finite). */
int or_mesh_tree(const float *tris, uint64_t n, or_tree *out);
/* units.rs:288-297 UnitBox::new; world: column-major 4x4 (glam Mat4). out: lo xyz, hi xyz. */
void or_unit_box(const float *world, const float *lo, const float *hi, float *out);
/* units.rs:303-321 top_tree: boxes n x 6, records n. Leaves' left_or_first are records. */
int or_top_tree(const float *boxes, const uint32_t *records, uint64_t n, or_tree *out);
/* units.rs:329-342 unit_record + translation_alone, in place. */
void or_unit_record(or_instance *r, uint32_t root, int cast_shadow);
/* units.rs:365-367 top_cost. */
double or_top_cost(const or_node *nodes, uint64_t count);

/* pathtrace.wgsl:732-740 ray_aabb. */
float or_ray_aabb(const float *o, const float *inv_d, const float *bmin, const float *bmax, float tmax);
/* pathtrace.wgsl:742-760 intersect_tri: writes (t, u, v); t = INF for a miss. */
void or_intersect_tri(const float *o, const float *d, const or_leaf *l, float tmax, float *tuv);
/* units.rs:1004-1006 Kernel::inv. */
void or_inv(const float *d, float *inv);

/* trace_units.wgsl:44-135 (units.rs:1031-1122); variant OR_WALK_*. reads may be NULL. Returns 0, or -1 on an
 * out-of-buffer read or a walk that does not stop within `budget` steps (0: 64 x (nodes + 3 leaves) + 1024). */
int or_trace(const or_scene *s, const float *o, const float *d, float tmax, uint32_t mode, uint32_t variant,
             uint64_t budget, or_hit *hit, or_reads *reads);
/* Every triangle of every unit of the top tree, in the unit's space (units.rs:1369-1388); variant REF keeps the
 * first found at the closest t, EXACT the lowest (instance, triangle). */
int or_brute(const or_scene *s, const float *o, const float *d, float tmax, uint32_t mode, uint32_t variant,
             or_hit *hit);
/* A batch through or_trace on `threads` OpenMP threads (0: all). Returns the number of failed rays. */
int64_t or_trace_batch(const or_scene *s, const or_ray *rays, uint64_t n, uint32_t variant, int threads,
                       or_hit *hits, or_reads *reads_total);
/* The closest hit in double precision by brute force: every triangle of every unit (the ray placed in the unit's
 * space as the walk places it, then the stored v0, e1, e2 by Moller-Trumbore, every operation in double); ties to the
 * lowest (instance, triangle). t64 = -1 and tri = OR_NO_TRI for a miss. The judge of rays whose float tests disagree. */
typedef struct {
    double t, u, v;
    uint32_t tri, inst;
} or_hit64;
int64_t or_brute64_batch(const or_scene *s, const or_ray *rays, uint64_t n, int threads, or_hit64 *hits);

int64_t or_brute_batch(const or_scene *s, const or_ray *rays, uint64_t n, uint32_t variant, int threads,
                       or_hit *hits);

#ifdef __cplusplus
}
#endif
#endif
