// Device code of the ray walks, shared by the ray-batch kernels (trace.hip) and the path tracer (render.hip).
//
// Two walks over the same scene, one answer: the lowest (t, instance, triangle) among the triangles the ray meets
// in (0, tmax), with every triangle tested by the reference's Moller-Trumbore (pathtrace.wgsl:742-760) in the
// reference's operation order (these files are built with -ffp-contract=off), so t, u and v are the reference's bits.
//   trace_binary  the reference's own 32-byte binary node buffer, trace_units.wgsl:44-135 statement for statement
//                 (64-entry stack, near child first), boxes culled at hit.t widened by kCullSlack and each box
//                 widened before its slab test by 2^-8 of its largest extent plus 2^-20 of its largest coordinate
//                 (a float slab test can miss a ray the triangle test inside the box accepts: an edge of a triangle
//                 on a flat box's face; the quantised boxes of the wide walk are wider than that by construction).
//   trace_wide    the library's 64-byte 4-wide quantised nodes (wide.cpp): each node's four child boxes are tested
//                 as fma(q, step / d, (origin - o) / d) per plane, children visited nearest first.
// The stack's first kLdsStack entries live in LDS (stride = block size, so lanes never share a bank); deeper entries
// spill to a per-thread region of device memory sized from the scene's deepest possible walk.
#pragma once

#include <hip/hip_runtime.h>

#include "trace_kernels.h"

namespace trt {
namespace walk {

constexpr float kCullSlack = 1.0f / 16384.0f;
constexpr uint32_t kTraceClosest = 0, kTraceShadow = 1, kTraceAny = 2;

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ F3 inv3(F3 d) {
    return {1.0f / (fabsf(d.x) < 1e-20f ? 1e-20f : d.x), 1.0f / (fabsf(d.y) < 1e-20f ? 1e-20f : d.y),
            1.0f / (fabsf(d.z) < 1e-20f ? 1e-20f : d.z)};
}
__device__ __forceinline__ float cull_t(float t) { return t + t * kCullSlack; }

// pathtrace.wgsl:732-740 ray_aabb.
__device__ __forceinline__ float ray_aabb(F3 o, F3 inv, F3 bmin, F3 bmax, float tmax) {
    const float t0x = (bmin.x - o.x) * inv.x, t0y = (bmin.y - o.y) * inv.y, t0z = (bmin.z - o.z) * inv.z;
    const float t1x = (bmax.x - o.x) * inv.x, t1y = (bmax.y - o.y) * inv.y, t1z = (bmax.z - o.z) * inv.z;
    const float lx = t0x < t1x ? t0x : t1x, ly = t0y < t1y ? t0y : t1y, lz = t0z < t1z ? t0z : t1z;
    const float hx = t0x > t1x ? t0x : t1x, hy = t0y > t1y ? t0y : t1y, hz = t0z > t1z ? t0z : t1z;
    const float tn = fmaxf(fmaxf(lx, ly), fmaxf(lz, 0.0f));
    const float tf = fminf(fminf(hx, hy), fminf(hz, tmax));
    return tn <= tf ? tn : kInf;
}

// The binary walk's box test: ray_aabb over the box widened by 2^-8 of its largest extent plus 2^-20 of its largest
// coordinate.
__device__ __forceinline__ float ray_aabb_wide(F3 o, F3 inv, F3 bmin, F3 bmax, float tmax) {
    const float ext = fmaxf(fmaxf(bmax.x - bmin.x, bmax.y - bmin.y), bmax.z - bmin.z);
    float mag = fmaxf(0.0f, fmaxf(fabsf(bmin.x), fabsf(bmax.x)));
    mag = fmaxf(mag, fmaxf(fabsf(bmin.y), fabsf(bmax.y)));
    mag = fmaxf(mag, fmaxf(fabsf(bmin.z), fabsf(bmax.z)));
    const float m = ext * 0.00390625f + mag * 9.5367431640625e-07f;
    return ray_aabb(o, inv, {bmin.x - m, bmin.y - m, bmin.z - m}, {bmax.x + m, bmax.y + m, bmax.z + m}, tmax);
}

// pathtrace.wgsl:742-760 intersect_tri; t = kInf for a miss.
__device__ __forceinline__ void intersect_tri(F3 o, F3 d, float4 l0, float4 l1, float4 l2, float tmax, float &t,
                                              float &u, float &v) {
    t = kInf;
    const F3 v0{l0.x, l0.y, l0.z}, e1{l1.x, l1.y, l1.z}, e2{l2.x, l2.y, l2.z};
    const F3 pv = cross3(d, e2);
    const float det = dot3(e1, pv);
    if (fabsf(det) < 1e-12f) return;
    const float inv = 1.0f / det;
    const F3 tv = sub3(o, v0);
    const float uu = dot3(tv, pv) * inv;
    if (!(uu >= 0.0f && uu <= 1.0f)) return;
    const F3 qv = cross3(tv, e1);
    const float vv = dot3(d, qv) * inv;
    if (vv < 0.0f || uu + vv > 1.0f) return;
    const float tt = dot3(e2, qv) * inv;
    if (tt <= 0.0f || tt >= tmax) return;
    t = tt, u = uu, v = vv;
}

// The watertight triangle test (Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2013): the
// triangle's vertices translated to the ray's origin and sheared so the ray runs along +z, then the 2-D edge functions
// U, V, W of the projected vertices (the same float operations for an edge however a triangle names it, so a ray on
// an edge shared by two triangles meets one of them); an edge function that rounds to exactly 0 is evaluated again
// in double. The vertices are v0, v0 + e1, v0 + e2 from the leaf. u = V / det, v = W / det weigh v1 and v2, as
// Moller-Trumbore's u and v do.
struct Shear {
    uint32_t kx, ky, kz;
    float sx, sy, sz;
};
__device__ __forceinline__ float comp(F3 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ Shear shear_of(F3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    Shear s;
    s.kz = ax >= ay ? (ax >= az ? 0u : 2u) : (ay >= az ? 1u : 2u);  // the largest component, the first of equals
    s.kx = s.kz == 2u ? 0u : s.kz + 1u;
    s.ky = s.kx == 2u ? 0u : s.kx + 1u;
    if (comp(d, s.kz) < 0.0f) {
        const uint32_t t = s.kx;
        s.kx = s.ky, s.ky = t;
    }
    s.sx = comp(d, s.kx) / comp(d, s.kz);
    s.sy = comp(d, s.ky) / comp(d, s.kz);
    s.sz = 1.0f / comp(d, s.kz);
    return s;
}
__device__ __forceinline__ void intersect_tri_wt(F3 o, const Shear &k, float4 l0, float4 l1, float4 l2, float tmax,
                                                 float &t, float &u, float &v) {
    t = kInf;
    const F3 v0{l0.x, l0.y, l0.z};
    const F3 v1{l0.x + l1.x, l0.y + l1.y, l0.z + l1.z}, v2{l0.x + l2.x, l0.y + l2.y, l0.z + l2.z};
    const F3 A = sub3(v0, o), B = sub3(v1, o), C = sub3(v2, o);
    const float Az0 = comp(A, k.kz), Bz0 = comp(B, k.kz), Cz0 = comp(C, k.kz);
    const float ax = comp(A, k.kx) - k.sx * Az0, ay = comp(A, k.ky) - k.sy * Az0;
    const float bx = comp(B, k.kx) - k.sx * Bz0, by = comp(B, k.ky) - k.sy * Bz0;
    const float cx = comp(C, k.kx) - k.sx * Cz0, cy = comp(C, k.ky) - k.sy * Cz0;
    float U = cx * by - cy * bx, V = ax * cy - ay * cx, W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = static_cast<float>(static_cast<double>(cx) * by - static_cast<double>(cy) * bx);
        V = static_cast<float>(static_cast<double>(ax) * cy - static_cast<double>(ay) * cx);
        W = static_cast<float>(static_cast<double>(bx) * ay - static_cast<double>(by) * ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return;
    const float det = U + V + W;
    if (det == 0.0f) return;
    const float T = U * (k.sz * Az0) + V * (k.sz * Bz0) + W * (k.sz * Cz0);
    const float rcp = 1.0f / det;
    const float tt = T * rcp;
    if (!(tt > 0.0f) || tt >= tmax) return;
    t = tt, u = V * rcp, v = W * rcp;
}

// The first out-of-range read of a launch, for the host's error message: guard word 1 = which check (a kRange id),
// words 2-3 = the index (low, high), word 4 = the size it was checked against. Word 0 holds the kFlag bits.
// Words 5-7 may carry three values of the reporting thread's context (ctx non-null). Returns whether this thread's
// report is the one kept.
__device__ __forceinline__ bool report_range(uint32_t *guard, uint32_t id, uint64_t i, uint64_t n,
                                             const uint32_t *ctx = nullptr) {
    if (atomicCAS(guard + 1, 0u, id) != 0u) return false;
    guard[2] = static_cast<uint32_t>(i);
    guard[3] = static_cast<uint32_t>(i >> 32);
    guard[4] = static_cast<uint32_t>(n);
    if (ctx) guard[5] = ctx[0], guard[6] = ctx[1], guard[7] = ctx[2];
    return true;
}

struct HitRec {
    float t;
    uint32_t tri;
    float u, v;
    uint32_t inst, part;
};

// The tie rule: a candidate at t in (0, tmax) replaces the hit when nearer, or as near with a lower
// (instance, triangle).
__device__ __forceinline__ bool accept(const HitRec &h, float x, uint32_t inst, uint32_t tri) {
    if (x < h.t) return true;
    return x == h.t && h.tri != kNone && (inst < h.inst || (inst == h.inst && tri < h.tri));
}

// The hit update as one select per field, with no branch around it. Written as a branch (`if (take) h = {...}`)
// and inlined into the path tracer, hipcc for gfx950 (ROCm 7.2, -O3) kept the section id of the replaced hit when
// the new one won an exact-t tie: the hit then named one triangle and another's section.
__device__ __forceinline__ void take_hit(HitRec &h, bool take, float x, uint32_t tri, float u, float v,
                                         uint32_t inst, uint32_t part) {
    h.t = take ? x : h.t;
    h.tri = take ? tri : h.tri;
    h.u = take ? u : h.u;
    h.v = take ? v : h.v;
    h.inst = take ? inst : h.inst;
    h.part = take ? part : h.part;
}

// pathtrace.wgsl proc_dir: n0 * d.x + n1 * d.y + n2 * d.z.
__device__ __forceinline__ F3 proc_dir(F3 n0, F3 n1, F3 n2, F3 d) {
    return {n0.x * d.x + n1.x * d.y + n2.x * d.z, n0.y * d.x + n1.y * d.y + n2.y * d.z,
            n0.z * d.x + n1.z * d.y + n2.z * d.z};
}

// The LDS part of a stack is reached through an LDS-qualified pointer: with a generic one the compiler merges the
// LDS and the spill read of a pop into one flat load, which waits for every memory access the wave has in flight.
#ifndef TRT_LDS
#define TRT_LDS __attribute__((address_space(3)))
#endif

struct Stack {
    TRT_LDS uint32_t *lds;  // this thread's entry 0; entry k at lds[k * kTraceBlock]
    uint32_t *spill;        // the launch's spill region: this thread's entry k at spill[k * stride + slot]
    uint32_t slot;
    uint32_t stride;
    uint32_t sp;
    uint32_t cap;     // entries in all: kLdsStack + the spill entries; the walks check it before they push
    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < kLdsStack)
            lds[sp * kTraceBlock] = v;
        else
            spill[uint64_t(sp - kLdsStack) * stride + slot] = v;
        sp++;
    }
    __device__ __forceinline__ uint32_t pop() {
        sp--;
        return sp < kLdsStack ? lds[sp * kTraceBlock] : spill[uint64_t(sp - kLdsStack) * stride + slot];
    }
};

__device__ __forceinline__ void store_hit(uint32_t *hits, uint64_t i, const HitRec &h) {
    uint32_t *p = hits + 6 * i;
    p[0] = __float_as_uint(h.t);
    p[1] = h.tri;
    p[2] = __float_as_uint(h.u);
    p[3] = __float_as_uint(h.v);
    p[4] = h.inst;
    p[5] = h.part;
}

// What a walk read (units.rs:922-927 Reads): nodes fetched, triangles tested, unit records read.
struct Reads {
    uint32_t nodes = 0, tris = 0, units = 0;
    uint32_t shaded = 0;  // the path tracer's counting launches: hits shaded
};

// ---- the reference's binary walk ------------------------------------------------------------------------------
template <bool kCount>
__device__ HitRec trace_binary(const TraceArgs &a, F3 o, F3 d, float tmax_in, uint32_t mode, Stack &st,
                               uint32_t &flags, Reads &cnt) {
    const float tmax = tmax_in < kInf ? tmax_in : kInf;
    HitRec h{tmax, kNone, 0.0f, 0.0f, kNone, 0};
    if (a.units == 0) return h;
    const F3 inv = inv3(d);
    const uint32_t top = a.top_first;
    if (top >= a.bnode_count) {
        flags |= kFlagWalk;
        report_range(a.guard_flag, 1, top, a.bnode_count);
        return h;
    }
    {
        if (kCount) cnt.nodes++;
        const float4 w0 = a.bnodes[2 * top], w1 = a.bnodes[2 * top + 1];
        if (ray_aabb_wide(o, inv, {w0.x, w0.y, w0.z}, {w1.x, w1.y, w1.z}, cull_t(h.t)) == kInf) return h;
    }
    st.sp = 0;
    uint32_t node = top, rec = kNone, off = top, first = 0, base = 0;
    F3 ro = o, rd = d, rinv = inv;
    uint32_t guard = 0;
    for (;;) {
        if (++guard > kGuardSteps) {
            flags |= kFlagGuard;
            return h;
        }
        if (node >= a.bnode_count) {
            flags |= kFlagWalk;
            report_range(a.guard_flag, 2, node, a.bnode_count);
            return h;
        }
        const float4 n0 = a.bnodes[2 * node], n1 = a.bnodes[2 * node + 1];
        const uint32_t left_or_first = __float_as_uint(n0.w), count = __float_as_uint(n1.w);
        if (kCount) cnt.nodes++;
        if (count == 0) {
            if (kCount) cnt.nodes += 2;
            const uint32_t l = off + left_or_first, r = l + 1;
            if (uint64_t(off) + left_or_first + 1 >= a.bnode_count) {
                flags |= kFlagWalk;
                report_range(a.guard_flag, 3, uint64_t(off) + left_or_first + 1, a.bnode_count);
                return h;
            }
            const float4 l0 = a.bnodes[2 * l], l1 = a.bnodes[2 * l + 1];
            const float4 r0 = a.bnodes[2 * r], r1 = a.bnodes[2 * r + 1];
            float tl = ray_aabb_wide(ro, rinv, {l0.x, l0.y, l0.z}, {l1.x, l1.y, l1.z}, cull_t(h.t));
            float tr = ray_aabb_wide(ro, rinv, {r0.x, r0.y, r0.z}, {r1.x, r1.y, r1.z}, cull_t(h.t));
            if (tl != kInf || tr != kInf) {
                uint32_t near = l, far = r;
                if (tr < tl) {
                    near = r, far = l;
                    const float tmp = tl;
                    tl = tr, tr = tmp;
                }
                node = near;
                if (tr != kInf && st.sp < kRefStack) {
                    if (st.sp >= st.cap) {
                        flags |= kFlagWalk;
                        report_range(a.guard_flag, 4, st.sp, st.cap);
                        return h;
                    }
                    st.push(far);
                }
                continue;
            }
        } else if (rec == kNone) {
            if (kCount) cnt.units++;
            if (left_or_first >= a.instance_count) {
                flags |= kFlagWalk;
                report_range(a.guard_flag, 5, left_or_first, a.instance_count);
                return h;
            }
            const float4 *u = a.instances + 10 * uint64_t(left_or_first);
            const float4 m0 = u[0], m1 = u[1], m2 = u[2], q0 = u[3], q1 = u[4], q2 = u[5];
            const uint32_t flags = __float_as_uint(q1.w);
            if (mode != kTraceShadow || (flags & kUnitShadow) != 0) {
                rec = left_or_first;
                ro = {o.x - m0.w, o.y - m1.w, o.z - m2.w};
                if ((flags & kUnitShifted) == 0) {
                    const F3 c0{q0.x, q0.y, q0.z}, c1{q1.x, q1.y, q1.z}, c2{q2.x, q2.y, q2.z};
                    ro = proc_dir(c0, c1, c2, ro);
                    rd = proc_dir(c0, c1, c2, d);
                    rinv = inv3(rd);
                }
                off = __float_as_uint(q0.w);
                first = __float_as_uint(u[6].x);
                base = st.sp;
                node = off;
                continue;
            }
        } else {
            if (uint64_t(first) + left_or_first + count > a.leaf_count) {
                flags |= kFlagWalk;
                report_range(a.guard_flag, 6, uint64_t(first) + left_or_first + count, a.leaf_count);
                return h;
            }
            for (uint32_t k = 0; k < count; k++) {
                const uint32_t ti = first + left_or_first + k;
                if (kCount) cnt.tris++;
                const float4 *l = a.leaves + 3 * uint64_t(ti);
                const float4 lv0 = l[0], lv1 = l[1], lv2 = l[2];
                float x, uu, vv;
                intersect_tri(ro, rd, lv0, lv1, lv2, tmax, x, uu, vv);
                const bool take = x < tmax && accept(h, x, rec, ti);
                take_hit(h, take, x, ti, uu, vv, rec, __float_as_uint(lv0.w));
                if (take && mode != kTraceClosest) return h;
            }
        }
        if (rec != kNone && st.sp == base) {
            rec = kNone;
            ro = o, rd = d, rinv = inv;
            off = top;
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    return h;
}

// ---- the 4-wide quantised walk ----------------------------------------------------------------------------------
__device__ __forceinline__ float qbyte(uint32_t w, int c) { return float((w >> (8 * c)) & 0xffu); }

// A wide walk in flight: the ray in world space and in the current unit's space, the next reference to visit or a
// state of kind 3 (kWalkLeave, kWalkDone), the unit being walked (its record, first triangle, and the stack depth at
// its entry) and the best hit so far.
//
// A wave runs its walks in two alternating phases. In the node phase, lanes whose next reference is a node test it
// and move on, while lanes that have reached anything else wait: a run of triangles, a unit to enter, or the end of
// a unit's tree (kWalkLeave). Once a set share of the walking lanes wait (or no lane has a node left), the other
// phase serves every waiting lane at once. Mixing them in every step would run the triangle loop, for a lane or two,
// on nearly every step of the wave; and with the ray's unit space fixed during the node phase, the node loop keeps
// no copies of it. Each lane still visits its references in the order of a lone walk, so the hits and the reads are
// a lone walk's.
constexpr uint32_t kWalkDone = 0xffffffffu;   // the walk is over
constexpr uint32_t kWalkLeave = 0xfffffffeu;  // the stack is back at the unit's entry depth: leave the unit, then pop
constexpr uint32_t kLeafShareBatch = 2;  // ray batches (trace.hip): the other phase once half of the lanes wait
constexpr uint32_t kLeafShareCall = 2;   // trace_wide (the path tracer's rays): once half of them wait
// (measured on the path tracer: 1, serving lanes as soon as they all wait, took the 720p city frame from 1.94 to
// 1.90 ms but the 3M-triangle 1080p image from 1.36 to 2.14 s; 3 made the city 2.02 ms and the large image no
// faster)

struct WideWalk {
    F3 o, d, ro, rd, rinv;
    float tmax;
    uint32_t mode, node, rec, first, base, guard;
    HitRec h;
};

__device__ __forceinline__ uint32_t ref_kind(uint32_t r) { return r >> kRefKindShift; }
// The lane takes node-phase steps.
__device__ __forceinline__ bool wide_at_node(const WideWalk &w) { return ref_kind(w.node) == kRefNode; }
// The lane waits for the other phase: triangles, a unit to enter, or a unit to leave.
__device__ __forceinline__ bool wide_waits(const WideWalk &w) {
    return ref_kind(w.node) != kRefNode && w.node != kWalkDone;
}
// The walk is over.
__device__ __forceinline__ bool wide_over(const WideWalk &w) { return w.node == kWalkDone; }

// Starts a walk; false when it is already finished (no units, or the ray misses the top box): w.h is the answer.
__device__ __forceinline__ bool wide_begin(const TraceArgs &a, F3 o, F3 d, float tmax_in, uint32_t mode, WideWalk &w,
                                           Stack &st) {
    w.tmax = tmax_in < kInf ? tmax_in : kInf;
    w.mode = mode;
    w.h = HitRec{w.tmax, kNone, 0.0f, 0.0f, kNone, 0};
    if (a.units == 0 || ref_kind(a.top_ref) == 3u) return false;
    const F3 inv = inv3(d);
    w.o = o, w.d = d;
    if (ray_aabb(o, inv, {a.top_box[0], a.top_box[1], a.top_box[2]}, {a.top_box[3], a.top_box[4], a.top_box[5]},
                 cull_t(w.h.t)) == kInf)
        return false;
    st.sp = 0;
    w.node = a.top_ref, w.rec = kNone, w.first = 0, w.base = 0, w.guard = 0;
    w.ro = o, w.rd = d, w.rinv = inv;
    return true;
}

// Moves past a finished node or leaf: kWalkLeave when the stack is back at the unit's entry depth, else the next
// reference from the stack, or the end of the walk.
__device__ __forceinline__ void wide_pop(WideWalk &w, Stack &st) {
    if (w.rec != kNone && st.sp == w.base)
        w.node = kWalkLeave;
    else
        w.node = st.sp == 0 ? kWalkDone : st.pop();
}

// One node-phase step (wide_at_node holds): the node's four children tested and the nearer ones queued.
template <bool kCount>
__device__ __forceinline__ void wide_node(const TraceArgs &a, WideWalk &w, Stack &st, uint32_t &flags, Reads &cnt) {
    const uint32_t node = w.node;
    if (++w.guard > kGuardSteps) {
        flags |= kFlagGuard;
        w.node = kWalkDone;
        return;
    }
    if (kCount) cnt.nodes++;
    if (node >= a.wnode_count) {
        flags |= kFlagWalk;
        report_range(a.guard_flag, 7, node, a.wnode_count);
        w.node = kWalkDone;
        return;
    }
        const uint4 *p = a.wnodes + 4 * uint64_t(node);
        const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3];
        const float sx = __uint_as_float((w0.w & 0xffu) << 23), sy = __uint_as_float(((w0.w >> 8) & 0xffu) << 23),
                    sz = __uint_as_float(((w0.w >> 16) & 0xffu) << 23);
        const F3 ro = w.ro, rinv = w.rinv;
        const float bx = (__uint_as_float(w0.x) - ro.x) * rinv.x, by = (__uint_as_float(w0.y) - ro.y) * rinv.y,
                    bz = (__uint_as_float(w0.z) - ro.z) * rinv.z;
        const float ix = sx * rinv.x, iy = sy * rinv.y, iz = sz * rinv.z;
        const uint32_t nx = rinv.x >= 0.0f ? w1.x : w1.w, fx = rinv.x >= 0.0f ? w1.w : w1.x;
        const uint32_t ny = rinv.y >= 0.0f ? w1.y : w2.x, fy = rinv.y >= 0.0f ? w2.x : w1.y;
        const uint32_t nz = rinv.z >= 0.0f ? w1.z : w2.y, fz = rinv.z >= 0.0f ? w2.y : w1.z;
        const float cull = cull_t(w.h.t);
        uint32_t ref[4] = {w2.z, w2.w, w3.x, w3.y}, key[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const float tnx = fmaf(qbyte(nx, c), ix, bx), tny = fmaf(qbyte(ny, c), iy, by),
                        tnz = fmaf(qbyte(nz, c), iz, bz);
            const float tfx = fmaf(qbyte(fx, c), ix, bx), tfy = fmaf(qbyte(fy, c), iy, by),
                        tfz = fmaf(qbyte(fz, c), iz, bz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, cull));
            // an empty slot holds kNone; no reference is of kind 3, so a walk never queues one
            key[c] = tn <= tf && ref_kind(ref[c]) != 3u ? ((__float_as_uint(tn) & ~3u) | uint32_t(c)) : kNone;
        }
        // nearest first; each reference moves with its key (selects only: a lookup by index compiles to branches)
#define TRT_CSWAP(i, j)                                                          \
    {                                                                            \
        const bool s_ = key[j] < key[i];                                         \
        const uint32_t ki_ = s_ ? key[j] : key[i], kj_ = s_ ? key[i] : key[j];   \
        const uint32_t ri_ = s_ ? ref[j] : ref[i], rj_ = s_ ? ref[i] : ref[j];   \
        key[i] = ki_, key[j] = kj_, ref[i] = ri_, ref[j] = rj_;                  \
    }
        TRT_CSWAP(0, 1);
        TRT_CSWAP(2, 3);
        TRT_CSWAP(0, 2);
        TRT_CSWAP(1, 3);
        TRT_CSWAP(1, 2);
#undef TRT_CSWAP
        if (st.sp + 3 <= kLdsStack) {
            // The common case, with no branch: the three push slots and the pop slot are all in LDS, so the lane
            // writes its three candidates (counting only the queued ones) and reads the top entry whether it then
            // moves to its nearest child or pops. (As two branches, the lanes that push and the lanes that pop ran
            // one after the other: 2.5 % of the 3M-triangle batch's time.)
            const bool any = key[0] != kNone;
            const uint32_t np = uint32_t(key[1] != kNone) + uint32_t(key[2] != kNone) + uint32_t(key[3] != kNone);
            uint32_t s0 = ref[1];
            s0 = np == 2 ? ref[2] : s0;
            s0 = np == 3 ? ref[3] : s0;
            const uint32_t s1 = np == 3 ? ref[2] : ref[1];
            const uint32_t sp = st.sp;
            TRT_LDS uint32_t *top = st.lds + sp * kTraceBlock;
            const uint32_t popped = st.lds[(sp > 0 ? sp - 1 : 0) * kTraceBlock];
            top[0] = s0;
            top[kTraceBlock] = s1;
            top[2 * kTraceBlock] = ref[1];
            // the pop (wide_pop): leave the unit at its entry depth, end at an empty stack, else the top entry
            const bool leave = w.rec != kNone && sp == w.base;
            const uint32_t next = leave ? kWalkLeave : (sp == 0 ? kWalkDone : popped);
            w.node = any ? ref[0] : next;
            st.sp = any ? sp + np : ((leave || sp == 0) ? sp : sp - 1);
            return;
        }
        if (key[0] != kNone) {
            // near the LDS part's end: the children to queue are ref[1..np]; they go on the stack farthest first, so
            // the second nearest is on top
            const uint32_t np = uint32_t(key[1] != kNone) + uint32_t(key[2] != kNone) + uint32_t(key[3] != kNone);
            if (st.sp + np > st.cap) {
                flags |= kFlagWalk;
                report_range(a.guard_flag, 8, st.sp + np, st.cap);
                w.node = kWalkDone;
                return;
            }
            if (key[3] != kNone) st.push(ref[3]);
            if (key[2] != kNone) st.push(ref[2]);
            if (key[1] != kNone) st.push(ref[1]);
            w.node = ref[0];
            return;
        }
    wide_pop(w, st);
}

// One step of the other phase (wide_waits holds): a run of triangles tested, a unit entered or left. Shadow and
// any-hit walks end at the first hit they take.
// kWatertight: triangles by intersect_tri_wt (TRT_WALK_WATERTIGHT) instead of the reference's Moller-Trumbore.
template <bool kCount, bool kWatertight = false>
__device__ __forceinline__ void wide_other(const TraceArgs &a, WideWalk &w, Stack &st, uint32_t &flags, Reads &cnt) {
    const uint32_t node = w.node;
    if (++w.guard > kGuardSteps) {
        flags |= kFlagGuard;
        w.node = kWalkDone;
        return;
    }
    if (node == kWalkLeave) {
        w.rec = kNone;
        if (st.sp == 0) {  // nothing left to walk: the world ray is not needed again
            w.node = kWalkDone;
            return;
        }
        w.ro = w.o, w.rd = w.d, w.rinv = inv3(w.d);  // the same bits as at the start: kept no register for it
        w.node = st.pop();
        return;
    }
    if (ref_kind(node) == kRefUnit) {
        const uint32_t r = node & 0x3fffffffu;
        if (kCount) cnt.units++;
        if (r >= a.wunit_count) {
            flags |= kFlagWalk;
            report_range(a.guard_flag, 9, r, a.wunit_count);
            w.node = kWalkDone;
            return;
        }
        const float4 *u = a.wunits + 4 * uint64_t(r);
        const float4 u0 = u[0];
        const uint32_t uflags = __float_as_uint(u0.w);
        if (w.mode != kTraceShadow || (uflags & kUnitShadow) != 0) {
            const float4 u1 = u[1], u2 = u[2], u3 = u[3];
            const uint32_t root = __float_as_uint(u1.w);
            if (ref_kind(root) == 3u) {
                flags |= kFlagWalk;
                report_range(a.guard_flag, 10, root, a.wnode_count);
                w.node = kWalkDone;
                return;
            }
            w.rec = r;
            w.ro = {w.o.x - u0.x, w.o.y - u0.y, w.o.z - u0.z};
            if ((uflags & kUnitShifted) == 0) {
                const F3 c0{u1.x, u1.y, u1.z}, c1{u2.x, u2.y, u2.z}, c2{u3.x, u3.y, u3.z};
                w.ro = proc_dir(c0, c1, c2, w.ro);
                w.rd = proc_dir(c0, c1, c2, w.d);
                w.rinv = inv3(w.rd);
            }
            w.first = __float_as_uint(u2.w);
            w.base = st.sp;
            w.node = root;
            return;
        }
        wide_pop(w, st);
        return;
    }
    const uint32_t count = ((node >> kRefTriFirstBits) & 15u) + 1u;
    const uint64_t f = uint64_t(w.first) + (node & kRefTriFirstMask);
    if (ref_kind(node) != kRefTris || f + count > a.leaf_count) {
        flags |= kFlagWalk;
        report_range(a.guard_flag, 10, f + count, a.leaf_count);
        w.node = kWalkDone;
        return;
    }
    bool taken = false;
    Shear sh{};
    if (kWatertight) sh = shear_of(w.rd);
    for (uint32_t k = 0; k < count; k++) {
        if (kCount) cnt.tris++;
        const float4 *l = a.wleaves + 3 * (f + k);
        const float4 lv0 = l[0], lv1 = l[1], lv2 = l[2];
        const uint32_t ti = __float_as_uint(lv2.w);  // the reference leaf index: what hits report and ties order
        float x, uu, vv;
        if (kWatertight)
            intersect_tri_wt(w.ro, sh, lv0, lv1, lv2, w.tmax, x, uu, vv);
        else
            intersect_tri(w.ro, w.rd, lv0, lv1, lv2, w.tmax, x, uu, vv);
        const bool take = x < w.tmax && accept(w.h, x, w.rec, ti);
        take_hit(w.h, take, x, ti, uu, vv, w.rec, __float_as_uint(lv0.w));
        taken = taken || take;
    }
    if (taken && w.mode != kTraceClosest)
        w.node = kWalkDone;
    else
        wide_pop(w, st);
}

template <bool kCount>
__device__ HitRec trace_wide(const TraceArgs &a, F3 o, F3 d, float tmax_in, uint32_t mode, Stack &st,
                             uint32_t &flags, Reads &cnt) {
    WideWalk w;
    if (!wide_begin(a, o, d, tmax_in, mode, w, st)) return w.h;
    // The lanes in this call walk their rays together (see WideWalk); a lane leaves when its walk is over.
    for (;;) {
        const uint32_t walking = __popcll(__ballot(true));
        for (;;) {
            const bool at_node = wide_at_node(w);
            if (__ballot(at_node) == 0 || kLeafShareCall * __popcll(__ballot(wide_waits(w))) >= walking) break;
            if (at_node) wide_node<kCount>(a, w, st, flags, cnt);
        }
        if (wide_waits(w)) wide_other<kCount>(a, w, st, flags, cnt);
        if (wide_over(w)) break;
    }
    return w.h;
}

}  // namespace walk
}  // namespace trt
