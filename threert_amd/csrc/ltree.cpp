// The light tree over a scene's emitter triangles (ltree.rs:409-459 build / build_with without moving emitters or
// the slot for emitters added in place, grow 355-403, encode 118-131, box_sphere 133-138): a binary tree whose nodes
// bound their emitters with a half-float sphere and sum their power, descended by importance in the path tracer
// (pathtrace.wgsl:504-618). Host code, as in the reference's `prepare` (lib.rs:256-264 light_tree_of).
//
// One difference in order, not in shape: the reference splits each node at the median of its centroids along the
// widest axis with `select_nth_unstable_by`, which leaves each side in an order of its own; here each side is sorted
// by the same key (stable), so a node's power sums its emitters in that order.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/threert.h"

int trt_fail(int code, const std::string &msg);
int trt_ok();

namespace trt {
namespace {

constexpr uint32_t kLeaf = 0x80000000u;       // ltree.rs:61 LEAF
constexpr uint32_t kPathBits = 26;            // ltree.rs:63 PATH_BITS
constexpr size_t kMinStatic = 256;            // ltree.rs:31 MIN_STATIC
constexpr size_t kLeafMax = 1;                // ltree.rs:34 LEAF_MAX

uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float from_bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

// IEEE binary16 from binary32, rounding to nearest even (the `half` crate's f16::from_f32): overflow to infinity,
// subnormals rounded, a NaN kept a quiet NaN with its sign.
uint16_t f32_to_f16(float f) {
    const uint32_t x = bits(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return static_cast<uint16_t>(sign | 0x7c00u | (man ? 0x200u | (man >> 13) : 0u));
    const int e = static_cast<int>(exp) - 127 + 15;
    if (e >= 0x1f) return static_cast<uint16_t>(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return static_cast<uint16_t>(sign);
        man |= 0x800000u;
        const uint32_t shift = static_cast<uint32_t>(14 - e);
        uint32_t h = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1u), mid = 1u << (shift - 1u);
        if (rem > mid || (rem == mid && (h & 1u))) h++;
        return static_cast<uint16_t>(sign | h);
    }
    uint32_t h = sign | (static_cast<uint32_t>(e) << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // a carry into the exponent is the right rounding
    return static_cast<uint16_t>(h);
}

// binary16 to binary32, exactly.
float f16_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    if (exp == 0x1fu) return from_bits(sign | 0x7f800000u | (man << 13));
    if (exp == 0) {
        if (man == 0) return from_bits(sign);
        int e = -14;
        while (!(man & 0x400u)) man <<= 1, e--;
        man &= 0x3ffu;
        return from_bits(sign | uint32_t(e + 127) << 23 | (man << 13));
    }
    return from_bits(sign | (exp - 15 + 127) << 23 | (man << 13));
}

namespace {

struct Item {
    uint32_t e;
    float lo[3], hi[3], c[3], w;
};

// ltree.rs:118-131 encode: the centre rounded to half floats; the radius grown by that rounding and rounded up.
void encode(const float c[3], float r, float out[2]) {
    float hc[3];
    for (int k = 0; k < 3; k++) hc[k] = f16_to_f32(f32_to_f16(c[k]));
    const float d0 = hc[0] - c[0], d1 = hc[1] - c[1], d2 = hc[2] - c[2];
    const float grow = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const float rg = r + grow;
    uint16_t hr = f32_to_f16(rg);
    if (f16_to_f32(hr) < rg) hr = static_cast<uint16_t>(hr + 1);
    out[0] = from_bits(uint32_t(f32_to_f16(hc[0])) | uint32_t(f32_to_f16(hc[1])) << 16);
    out[1] = from_bits(uint32_t(f32_to_f16(hc[2])) | uint32_t(f32_to_f16(f16_to_f32(hr))) << 16);
}

// ltree.rs:133-138 box_sphere
void box_sphere(const float lo[3], const float hi[3], float out[2]) {
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = 0.5f * (lo[k] + hi[k]);
    const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    encode(c, 0.5f * std::sqrt(x * x + y * y + z * z), out);
}

// f32::total_cmp as an integer key
int32_t total_key(float f) {
    int32_t k;
    std::memcpy(&k, &f, 4);
    return k ^ static_cast<int32_t>(static_cast<uint32_t>(k >> 31) >> 1);
}

struct Tree {
    std::vector<float> nodes;  // 4 a node
    std::vector<uint32_t> order, code;
};

// ltree.rs:355-403 grow, from the root (node 0) at depth 0, path 0.
bool grow(std::vector<Item> &items, Tree &t, std::string *why) {
    struct Open {
        size_t a, b, node;
        uint32_t depth, path;
    };
    t.nodes.assign(4, 0.0f);
    std::vector<Open> stack{{0, items.size(), 0, 0, 0}};
    while (!stack.empty()) {
        const Open o = stack.back();
        stack.pop_back();
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX}, w = 0.0f;
        for (size_t i = o.a; i < o.b; i++) {
            for (int k = 0; k < 3; k++) {
                lo[k] = std::fmin(lo[k], items[i].lo[k]);
                hi[k] = std::fmax(hi[k], items[i].hi[k]);
            }
            w += items[i].w;
        }
        float s[2];
        box_sphere(lo, hi, s);
        float *n = &t.nodes[4 * o.node];
        n[0] = s[0], n[1] = s[1], n[2] = w, n[3] = 0.0f;
        const size_t len = o.b - o.a;
        if (len <= kLeafMax) {
            if (o.depth > kPathBits) {
                *why = "light tree deeper than 26 levels";
                return false;
            }
            n[3] = from_bits(kLeaf | uint32_t(len - 1) << 29 | uint32_t(o.a));
            for (size_t k = 0; k < len; k++) {
                t.order[o.a + k] = items[o.a + k].e;
                t.code[o.a + k] = (o.depth + 1) << kPathBits | o.path;
            }
            continue;
        }
        // the widest axis of the centroids (the last of equal widths, as Iterator::max_by), median split
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t i = o.a; i < o.b; i++)
            for (int k = 0; k < 3; k++) {
                clo[k] = std::fmin(clo[k], items[i].c[k]);
                chi[k] = std::fmax(chi[k], items[i].c[k]);
            }
        int axis = 0;
        for (int k = 1; k < 3; k++)
            if (total_key(chi[k] - clo[k]) >= total_key(chi[axis] - clo[axis])) axis = k;
        const size_t mid = len / 2;
        std::stable_sort(items.begin() + static_cast<long>(o.a), items.begin() + static_cast<long>(o.b),
                         [axis](const Item &p, const Item &q) { return total_key(p.c[axis]) < total_key(q.c[axis]); });
        const size_t left = t.nodes.size() / 4;
        t.nodes.resize(t.nodes.size() + 8, 0.0f);
        t.nodes[4 * o.node + 3] = from_bits(static_cast<uint32_t>(left));
        stack.push_back({o.a, o.a + mid, left, o.depth + 1, o.path});
        stack.push_back({o.a + mid, o.b, left + 1, o.depth + 1, o.path | 1u << o.depth});
    }
    return true;
}

}  // namespace
}  // namespace trt

extern "C" {

int trt_build_light_tree(const float *verts, const float *centroids, const float *weights, uint64_t n,
                         trt_light_tree *out) {
    if (!out || (n && (!verts || !centroids || !weights)))
        return trt_fail(TRT_E_INVALID, "trt_build_light_tree: null argument");
    std::memset(out, 0, sizeof(*out));
    if (n >= (uint64_t{1} << 29)) return trt_fail(TRT_E_INVALID, "trt_build_light_tree: entries beyond the child word");
    std::vector<trt::Item> items;
    for (uint64_t i = 0; i < n; i++) {
        const float w = weights[i];
        if (!(w > 0.0f)) continue;
        trt::Item it{};
        it.e = static_cast<uint32_t>(i);
        for (int k = 0; k < 3; k++) {
            it.lo[k] = FLT_MAX, it.hi[k] = -FLT_MAX;
            for (int j = 0; j < 3; j++) {
                it.lo[k] = std::fmin(it.lo[k], verts[9 * i + 3 * j + k]);
                it.hi[k] = std::fmax(it.hi[k], verts[9 * i + 3 * j + k]);
            }
            it.c[k] = centroids[3 * i + k];
        }
        it.w = w;
        items.push_back(it);
    }
    if (items.size() < trt::kMinStatic) return trt_ok();  // the CDF alone (ltree.rs:431-433)
    trt::Tree t;
    t.order.assign(items.size(), 0);
    t.code.assign(items.size(), 0);
    std::string why;
    if (!trt::grow(items, t, &why)) return trt_fail(TRT_E_INVALID, "trt_build_light_tree: " + why);
    const uint64_t nodes = t.nodes.size() / 4;
    out->nodes = static_cast<float *>(std::malloc(t.nodes.size() * sizeof(float)));
    out->order = static_cast<uint32_t *>(std::malloc(t.order.size() * sizeof(uint32_t)));
    out->code = static_cast<uint32_t *>(std::malloc(t.code.size() * sizeof(uint32_t)));
    if (!out->nodes || !out->order || !out->code) {
        trt_light_tree_free(out);
        return trt_fail(TRT_E_OOM, "trt_build_light_tree: out of host memory");
    }
    std::memcpy(out->nodes, t.nodes.data(), t.nodes.size() * sizeof(float));
    std::memcpy(out->order, t.order.data(), t.order.size() * sizeof(uint32_t));
    std::memcpy(out->code, t.code.data(), t.code.size() * sizeof(uint32_t));
    out->node_count = nodes;
    out->entry_count = t.order.size();
    out->power = t.nodes[2];
    return trt_ok();
}

void trt_light_tree_free(trt_light_tree *t) {
    if (!t) return;
    std::free(t->nodes);
    std::free(t->order);
    std::free(t->code);
    std::memset(t, 0, sizeof(*t));
}

}  // extern "C"
