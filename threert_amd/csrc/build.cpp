// Host tree builders: the reference's prepare-side BVH construction, restated for this library.
//   bvh.rs:81-239   build_tris / build_from   binned SAH, 16 bins, leaves of <= 4 (<= 16 when SAH prefers a leaf),
//                                             depth cap with a median fallback, non-finite items last
//   units.rs:84-118 build_median              median splits along the widest centre axis
//   units.rs:140-223 build_sah                32-bin SAH over unit boxes, one box a leaf, depth-capped by room
//   units.rs:231-246 mesh_tree, 288-297 UnitBox::new, 303-321 top_tree, 329-342 unit_record
// Lane arithmetic follows glam's SSE2 Vec3A (min/max keep the second operand on NaN) so the trees come out the
// same bytes as the reference's. Where the reference leaves an order unspecified (select_nth_unstable), ties are
// broken by item index.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "trt_internal.h"

namespace trt {
namespace {

struct V3 {
    float x, y, z;
    float operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
inline float lmin(float a, float b) { return a < b ? a : b; }
inline float lmax(float a, float b) { return a > b ? a : b; }
inline V3 vmin(V3 a, V3 b) { return {lmin(a.x, b.x), lmin(a.y, b.y), lmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {lmax(a.x, b.x), lmax(a.y, b.y), lmax(a.z, b.z)}; }
inline bool vfinite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
constexpr float kBig = INFINITY;  // (not the kernel INF)

struct Box {
    V3 lo{kBig, kBig, kBig}, hi{-kBig, -kBig, -kBig};
    void grow(V3 p) { lo = vmin(lo, p), hi = vmax(hi, p); }
    void add(const Box &b) { lo = vmin(lo, b.lo), hi = vmax(hi, b.hi); }
    bool finite() const {
        return vfinite(lo) && vfinite(hi) && lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z;
    }
};
inline float surface(V3 lo, V3 hi) {
    float ex = lmax(hi.x - lo.x, 0.0f), ey = lmax(hi.y - lo.y, 0.0f), ez = lmax(hi.z - lo.z, 0.0f);
    return 2.0f * (ex * ey + ey * ez + ez * ex);
}
inline float surface(const Box &b) { return surface(b.lo, b.hi); }

// Rust `f32 as usize`.
inline uint64_t to_index(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return static_cast<uint64_t>(f);
}
inline int32_t order_key(float f) {
    int32_t b;
    std::memcpy(&b, &f, 4);
    return b ^ static_cast<int32_t>(static_cast<uint32_t>(b >> 31) >> 1);
}
inline uint32_t depth_for(uint64_t count) {  // bvh.rs:119-121 levels
    uint64_t q = (count + 3) / 4, p = 1;
    uint32_t l = 0;
    while (p < q) p <<= 1, l++;
    return l;
}

Node make_node(const Box &b, uint32_t first_or_left, uint32_t count) {
    Node n;
    n.bmin[0] = b.lo.x, n.bmin[1] = b.lo.y, n.bmin[2] = b.lo.z;
    n.bmax[0] = b.hi.x, n.bmax[1] = b.hi.y, n.bmax[2] = b.hi.z;
    n.left_or_first = first_or_left;
    n.count = count;
    return n;
}
const Node kZero = {{0, 0, 0}, 0, {0, 0, 0}, 0};
const Node kFill = {{0, 0, 0}, 0, {0, 0, 0}, 1};

void median_order(uint32_t *items, uint64_t n, const std::vector<V3> &keys, int axis) {
    std::sort(items, items + n, [&](uint32_t a, uint32_t b) {
        int32_t ka = order_key(keys[a][axis]), kb = order_key(keys[b][axis]);
        return ka != kb ? ka < kb : a < b;
    });
}

struct Range {
    uint32_t node;
    uint64_t start, end;
    uint32_t depth;
};

// bvh.rs:130-239
bool sah_tris(const std::vector<Box> &boxes, const std::vector<V3> &cent, uint32_t max_depth, Tree &out) {
    constexpr int kBins = 16;
    const uint64_t total = boxes.size();
    out.order.clear();
    out.order.reserve(total);
    for (uint64_t i = 0; i < total; i++)
        if (boxes[i].finite()) out.order.push_back(static_cast<uint32_t>(i));
    const uint64_t n = out.order.size();
    for (uint64_t i = 0; i < total; i++)
        if (!boxes[i].finite()) out.order.push_back(static_cast<uint32_t>(i));
    out.nodes.assign(1, kZero);
    out.depth = 0;
    if (n == 0) return true;
    if (depth_for(n) > max_depth) return false;
    std::vector<Range> todo{{0, 0, n, 0}};
    uint32_t deepest = 0;
    uint32_t *order = out.order.data();
    while (!todo.empty()) {
        Range r = todo.back();
        todo.pop_back();
        deepest = std::max(deepest, r.depth);
        Box bounds, cb;
        for (uint64_t q = r.start; q < r.end; q++) bounds.add(boxes[order[q]]), cb.grow(cent[order[q]]);
        const uint64_t count = r.end - r.start;
        if (count <= 4) {
            out.nodes[r.node] = make_node(bounds, static_cast<uint32_t>(r.start), static_cast<uint32_t>(count));
            continue;
        }
        const float ext[3] = {cb.hi.x - cb.lo.x, cb.hi.y - cb.lo.y, cb.hi.z - cb.lo.z};
        const float cmin[3] = {cb.lo.x, cb.lo.y, cb.lo.z};
        float best = kBig;
        int best_axis = 0;
        int best_split = 0;
        for (int a = 0; a < 3; a++) {
            if (ext[a] <= 1e-9f) continue;
            Box bb[kBins];
            uint64_t bn[kBins] = {};
            const float scale = static_cast<float>(kBins) / ext[a];
            for (uint64_t q = r.start; q < r.end; q++) {
                const uint32_t t = order[q];
                const uint64_t b = std::min<uint64_t>(to_index((cent[t][a] - cmin[a]) * scale), kBins - 1);
                bn[b]++;
                bb[b].add(boxes[t]);
            }
            float la[kBins];
            uint64_t lc[kBins];
            Box acc;
            uint64_t c = 0;
            for (int i = 0; i < kBins; i++) acc.add(bb[i]), c += bn[i], la[i] = surface(acc), lc[i] = c;
            acc = Box();
            c = 0;
            for (int i = kBins - 1; i >= 1; i--) {
                acc.add(bb[i]);
                c += bn[i];
                const float cost = la[i - 1] * static_cast<float>(lc[i - 1]) + surface(acc) * static_cast<float>(c);
                const bool fits = r.depth + 1 + depth_for(std::max(lc[i - 1], c)) <= max_depth;
                if (lc[i - 1] > 0 && c > 0 && cost < best && fits) best = cost, best_axis = a, best_split = i;
            }
        }
        const float leaf_cost = surface(bounds) * static_cast<float>(count);
        const bool splittable = std::isfinite(best);
        if (count <= 16 && (!splittable || best >= leaf_cost)) {
            out.nodes[r.node] = make_node(bounds, static_cast<uint32_t>(r.start), static_cast<uint32_t>(count));
            continue;
        }
        uint64_t mid;
        if (splittable) {
            const float scale = static_cast<float>(kBins) / ext[best_axis];
            uint64_t i = r.start;
            for (uint64_t q = r.start; q < r.end; q++) {
                const uint64_t b =
                    std::min<uint64_t>(to_index((cent[order[q]][best_axis] - cmin[best_axis]) * scale), kBins - 1);
                if (b < static_cast<uint64_t>(best_split)) std::swap(order[i++], order[q]);
            }
            mid = i;
        } else if (ext[0] <= 1e-9f && ext[1] <= 1e-9f && ext[2] <= 1e-9f) {
            mid = r.start + count / 2;
        } else {
            int axis = 0;
            for (int k = 1; k < 3; k++)
                if (order_key(ext[axis]) <= order_key(ext[k])) axis = k;
            median_order(order + r.start, count, cent, axis);
            mid = r.start + count / 2;
        }
        const uint32_t left = static_cast<uint32_t>(out.nodes.size());
        out.nodes.push_back(kZero);
        out.nodes.push_back(kZero);
        out.nodes[r.node] = make_node(bounds, left, 0);
        todo.push_back({left + 1, mid, r.end, r.depth + 1});
        todo.push_back({left, r.start, mid, r.depth + 1});
    }
    out.depth = deepest;
    return true;
}

void boxes_of(const float *b, uint64_t n, std::vector<Box> &bx, std::vector<V3> &ce) {
    bx.resize(n);
    ce.resize(n);
    for (uint64_t i = 0; i < n; i++) {
        bx[i].lo = {b[6 * i], b[6 * i + 1], b[6 * i + 2]};
        bx[i].hi = {b[6 * i + 3], b[6 * i + 4], b[6 * i + 5]};
        ce[i] = {(bx[i].lo.x + bx[i].hi.x) * 0.5f, (bx[i].lo.y + bx[i].hi.y) * 0.5f,
                 (bx[i].lo.z + bx[i].hi.z) * 0.5f};
    }
}

inline int widest(const float *e) { return (e[0] >= e[1] && e[0] >= e[2]) ? 0 : (e[1] >= e[2] ? 1 : 2); }

}  // namespace

// bvh.rs:81-106
bool build_tris(const float *tris, uint64_t n, uint32_t max_depth, Tree &out) {
    out = Tree();
    if (n == 0) {
        out.nodes.assign(1, kFill);
        return true;
    }
    std::vector<Box> boxes(n);
    std::vector<V3> cent(n);
    for (uint64_t i = 0; i < n; i++) {
        const float *t = tris + 9 * i;
        const V3 p0{t[0], t[1], t[2]}, p1{t[3], t[4], t[5]}, p2{t[6], t[7], t[8]};
        if (vfinite(p0) && vfinite(p1) && vfinite(p2)) boxes[i].grow(p0), boxes[i].grow(p1), boxes[i].grow(p2);
        cent[i] = {(p0.x + p1.x + p2.x) / 3.0f, (p0.y + p1.y + p2.y) / 3.0f, (p0.z + p1.z + p2.z) / 3.0f};
    }
    if (!sah_tris(boxes, cent, max_depth, out)) return false;
    if (out.nodes.size() == 1 && out.nodes[0].count == 0) out.nodes[0] = kFill;
    return true;
}

// units.rs:84-118
void build_median(const float *b, uint64_t n, uint32_t max_leaf, Tree &out) {
    out = Tree();
    out.order.resize(n);
    for (uint64_t i = 0; i < n; i++) out.order[i] = static_cast<uint32_t>(i);
    out.nodes.assign(1, kFill);
    if (n == 0) return;
    std::vector<Box> bx;
    std::vector<V3> ce;
    boxes_of(b, n, bx, ce);
    std::vector<Range> todo{{0, 0, n, 0}};
    uint32_t *order = out.order.data();
    while (!todo.empty()) {
        Range r = todo.back();
        todo.pop_back();
        Box all, cb;
        for (uint64_t q = r.start; q < r.end; q++) all.add(bx[order[q]]), cb.grow(ce[order[q]]);
        const uint64_t count = r.end - r.start;
        if (count <= max_leaf) {
            out.nodes[r.node] = make_node(all, static_cast<uint32_t>(r.start), static_cast<uint32_t>(count));
            continue;
        }
        const float e[3] = {cb.hi.x - cb.lo.x, cb.hi.y - cb.lo.y, cb.hi.z - cb.lo.z};
        median_order(order + r.start, count, ce, widest(e));
        const uint64_t mid = count / 2;
        const uint32_t left = static_cast<uint32_t>(out.nodes.size());
        out.nodes.push_back(kFill);
        out.nodes.push_back(kFill);
        out.nodes[r.node] = make_node(all, left, 0);
        todo.push_back({left + 1, r.start + mid, r.end, 0});
        todo.push_back({left, r.start, r.start + mid, 0});
    }
}

// units.rs:140-223
void build_sah(const float *b, uint64_t n, uint32_t max_depth, Tree &out) {
    constexpr int kBins = 32;
    out = Tree();
    out.order.resize(n);
    for (uint64_t i = 0; i < n; i++) out.order[i] = static_cast<uint32_t>(i);
    out.nodes.assign(1, kFill);
    if (n == 0) return;
    std::vector<Box> bx;
    std::vector<V3> ce;
    boxes_of(b, n, bx, ce);
    std::vector<Range> todo{{0, 0, n, 0}};
    uint32_t *order = out.order.data();
    uint32_t deepest = 0;
    while (!todo.empty()) {
        Range r = todo.back();
        todo.pop_back();
        deepest = std::max(deepest, r.depth);
        Box all, cb;
        for (uint64_t q = r.start; q < r.end; q++) all.add(bx[order[q]]), cb.grow(ce[order[q]]);
        const uint64_t count = r.end - r.start;
        if (count == 1) {
            out.nodes[r.node] = make_node(all, static_cast<uint32_t>(r.start), 1);
            continue;
        }
        const uint64_t shift = static_cast<uint64_t>(max_depth) - r.depth - 1;
        const uint64_t room = shift >= 64 ? UINT64_MAX : (uint64_t{1} << shift);
        const float e[3] = {cb.hi.x - cb.lo.x, cb.hi.y - cb.lo.y, cb.hi.z - cb.lo.z};
        const float cmin[3] = {cb.lo.x, cb.lo.y, cb.lo.z};
        auto bin = [&](uint32_t k, int a) {
            return std::min<uint64_t>(to_index((ce[k][a] - cmin[a]) * (static_cast<float>(kBins) / e[a])), kBins - 1);
        };
        float best = kBig;
        int best_axis = 0, best_split = 0;
        for (int a = 0; a < 3; a++) {
            if (!(e[a] > 0.0f)) continue;
            Box bb[kBins];
            uint64_t bn[kBins] = {};
            for (uint64_t q = r.start; q < r.end; q++) {
                const uint64_t k = bin(order[q], a);
                bb[k].add(bx[order[q]]);
                bn[k]++;
            }
            float right[kBins] = {};
            Box acc;
            for (int i = kBins - 1; i >= 1; i--) acc.add(bb[i]), right[i] = surface(acc);
            acc = Box();
            uint64_t nl = 0;
            for (int i = 1; i < kBins; i++) {
                acc.add(bb[i - 1]);
                nl += bn[i - 1];
                const uint64_t nr = count - nl;
                if (nl == 0 || nr == 0 || nl > room || nr > room) continue;
                const float cost = surface(acc) * static_cast<float>(nl) + right[i] * static_cast<float>(nr);
                if (cost < best) best = cost, best_axis = a, best_split = i;
            }
        }
        uint64_t mid;
        if (std::isfinite(best)) {
            uint64_t i = r.start;
            for (uint64_t q = r.start; q < r.end; q++)
                if (bin(order[q], best_axis) < static_cast<uint64_t>(best_split)) std::swap(order[i++], order[q]);
            mid = i - r.start;
        } else {
            median_order(order + r.start, count, ce, widest(e));
            mid = count / 2;
        }
        const uint32_t left = static_cast<uint32_t>(out.nodes.size());
        out.nodes.push_back(kFill);
        out.nodes.push_back(kFill);
        out.nodes[r.node] = make_node(all, left, 0);
        todo.push_back({left + 1, r.start + mid, r.end, r.depth + 1});
        todo.push_back({left, r.start, r.start + mid, r.depth + 1});
    }
    out.depth = deepest;
}

// bvh.rs:246-282 check_tree, with the leaf rule `count <= max_leaf`, `first + count <= items`; records: when not
// null, leaves must be single, distinct records below `records_n` (units.rs:68-79 check_top). Returns the depth or -1.
int check_tree(const Node *nodes, uint64_t count, uint32_t max_depth, uint64_t items, uint32_t max_leaf,
               uint64_t records_n, std::string *why) {
    if (count == 0) {
        if (why) *why = "a tree without nodes";
        return -1;
    }
    std::vector<uint8_t> seen(count, 0);
    std::vector<uint8_t> rec_seen(records_n ? records_n : 0, 0);
    std::vector<std::pair<uint64_t, uint32_t>> st{{0, 0}};
    uint64_t reached = 0;
    uint32_t deepest = 0;
    auto fail = [&](const std::string &s) {
        if (why) *why = s;
        return -1;
    };
    while (!st.empty()) {
        auto [i, depth] = st.back();
        st.pop_back();
        if (seen[i]) return fail("node " + std::to_string(i) + " is reached twice");
        seen[i] = 1;
        reached++;
        deepest = std::max(deepest, depth);
        const Node &n = nodes[i];
        if (n.count > 0) {
            if (records_n) {
                if (n.count != 1 || n.left_or_first >= records_n) return fail("a top leaf is not one unit");
                if (rec_seen[n.left_or_first]) return fail("record " + std::to_string(n.left_or_first) + " twice");
                rec_seen[n.left_or_first] = 1;
            } else if (n.count > max_leaf || uint64_t{n.left_or_first} + n.count > items) {
                return fail("leaf " + std::to_string(i) + " names items outside the mesh");
            }
            continue;
        }
        for (int k = 0; k < 3; k++)
            if (!(std::isfinite(n.bmin[k]) && std::isfinite(n.bmax[k]) && n.bmin[k] <= n.bmax[k]))
                return fail("node " + std::to_string(i) + "'s box is not finite");
        const uint64_t l = n.left_or_first;
        if (l <= i || l + 1 >= count) return fail("node " + std::to_string(i) + "'s children are outside the tree");
        if (depth + 1 > max_depth) return fail("deeper than " + std::to_string(max_depth));
        st.push_back({l, depth + 1});
        st.push_back({l + 1, depth + 1});
    }
    if (reached != count) return fail(std::to_string(count - reached) + " nodes are not reached");
    return static_cast<int>(deepest);
}

// units.rs:231-246
bool mesh_tree(const float *tris, uint64_t n, Tree &out, std::string *why) {
    if (n == 0) {
        if (why) *why = "a mesh without triangles has no tree";
        return false;
    }
    for (uint64_t i = 0; i < 9 * n; i++)
        if (!std::isfinite(tris[i])) {
            if (why) *why = "a mesh's vertex is not finite";
            return false;
        }
    if (!build_tris(tris, n, kBvhStack, out)) {
        if (why) *why = "too many triangles for a tree";
        return false;
    }
    int d = check_tree(out.nodes.data(), out.nodes.size(), kMeshDepth, n, kMaxLeaf, 0, nullptr);
    if (d < 0) {
        std::vector<float> boxes(6 * n);
        for (uint64_t i = 0; i < n; i++) {
            const float *t = tris + 9 * i;
            for (int k = 0; k < 3; k++) {
                boxes[6 * i + k] = lmin(lmin(t[k], t[3 + k]), t[6 + k]);
                boxes[6 * i + 3 + k] = lmax(lmax(t[k], t[3 + k]), t[6 + k]);
            }
        }
        build_median(boxes.data(), n, 4, out);
        d = check_tree(out.nodes.data(), out.nodes.size(), kMeshDepth, n, kMaxLeaf, 0, why);
        if (d < 0) return false;
    }
    out.depth = static_cast<uint32_t>(d);
    V3 lo{kBig, kBig, kBig}, hi{-kBig, -kBig, -kBig};
    for (uint64_t i = 0; i < 3 * n; i++) {
        const V3 p{tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
        lo = vmin(lo, p), hi = vmax(hi, p);
    }
    out.lo[0] = lo.x, out.lo[1] = lo.y, out.lo[2] = lo.z;
    out.hi[0] = hi.x, out.hi[1] = hi.y, out.hi[2] = hi.z;
    return true;
}

// units.rs:288-297
void unit_box(const float *w, const float *lo, const float *hi, float *out) {
    V3 a{kBig, kBig, kBig}, b{-kBig, -kBig, -kBig};
    for (int c = 0; c < 8; c++) {
        const float p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
        float r[3];
        for (int k = 0; k < 3; k++) {
            float s = w[k] * p[0];
            s = w[4 + k] * p[1] + s;
            s = w[8 + k] * p[2] + s;
            r[k] = w[12 + k] + s;
        }
        const V3 q{r[0], r[1], r[2]};
        a = vmin(a, q), b = vmax(b, q);
    }
    const float ex = lmax(lmax(b.x - a.x, b.z - a.z), b.y - a.y);
    const float ab =
        lmax(lmax(lmax(std::fabs(a.x), std::fabs(b.x)), lmax(std::fabs(a.z), std::fabs(b.z))),
             lmax(std::fabs(a.y), std::fabs(b.y)));
    const float pad = ex * 1e-4f + ab * 1e-5f + 1e-6f;
    out[0] = a.x - pad, out[1] = a.y - pad, out[2] = a.z - pad;
    out[3] = b.x + pad, out[4] = b.y + pad, out[5] = b.z + pad;
}

// units.rs:303-321
bool top_tree(const float *boxes, const uint32_t *records, uint64_t n, Tree &out, std::string *why) {
    if (n == 0 || n > kMaxUnits) {
        if (why) *why = n == 0 ? "a top tree needs a unit" : "more units than a top tree holds";
        return false;
    }
    for (uint64_t i = 0; i < 6 * n; i++)
        if (!std::isfinite(boxes[i])) {
            if (why) *why = "a unit's box is not finite";
            return false;
        }
    build_sah(boxes, n, kTopDepth - 1, out);
    uint32_t most = 0;
    for (uint64_t i = 0; i < n; i++) most = std::max(most, records[i]);
    for (auto &nd : out.nodes)
        if (nd.count > 0) nd.left_or_first = records[out.order[nd.left_or_first]];
    const int d = check_tree(out.nodes.data(), out.nodes.size(), kTopDepth, 0, 1, uint64_t{most} + 1, why);
    if (d < 0) return false;
    out.depth = static_cast<uint32_t>(d);
    return true;
}

// units.rs:329-342
void unit_record(Instance &r, uint32_t root, bool cast_shadow) {
    bool alone = true;
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) alone = alone && r.n[k][c] == (k == c ? 1.0f : 0.0f);
    r.n[0][3] = from_bits(root);
    r.n[1][3] = from_bits((to_bits(r.n[1][3]) & kInstBits) | (cast_shadow ? kUnitShadow : 0u) |
                          (alone ? kUnitShifted : 0u));
}

}  // namespace trt
