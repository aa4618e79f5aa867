// The library's own node format: 4-wide nodes of 64 bytes with 8-bit child bounds against an FP32 origin and a
// power-of-two step per axis.
//
// Mesh trees: per mesh, the library builds its own binary SAH tree over the mesh's triangles (16 bins on each axis,
// leaves of at most kOwnLeaf triangles) and lays the mesh's leaves out in that tree's order in a copy of the leaf
// buffer (`wleaves`), each carrying its reference leaf index in e2.w: the walk reports and breaks ties by that index,
// so hits are the reference's whatever the layout. (The reference's trees allow leaves of 16, which costs the
// incoherent workload about 23 triangle tests a ray.) The top tree is the reference's.
//
// Collapse: a binary interior node becomes a wide node whose children are found by opening, while fewer than four,
// the child of largest surface area that is itself interior. Binary leaves stay leaves (a run of triangles, or one
// unit in the top tree).
//
// Quantisation is conservative: origin = (node low - one step) rounded down; each child's low byte is floored and
// lowered by one step, its high byte ceiled and raised by one, checked in exact (double) arithmetic. The extra step
// covers the walk's rounding when it evaluates a plane as fma(q, step / d, (origin - o) / d).
#include <algorithm>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#define TRT_HD
#include "gbuild_core.h"
#include "trt_internal.h"

namespace trt {
namespace {

inline float node_area(const Node &n) { return gb::open_area(n); }

bool quantise(WideNode &w, const Node *const *kids, int nc) { return gb::quantise_node(w, kids, nc); }

struct Collapser {
    const Node *nodes;
    uint64_t node_count;
    WideScene &out;
    // The stack bound of a unit's mesh tree, by record (top tree leaves).
    const std::vector<uint32_t> *unit_bound = nullptr;
    std::string err;

    // Converts the binary subtree at absolute index b (children relative to `off`). Returns its reference and
    // writes the most entries a walk under it can push.
    uint32_t convert(uint32_t b, uint32_t off, bool top, uint32_t &bound) {
        const Node &n = nodes[b];
        if (n.count > 0) {
            if (top) {
                bound = (*unit_bound)[n.left_or_first];
                return ref_unit(n.left_or_first);
            }
            bound = 0;
            if (n.left_or_first > kRefTriFirstMask) err = "a mesh has more triangles than a reference holds (2^26)";
            return ref_tris(n.left_or_first, n.count);
        }
        uint32_t kid[4];
        const int nc = gb::open_kids(nodes, off, b, kid);
        const uint32_t wi = static_cast<uint32_t>(out.node_base + out.nodes.size());
        out.nodes.emplace_back();
        WideNode w{};
        const Node *kp[4];
        uint32_t deepest = 0;
        for (int c = 0; c < nc; c++) {
            uint32_t cb = 0;
            w.child[c] = convert(kid[c], off, top, cb);
            deepest = std::max(deepest, cb);
            kp[c] = &nodes[kid[c]];
        }
        for (int c = nc; c < 4; c++) w.child[c] = kNone;
        if (!quantise(w, kp, nc)) err = "a node's box cannot be quantised";
        out.nodes[wi - out.node_base] = w;
        bound = deepest + static_cast<uint32_t>(nc - 1);
        return ref_node(wi);
    }
};

// Walks a binary tree from `root` (children relative to `off`) the way the kernel would, checking every index, the
// depth and the leaves; returns false with a reason. (units.rs check_mesh / check_top on the uploaded buffer: the
// node buffer may hold room after a tree, so nodes past the tree are not required to be reached.)
bool check_walkable(const Node *nodes, uint64_t count, uint32_t root, uint32_t off, uint32_t max_depth, bool top,
                    uint64_t leaves_or_records, uint32_t first, std::vector<uint32_t> *records, std::string *why) {
    std::vector<std::pair<uint32_t, uint32_t>> st{{root, 0}};
    uint64_t visits = 0;
    while (!st.empty()) {
        auto [i, depth] = st.back();
        st.pop_back();
        if (i >= count) {
            *why = "a node index is outside the node buffer";
            return false;
        }
        if (++visits > count) {
            *why = "a tree reaches a node twice";
            return false;
        }
        const Node &n = nodes[i];
        if (n.count > 0) {
            if (top) {
                if (n.count != 1 || n.left_or_first >= leaves_or_records) {
                    *why = "a top-tree leaf does not name one record";
                    return false;
                }
                records->push_back(n.left_or_first);
            } else if (n.count > kMaxLeaf || uint64_t(first) + n.left_or_first + n.count > leaves_or_records) {
                *why = "a mesh-tree leaf names triangles outside the leaves";
                return false;
            }
            continue;
        }
        for (int k = 0; k < 3; k++)
            if (!(std::isfinite(n.bmin[k]) && std::isfinite(n.bmax[k]) && n.bmin[k] <= n.bmax[k])) {
                *why = "an interior node's box is not finite";
                return false;
            }
        const uint64_t l = uint64_t(off) + n.left_or_first;
        if (l <= i || l + 1 >= count) {
            *why = "an interior node's children are not after it inside the buffer";
            return false;
        }
        if (depth + 1 > max_depth) {
            *why = "a tree is deeper than the walk's stack allows";
            return false;
        }
        st.push_back({uint32_t(l), depth + 1});
        st.push_back({uint32_t(l + 1), depth + 1});
    }
    return true;
}

// The number of triangles a reference mesh tree covers (its leaves name positions 0 .. n-1 past the mesh's first).
uint32_t mesh_size(const Node *nodes, uint32_t root) {
    uint32_t n = 0;
    std::vector<uint32_t> st{root};
    while (!st.empty()) {
        const Node &x = nodes[st.back()];
        st.pop_back();
        if (x.count > 0) {
            n = std::max(n, x.left_or_first + x.count);
        } else {
            st.push_back(root + x.left_or_first);
            st.push_back(root + x.left_or_first + 1);
        }
    }
    return n;
}

inline float down(double v) {
    float f = static_cast<float>(v);
    if (double(f) > v) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float up(double v) {
    float f = static_cast<float>(v);
    if (double(f) < v) f = std::nextafter(f, INFINITY);
    return f;
}

// The library's binary SAH tree over the triangles leaves[first .. first + n): nodes[0] is the root, children at
// left_or_first and left_or_first + 1 (absolute), leaves name positions in `order` (relative to first), whose
// entries are positions in the reference order. A triangle whose corners are not finite can never be hit and is
// left out of the tree (it keeps a place in `order`).
struct OwnTree {
    std::vector<Node> nodes;
    std::vector<uint32_t> order;

    struct Box {
        float lo[3], hi[3];
    };
    static float area(const Box &b) {
        const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
        return 2.0f * (x * y + y * z + z * x);
    }
    static void grow(Box &b, const Box &o) {
        for (int k = 0; k < 3; k++) b.lo[k] = std::min(b.lo[k], o.lo[k]), b.hi[k] = std::max(b.hi[k], o.hi[k]);
    }
    static Box empty() { return Box{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}}; }

    void build(const Leaf *leaves, uint32_t first, uint32_t n) {
        std::vector<Box> box(n);
        std::vector<float> cen(3 * size_t(n));
        std::vector<uint32_t> live, dead;
        live.reserve(n);
        for (uint32_t i = 0; i < n; i++) {
            const Leaf &l = leaves[uint64_t(first) + i];
            bool finite = true;
            for (int k = 0; k < 3; k++) {
                const double v0 = l.r[0][k], v1 = v0 + double(l.r[1][k]), v2 = v0 + double(l.r[2][k]);
                if (!std::isfinite(v0) || !std::isfinite(v1) || !std::isfinite(v2)) finite = false;
                // + 0.0f makes a zero bound +0: min and max then do not depend on the order they see items in
                box[i].lo[k] = down(std::min(v0, std::min(v1, v2))) + 0.0f;
                box[i].hi[k] = up(std::max(v0, std::max(v1, v2))) + 0.0f;
                cen[3 * i + k] = 0.5f * box[i].lo[k] + 0.5f * box[i].hi[k];
            }
            (finite ? live : dead).push_back(i);
        }
        nodes.clear();
        nodes.reserve(2 * size_t(live.size()) / 2 + 2);
        order = live;
        nodes.emplace_back();
        if (live.empty()) {
            // nothing can be hit: one leaf over the first triangle, with a point box
            order = dead;
            Node &r = nodes[0];
            for (int k = 0; k < 3; k++) r.bmin[k] = r.bmax[k] = 0.0f;
            r.left_or_first = 0;
            r.count = 1;
            return;
        }
        struct Item {
            uint32_t node, b, e, depth;
        };
        std::vector<Item> st{{0, 0, uint32_t(live.size()), 0}};
        while (!st.empty()) {
            const Item it = st.back();
            st.pop_back();
            Box nb = empty(), cb = empty();
            for (uint32_t j = it.b; j < it.e; j++) {
                const uint32_t p = order[j];
                grow(nb, box[p]);
                const Box c{{cen[3 * p], cen[3 * p + 1], cen[3 * p + 2]}, {cen[3 * p], cen[3 * p + 1], cen[3 * p + 2]}};
                grow(cb, c);
            }
            Node &nd = nodes[it.node];
            for (int k = 0; k < 3; k++) nd.bmin[k] = nb.lo[k], nd.bmax[k] = nb.hi[k];
            const uint32_t cnt = it.e - it.b;
            const float leaf_cost = float(cnt);
            int best_axis = -1, best_bin = -1;
            float best_cost = INFINITY;
            if (cnt > 1 && it.depth < kOwnDepth) {
                for (int k = 0; k < 3; k++) {
                    const float ext = cb.hi[k] - cb.lo[k];
                    const float scale = float(kOwnBins) / ext;
                    // no bins over an extent past the float range or a subnormal one: (c - lo) * scale would not
                    // be a number, or the scale not finite
                    if (!(ext > 0.0f) || !std::isfinite(ext) || !std::isfinite(scale)) continue;
                    Box bins[kOwnBins];
                    uint32_t counts[kOwnBins] = {};
                    for (auto &x : bins) x = empty();
                    for (uint32_t j = it.b; j < it.e; j++) {
                        const uint32_t p = order[j];
                        const int bi = std::min(kOwnBins - 1, int((cen[3 * p + k] - cb.lo[k]) * scale));
                        counts[bi]++;
                        grow(bins[bi], box[p]);
                    }
                    float right_area[kOwnBins];
                    uint32_t right_count[kOwnBins];
                    Box acc = empty();
                    uint32_t ac = 0;
                    for (int bi = kOwnBins - 1; bi > 0; bi--) {
                        grow(acc, bins[bi]);
                        ac += counts[bi];
                        right_area[bi] = ac ? area(acc) : 0.0f;
                        right_count[bi] = ac;
                    }
                    acc = empty();
                    ac = 0;
                    for (int bi = 0; bi < kOwnBins - 1; bi++) {
                        grow(acc, bins[bi]);
                        ac += counts[bi];
                        if (ac == 0 || right_count[bi + 1] == 0) continue;
                        const float cost = area(acc) * float(ac) + right_area[bi + 1] * float(right_count[bi + 1]);
                        if (cost < best_cost) best_cost = cost, best_axis = k, best_bin = bi;
                    }
                }
            }
            const float na = area(nb);
            const float split_cost = best_axis >= 0 ? 1.0f + (na > 0.0f ? best_cost / na : float(cnt)) : INFINITY;
            if (cnt <= kOwnLeaf && (cnt == 1 || !(split_cost < leaf_cost))) {
                nd.left_or_first = it.b;
                nd.count = cnt;
                continue;
            }
            uint32_t mid;
            if (best_axis >= 0) {
                const int k = best_axis;
                const float scale = float(kOwnBins) / (cb.hi[k] - cb.lo[k]);
                const float lo = cb.lo[k];
                // stable: each side keeps the order it had, so a node's triangles stay in ascending index order
                // above the depth limit (the device build's partition by prefix sums gives the same)
                auto pm = std::stable_partition(order.begin() + it.b, order.begin() + it.e, [&](uint32_t p) {
                    return std::min(kOwnBins - 1, int((cen[3 * p + k] - lo) * scale)) <= best_bin;
                });
                mid = uint32_t(pm - order.begin());
            } else {
                // all centroids equal, or too deep: split the run at its median along the widest centroid axis
                int k = 0;
                for (int a = 1; a < 3; a++)
                    if (cb.hi[a] - cb.lo[a] > cb.hi[k] - cb.lo[k]) k = a;
                mid = it.b + cnt / 2;
                std::sort(order.begin() + it.b, order.begin() + it.e, [&](uint32_t x, uint32_t y) {
                    return cen[3 * x + k] < cen[3 * y + k] || (cen[3 * x + k] == cen[3 * y + k] && x < y);
                });
            }
            if (mid == it.b || mid == it.e) mid = it.b + cnt / 2;
            const uint32_t l = static_cast<uint32_t>(nodes.size());
            nodes.emplace_back();
            nodes.emplace_back();
            nodes[it.node].left_or_first = l;
            nodes[it.node].count = 0;
            st.push_back({l, it.b, mid, it.depth + 1});
            st.push_back({l + 1, mid, it.e, it.depth + 1});
        }
        for (uint32_t p : dead) order.push_back(p);
    }
};

}  // namespace

void own_tree_host(const Leaf *leaves, uint32_t first, uint32_t n, OwnBuilt &out) {
    OwnTree own;
    own.build(leaves, first, n);
    out.nodes = std::move(own.nodes);
    out.order = std::move(own.order);
}

namespace {

bool plan_of(const Node *nodes, uint64_t node_count, const Instance *inst, uint64_t inst_count, uint64_t leaf_count,
             uint32_t top_first, WidePlan &p, std::string *why) {
    if (top_first >= node_count) {
        *why = "the top tree's root is outside the node buffer";
        return false;
    }
    if (!check_walkable(nodes, node_count, top_first, top_first, kTopDepth, true, inst_count, 0, &p.records, why))
        return false;
    std::vector<uint8_t> seen(inst_count, 0);
    for (uint32_t r : p.records) {
        if (seen[r]) {
            *why = "a record is in the top tree twice";
            return false;
        }
        seen[r] = 1;
    }
    p.mesh_of.assign(inst_count, kNone);
    std::unordered_map<uint64_t, uint32_t> index;  // (root, first) -> mesh
    for (uint32_t r : p.records) {
        const Instance &u = inst[r];
        // A record whose sections are all glass casts no shadow (scene.rs:127 rows_cast) and its walks need no row
        // flags; glass among a shadow caster's sections (the reference's F_IGLASS) and cutouts do.
        const uint32_t bits = to_bits(u.n[1][3]);
        if ((bits & kInstCutout) || ((bits & kInstGlass) && (bits & kUnitShadow))) {
            *why = (bits & kInstCutout) ? "records with cutout sections are not supported"
                                        : "records with glass among shadow-casting sections are not supported";
            return false;
        }
        const uint32_t root = to_bits(u.n[0][3]);
        if (!check_walkable(nodes, node_count, root, root, kMeshDepth, false, leaf_count, u.info[0], nullptr, why))
            return false;
        const uint64_t key = (uint64_t(root) << 32) | u.info[0];
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, uint32_t(p.meshes.size())).first;
            p.meshes.push_back(MeshRef{root, u.info[0], mesh_size(nodes, root)});
        }
        p.mesh_of[r] = it->second;
    }
    // each mesh's leaves are laid out again in its own tree's order: two meshes may not share leaves
    std::vector<std::pair<uint64_t, uint64_t>> span;
    for (const MeshRef &m : p.meshes)
        if (m.n) span.push_back({m.first, uint64_t(m.first) + m.n});
    std::sort(span.begin(), span.end());
    for (size_t k = 1; k < span.size(); k++)
        if (span[k].first < span[k - 1].second) {
            *why = "two mesh trees name the same leaves";
            return false;
        }
    return true;
}

}  // namespace

bool wide_plan(const Node *nodes, uint64_t node_count, const Instance *inst, uint64_t inst_count,
               uint64_t leaf_count, uint32_t top_first, uint32_t unit_count, WidePlan &plan, std::string *why) {
    plan = WidePlan{};
    if (unit_count == 0) return true;
    return plan_of(nodes, node_count, inst, inst_count, leaf_count, top_first, plan, why);
}

bool build_wide(const Node *nodes, uint64_t node_count, const Leaf *leaves, const Instance *inst, uint64_t inst_count,
                uint64_t leaf_count, uint32_t top_first, uint32_t unit_count, WideScene &out, std::string *why,
                const WidePlan *plan, const std::vector<OwnBuilt> *pre, const MeshWide *mesh_wide) {
    out = WideScene();
    out.units.assign(inst_count, WideUnit{});
    if (!mesh_wide) {
        out.wleaves.assign(leaves, leaves + leaf_count);
        for (uint64_t i = 0; i < leaf_count; i++) out.wleaves[i].r[2][3] = from_bits(static_cast<uint32_t>(i));
    }
    if (unit_count == 0) return true;
    WidePlan own_plan;
    if (!plan) {
        if (!plan_of(nodes, node_count, inst, inst_count, leaf_count, top_first, own_plan, why)) return false;
        plan = &own_plan;
    }
    const WidePlan &p = *plan;
    if ((pre && pre->size() != p.meshes.size()) ||
        (mesh_wide && (mesh_wide->ref.size() != p.meshes.size() || mesh_wide->bound.size() != p.meshes.size()))) {
        *why = "prebuilt mesh trees do not match the scene's meshes";
        return false;
    }
    Collapser col{nodes, node_count, out};
    std::vector<std::pair<uint32_t, uint32_t>> mesh(p.meshes.size());  // wide ref, stack bound
    if (mesh_wide) out.node_base = mesh_wide->count;  // (the device's, ahead of the top tree's)
    for (size_t i = 0; i < p.meshes.size(); i++) {
        const MeshRef m = p.meshes[i];
        if (mesh_wide) {
            mesh[i] = {mesh_wide->ref[i], mesh_wide->bound[i]};
            continue;
        }
        OwnBuilt host;
        const OwnBuilt *own = pre && !(*pre)[i].nodes.empty() ? &(*pre)[i] : nullptr;
        if (!own) {
            own_tree_host(leaves, m.first, m.n, host);
            own = &host;
        }
        if (own->order.size() != m.n) {
            *why = "a prebuilt mesh tree's order has the wrong size";
            return false;
        }
        for (uint32_t j = 0; j < m.n; j++) {
            Leaf l = leaves[uint64_t(m.first) + own->order[j]];
            l.r[2][3] = from_bits(m.first + own->order[j]);
            out.wleaves[uint64_t(m.first) + j] = l;
        }
        Collapser mc{own->nodes.data(), own->nodes.size(), out};
        uint32_t b = 0;
        const uint32_t ref = mc.convert(0, 0, false, b);
        if (!mc.err.empty()) {
            *why = mc.err;
            return false;
        }
        mesh[i] = {ref, b};
    }
    std::vector<uint32_t> bound(inst_count, 0);
    for (uint32_t r : p.records) {
        const Instance &u = inst[r];
        WideUnit w{};
        w.t[0] = u.m[0][3], w.t[1] = u.m[1][3], w.t[2] = u.m[2][3];
        w.flags = to_bits(u.n[1][3]);
        for (int k = 0; k < 3; k++) w.n0[k] = u.n[0][k], w.n1[k] = u.n[1][k], w.n2[k] = u.n[2][k];
        w.root = mesh[p.mesh_of[r]].first;
        w.first = u.info[0];
        out.units[r] = w;
        bound[r] = mesh[p.mesh_of[r]].second;
    }
    col.unit_bound = &bound;
    uint32_t b = 0;
    out.top_ref = col.convert(top_first, top_first, true, b);
    if (!col.err.empty()) {
        *why = col.err;
        return false;
    }
    out.max_stack = b;
    const Node &t = nodes[top_first];
    for (int k = 0; k < 3; k++) out.top_box[k] = t.bmin[k], out.top_box[3 + k] = t.bmax[k];
    return true;
}

}  // namespace trt
