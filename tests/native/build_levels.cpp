// The device build's level-synchronous algorithm (threert_amd/csrc/gbuild.hip) run serially on the host with the
// same per-item and per-node steps (gbuild_core.h), against the host build (wide.cpp own_tree_host): the same node
// bytes and the same order, on random soups of several meshes at once and the degenerate cases (copies, non-finite
// corners, coordinates near the float limit, a needle, one triangle, sizes spaced by powers of two, signed zeros).
// It checks the plan (bounds, binning, SAH decision, stable partition by prefix sums, the host's numbering from
// subtree sizes) where a GPU is not at hand; tests/test_gpu_build.py checks the kernels themselves.
#define TRT_HD
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "gbuild_core.h"

using namespace trt;
using namespace trt::gb;

namespace {

struct MeshDev {
    uint32_t first, base, n, live;
};

std::vector<uint32_t> exclusive_scan(const std::vector<uint32_t> &in) {
    std::vector<uint32_t> out(in.size());
    uint32_t s = 0;
    for (size_t i = 0; i < in.size(); i++) out[i] = s, s += in[i];
    return out;
}

bool is_min_word(uint32_t w) { return w < 3 || (w >= 6 && w < 9); }

// gbuild.hip device_wide's tree part, one step at a time; out[i] is meshes[i]'s tree (empty: left to the host).
void levels_build(const Leaf *leaves, const std::vector<MeshRef> &meshes, std::vector<OwnBuilt> &out) {
    const uint32_t mc = uint32_t(meshes.size());
    out.assign(mc, OwnBuilt{});
    uint32_t n = 0;
    std::vector<MeshDev> hm(mc);
    for (uint32_t i = 0; i < mc; i++) hm[i] = {meshes[i].first, n, meshes[i].n, 0}, n += meshes[i].n;
    auto mesh_of = [&](uint32_t g) {
        uint32_t lo = 0, hi = mc;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (hm[mid].base <= g) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    // k_boxes, k_place: the finite triangles of each mesh first, in index order
    std::vector<GBox> box(n);
    std::vector<uint32_t> finite(n + 1, 0);
    for (uint32_t g = 0; g < n; g++) {
        const MeshDev &m = hm[mesh_of(g)];
        finite[g] = tri_box(leaves[m.first + (g - m.base)], box[g]);
    }
    const std::vector<uint32_t> before = exclusive_scan(finite);
    std::vector<int32_t> mesh_slot(mc, -1);
    std::vector<uint32_t> roots;
    for (uint32_t i = 0; i < mc; i++) {
        hm[i].live = before[hm[i].base + hm[i].n] - before[hm[i].base];
        if (hm[i].live) mesh_slot[i] = int32_t(roots.size()), roots.push_back(i);
    }
    std::vector<uint32_t> order(n), order2(n);
    std::vector<int32_t> pos(n), pos2(n);
    for (uint32_t g = 0; g < n; g++) {
        const uint32_t mi = mesh_of(g);
        const MeshDev &m = hm[mi];
        const uint32_t rank = before[g] - before[m.base];
        const uint32_t p = finite[g] ? m.base + rank : m.base + m.live + (g - m.base - rank);
        order[p] = g;
        pos[p] = finite[g] ? mesh_slot[mi] : -1;
    }
    uint64_t live = 0;
    for (const MeshDev &m : hm) live += m.live;
    const uint64_t cap = 2 * live + mc;
    std::vector<TNode> tn(cap);
    std::vector<uint32_t> inner(cap), rank(cap), index(cap), host_mesh(mc, 0);
    for (size_t k = 0; k < roots.size(); k++) {
        TNode t{};
        t.b = hm[roots[k]].base, t.e = t.b + hm[roots[k]].live, t.left = kLeafMark, t.mesh = roots[k];
        tn[k] = t;
    }
    std::vector<std::pair<uint32_t, uint32_t>> levels;
    uint32_t level_base = 0, m = uint32_t(roots.size());
    uint64_t total = m;
    while (m) {
        levels.push_back({level_base, m});
        // k_clear, k_bounds
        std::vector<uint32_t> bounds(uint64_t(m) * kBoundWords), bins(uint64_t(m) * kSlotBinWords);
        for (uint64_t i = 0; i < bounds.size(); i++) bounds[i] = is_min_word(i % kBoundWords) ? kKeyMin : kKeyMax;
        for (uint64_t i = 0; i < bins.size(); i++) {
            const uint32_t w = i % kBinWords;
            bins[i] = w == 0 ? 0u : (w < 4 ? kKeyMin : kKeyMax);
        }
        for (uint32_t j = 0; j < n; j++) {
            if (pos[j] < 0) continue;
            const GBox &b = box[order[j]];
            uint32_t *d = &bounds[uint64_t(pos[j]) * kBoundWords];
            for (int q = 0; q < 3; q++) {
                const uint32_t c = fkey(centroid(b, q));
                d[q] = std::min(d[q], fkey(b.lo[q]));
                d[3 + q] = std::max(d[3 + q], fkey(b.hi[q]));
                d[6 + q] = std::min(d[6 + q], c);
                d[9 + q] = std::max(d[9 + q], c);
            }
        }
        // k_slots, k_bins, k_split
        std::vector<Slot> slots(m);
        for (uint32_t a = 0; a < m; a++) {
            TNode nd = tn[level_base + a];
            Slot s;
            if (!slot_setup(nd, &bounds[uint64_t(a) * kBoundWords], s)) {
                host_mesh[nd.mesh] = 1;
                s.axes = 0, s.mode = kModeLeaf;
            }
            tn[level_base + a] = nd;
            slots[a] = s;
        }
        for (uint32_t j = 0; j < n; j++) {
            if (pos[j] < 0 || !(slots[pos[j]].axes & 7u) || (slots[pos[j]].axes & kSmall)) continue;
            const Slot &s = slots[pos[j]];
            const GBox &b = box[order[j]];
            for (int ax = 0; ax < 3; ax++) {
                if (!(s.axes >> ax & 1u)) continue;
                const int bi = bin_of(centroid(b, ax), s.clo[ax], s.scale[ax]);
                uint32_t *d = &bins[uint64_t(pos[j]) * kSlotBinWords + (ax * kOwnBins + bi) * kBinWords];
                d[0]++;
                for (int q = 0; q < 3; q++) {
                    d[1 + q] = std::min(d[1 + q], fkey(b.lo[q]));
                    d[4 + q] = std::max(d[4 + q], fkey(b.hi[q]));
                }
            }
        }
        std::vector<uint32_t> splits(m + 1, 0);
        for (uint32_t a = 0; a < m; a++) {
            const TNode &nd = tn[level_base + a];
            int r = 0;
            if (slots[a].axes != 0 || nd.e - nd.b <= kOwnLeaf) {
                if (slots[a].axes & kSmall)
                    r = split_decide(nd, BinItems(box.data(), order.data(), nd.b, nd.e, slots[a]), slots[a]);
                else r = split_decide(nd, BinWords{&bins[uint64_t(a) * kSlotBinWords]}, slots[a]);
            }
            if (r < 0) host_mesh[nd.mesh] = 1, r = 0, slots[a].mode = kModeLeaf;
            splits[a] = r > 0;
        }
        // k_children, k_flags, k_move
        const std::vector<uint32_t> split_before = exclusive_scan(splits);
        const uint32_t nsplit = split_before[m], next_base = level_base + m;
        for (uint32_t a = 0; a < m; a++) {
            TNode &nd = tn[level_base + a];
            Slot &s = slots[a];
            if (s.mode == kModeLeaf) {
                nd.left = kLeafMark;
                continue;
            }
            const uint32_t c = 2 * split_before[a];
            s.child = c;
            nd.left = next_base + c;
            TNode l{}, r{};
            l.b = nd.b, l.e = s.mid, r.b = s.mid, r.e = nd.e;
            l.depth = r.depth = nd.depth + 1;
            l.mesh = r.mesh = nd.mesh;
            l.left = r.left = kLeafMark;
            tn[next_base + c] = l;
            tn[next_base + c + 1] = r;
        }
        if (nsplit) {
            std::vector<uint32_t> flag(n + 1, 0);
            for (uint32_t j = 0; j < n; j++)
                if (pos[j] >= 0) flag[j] = goes_left(slots[pos[j]], box[order[j]], j);
            const std::vector<uint32_t> left_before = exclusive_scan(flag);
            for (uint32_t j = 0; j < n; j++) {
                const int32_t a = pos[j];
                if (a < 0 || slots[a].mode == kModeLeaf) {
                    order2[j] = order[j], pos2[j] = -1;
                    continue;
                }
                const TNode &nd = tn[level_base + a];
                const uint32_t lr = left_before[j] - left_before[nd.b];
                const uint32_t nl = left_before[nd.e] - left_before[nd.b];
                const bool left = left_before[j + 1] != left_before[j];
                const uint32_t p = left ? nd.b + lr : nd.b + nl + (j - nd.b - lr);
                order2[p] = order[j];
                pos2[p] = int32_t(slots[a].child + (left ? 0u : 1u));
            }
            std::swap(order, order2);
            std::swap(pos, pos2);
        }
        level_base = next_base;
        m = 2 * nsplit;
        total += m;
    }
    // k_inner, k_rank, k_emit: the host's numbering
    for (size_t k = levels.size(); k-- > 0;)
        for (uint32_t i = 0; i < levels[k].second; i++) {
            const uint32_t t = levels[k].first + i, l = tn[t].left;
            inner[t] = l == kLeafMark ? 0 : 1 + inner[l] + inner[l + 1];
        }
    for (const auto &lv : levels)
        for (uint32_t i = 0; i < lv.second; i++) {
            const uint32_t t = lv.first + i, l = tn[t].left;
            if (l == kLeafMark) continue;
            rank[l + 1] = rank[t] + 1;
            rank[l] = rank[t] + 1 + inner[l + 1];
            index[l] = 2 * rank[t] + 1;
            index[l + 1] = 2 * rank[t] + 2;
        }
    std::vector<uint64_t> node_base(mc, 0);
    uint64_t nodes = 0;
    for (size_t k = 0; k < roots.size(); k++) node_base[roots[k]] = nodes, nodes += 2 * uint64_t(inner[k]) + 1;
    std::vector<Node> emitted(nodes);
    for (uint64_t t = 0; t < total; t++) {
        const TNode &nd = tn[t];
        if (host_mesh[nd.mesh]) continue;
        Node o;
        for (int k = 0; k < 3; k++) o.bmin[k] = nd.lo[k], o.bmax[k] = nd.hi[k];
        if (nd.left == kLeafMark) o.left_or_first = nd.b - hm[nd.mesh].base, o.count = nd.e - nd.b;
        else o.left_or_first = 2 * rank[t] + 1, o.count = 0;
        emitted[node_base[nd.mesh] + index[t]] = o;
    }
    for (size_t k = 0; k < roots.size(); k++) {
        const uint32_t i = roots[k];
        if (host_mesh[i]) continue;
        const uint64_t count = 2 * uint64_t(inner[k]) + 1;
        out[i].nodes.assign(emitted.begin() + node_base[i], emitted.begin() + node_base[i] + count);
        out[i].order.resize(hm[i].n);
        for (uint32_t q = 0; q < hm[i].n; q++) out[i].order[q] = order[hm[i].base + q] - hm[i].base;
    }
}

int check(const std::vector<Leaf> &leaves, const std::vector<MeshRef> &meshes, const char *name) {
    std::vector<OwnBuilt> got;
    levels_build(leaves.data(), meshes, got);
    int bad = 0;
    for (size_t i = 0; i < meshes.size(); i++) {
        if (got[i].nodes.empty()) continue;  // left to the host build
        OwnBuilt want;
        own_tree_host(leaves.data(), meshes[i].first, meshes[i].n, want);
        const bool same = want.nodes.size() == got[i].nodes.size() && want.order == got[i].order &&
                          !std::memcmp(want.nodes.data(), got[i].nodes.data(), want.nodes.size() * sizeof(Node));
        if (!same) {
            bad++;
            std::printf("%s mesh %zu differs: %zu nodes against %zu\n", name, i, got[i].nodes.size(), want.nodes.size());
        }
    }
    std::printf("%s: %zu meshes, %s\n", name, meshes.size(), bad ? "BAD" : "equal");
    return bad;
}

Leaf tri(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz) {
    Leaf l{};
    l.r[0][0] = ax, l.r[0][1] = ay, l.r[0][2] = az;
    l.r[1][0] = bx - ax, l.r[1][1] = by - ay, l.r[1][2] = bz - az;
    l.r[2][0] = cx - ax, l.r[2][1] = cy - ay, l.r[2][2] = cz - az;
    return l;
}

}  // namespace

int main() {
    int bad = 0;
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(-10, 10), S(0.001f, 1.0f);
    for (int trial = 0; trial < 6; trial++) {  // random soups, 1 to 6 meshes built together
        std::vector<Leaf> L;
        std::vector<MeshRef> ms;
        for (int mi = 0; mi <= trial; mi++) {
            const uint32_t n = 1 + rng() % (trial < 3 ? 50 : 20000);
            ms.push_back(MeshRef{0, uint32_t(L.size()), n});
            for (uint32_t i = 0; i < n; i++) {
                const float x = U(rng), y = U(rng), z = U(rng), s = S(rng);
                L.push_back(tri(x, y, z, x + s * U(rng) * 0.1f, y + s, z, x, y + s * U(rng) * 0.1f, z + s));
            }
        }
        char name[32];
        std::snprintf(name, sizeof name, "soup%d", trial);
        bad += check(L, ms, name);
    }
    std::vector<Leaf> L;
    for (int i = 0; i < 1000; i++) L.push_back(tri(0, 0, 0, 1, 0, 0, 0, 1, 0));
    bad += check(L, {MeshRef{0, 0, 1000}}, "copies");
    L.clear();
    for (int i = 0; i < 600; i++) {
        const float x = U(rng);
        if (i % 7 == 0) L.push_back(tri(NAN, 0, 0, 1, 1, 1, 2, 2, 2));
        else if (i % 11 == 0) L.push_back(tri(-3e38f, 0, 0, 3e38f, 1, 0, 0, 3e38f, 1));
        else L.push_back(tri(x, x, x, x + 1, x, x, x, x + 1, x));
    }
    bad += check(L, {MeshRef{0, 0, 600}}, "garbage");
    L.clear();
    for (int i = 0; i < 500; i++) {
        const float t = i * 0.001f;
        L.push_back(tri(t, 0, 0, t + 0.001f, 1e-6f, 0, t, 0, 1e-6f));
    }
    bad += check(L, {MeshRef{0, 0, 500}}, "needle");
    L.assign(1, tri(1, 2, 3, 4, 5, 6, 7, 8, 10));
    bad += check(L, {MeshRef{0, 0, 1}}, "one");
    L.clear();
    for (int i = 0; i < 120; i++) {
        const float s = std::ldexp(1.0f, i - 60);
        L.push_back(tri(s, 0, 0, s * 1.5f, s, 0, s, 0, s));
    }
    bad += check(L, {MeshRef{0, 0, 120}}, "powers");
    L.assign(10, tri(NAN, 0, 0, 1, 1, 1, 2, 2, 2));
    bad += check(L, {MeshRef{0, 0, 10}}, "allnan");
    L.clear();
    for (int i = 0; i < 64; i++) {
        L.push_back(tri(-0.0f, 0, 0, 0, 1, 0, 0, 0, 1));
        L.push_back(tri(0.0f, -0.0f, 0, -1, 1, 0, 0, 0, -1));
    }
    bad += check(L, {MeshRef{0, 0, 128}}, "zeros");
    std::printf("%s\n", bad ? "FAIL" : "ALL EQUAL");
    return bad ? 1 : 0;
}
