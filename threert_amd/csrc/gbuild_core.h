// The per-item and per-node steps of the device build of the library's own mesh trees (gbuild.hip), written once
// for the device. Each step does the host build's arithmetic (wide.cpp OwnTree) in the same order, so the trees
// are the same bytes: float min / max, the bin index int((c - lo) * (16 / ext)), the surface area
// 2 (xy + yz + zx), the SAH cost and the leaf rule.
#pragma once

#include <cstdint>

#include "trt_internal.h"

#ifndef TRT_HD
#define TRT_HD __host__ __device__
#endif

namespace trt {
namespace gb {

constexpr uint32_t kLeafMark = 0xffffffffu;  // TNode::left of a leaf
constexpr uint32_t kKeyMin = 0xffffffffu;    // the identity of a min over keys
constexpr uint32_t kKeyMax = 0u;             // the identity of a max over keys
constexpr int kBinWords = 7;                 // count, low keys x3, high keys x3
constexpr int kSlotBinWords = 3 * kOwnBins * kBinWords;
constexpr int kBoundWords = 12;              // node low / high keys, centroid low / high keys

enum Mode : uint32_t { kModeLeaf = 0, kModeSah = 1, kModeMiddle = 2 };
constexpr uint32_t kAllEqual = 1u << 3;  // Slot::axes: every centroid of the node is the same point
constexpr uint32_t kSmall = 1u << 4;     // Slot::axes: the node is small (binned from its triangles in k_split)
constexpr uint32_t kSmallNode = 8;       // nodes of at most this many triangles are small

// A node of the build, in level order (its final index in the mesh's tree comes at the end).
struct TNode {
    float lo[3], hi[3];
    uint32_t b, e;        // positions (global) of its triangles
    uint32_t depth, left; // left: the level-order index of its left child, kLeafMark for a leaf
    uint32_t mesh, pad;
};
static_assert(sizeof(TNode) == 48, "TNode is 48 bytes");

// A node of the level being split.
struct Slot {
    float clo[3], scale[3];  // centroid box low corner; bins per unit of length, by axis
    uint32_t axes;           // bit k: axis k is binned; kAllEqual
    uint32_t mode, axis, bin, mid, child;
};

struct GBox {
    float lo[3], hi[3];
};

// Floats as unsigned keys in the same order (for min / max by integer atomics). -0 never occurs: boxes are
// made +0 when zero.
TRT_HD inline uint32_t fkey(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
TRT_HD inline float kfloat(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

TRT_HD inline float next_down(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) == 0) u = 0x80000001u;  // +-0 -> the negative subnormal next to it
    else if (u & 0x80000000u) u += 1;
    else u -= 1;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
TRT_HD inline float next_up(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) == 0) u = 0x00000001u;
    else if (u & 0x80000000u) u -= 1;
    else u += 1;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// The float just at or below / above a double (wide.cpp down / up), for finite values.
TRT_HD inline float down_of(double v) {
    float f = static_cast<float>(v);
    if (double(f) > v) f = next_down(f);
    return f;
}
TRT_HD inline float up_of(double v) {
    float f = static_cast<float>(v);
    if (double(f) < v) f = next_up(f);
    return f;
}

TRT_HD inline bool finite_d(double v) { return v - v == 0.0; }

// A triangle's box (its corners from the leaf in double, rounded outwards, zeros made +0); false if a corner is
// not finite.
TRT_HD inline bool tri_box(const Leaf &l, GBox &b) {
    bool finite = true;
    for (int k = 0; k < 3; k++) {
        const double v0 = l.r[0][k], v1 = v0 + double(l.r[1][k]), v2 = v0 + double(l.r[2][k]);
        if (!finite_d(v0) || !finite_d(v1) || !finite_d(v2)) finite = false;
        const double mn = v0 < (v1 < v2 ? v1 : v2) ? v0 : (v1 < v2 ? v1 : v2);
        const double mx = v0 > (v1 > v2 ? v1 : v2) ? v0 : (v1 > v2 ? v1 : v2);
        b.lo[k] = down_of(mn) + 0.0f;
        b.hi[k] = up_of(mx) + 0.0f;
    }
    return finite;
}

TRT_HD inline float centroid(const GBox &b, int k) { return 0.5f * b.lo[k] + 0.5f * b.hi[k]; }

TRT_HD inline int bin_of(float c, float lo, float scale) {
    const int bi = int((c - lo) * scale);
    return bi < kOwnBins - 1 ? bi : kOwnBins - 1;
}

TRT_HD inline float area3(const float *lo, const float *hi) {
    const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    return 2.0f * (x * y + y * z + z * x);
}

// The node's box and centroid box from its keys; which axes it bins (wide.cpp: cnt > 1, above the depth limit,
// a finite extent > 0 with a finite scale). Returns false if the node needs the host (a split past the depth limit,
// or a median split whose order the device does not keep).
TRT_HD inline bool slot_setup(TNode &n, const uint32_t *bk, Slot &s) {
    float clo[3], chi[3];
    for (int k = 0; k < 3; k++) {
        n.lo[k] = kfloat(bk[k]);
        n.hi[k] = kfloat(bk[3 + k]);
        clo[k] = kfloat(bk[6 + k]);
        chi[k] = kfloat(bk[9 + k]);
    }
    const uint32_t cnt = n.e - n.b;
    s.axes = 0;
    s.mode = kModeLeaf;
    s.axis = s.bin = 0;
    s.mid = n.e;
    s.child = 0;
    bool all_equal = true;
    for (int k = 0; k < 3; k++) {
        s.clo[k] = clo[k];
        const float ext = chi[k] - clo[k];
        const float scale = float(kOwnBins) / ext;
        s.scale[k] = scale;
        if (chi[k] != clo[k]) all_equal = false;
        if (cnt > 1 && n.depth < kOwnDepth && ext > 0.0f && ext - ext == 0.0f && scale - scale == 0.0f)
            s.axes |= 1u << k;
    }
    // a node the host would split at its median: the device keeps the order of such a split only when every
    // centroid is the same point (its triangles are then sorted by index already, and stay so)
    if (cnt > kOwnLeaf && s.axes == 0 && !(all_equal && n.depth < kOwnDepth)) return false;
    if (all_equal) s.axes |= kAllEqual;
    if (cnt <= kSmallNode) s.axes |= kSmall;
    return true;
}

// Bins of a node as the sweep reads them: count and box of bin bi on axis k.
struct BinWords {  // from the binning kernel's words
    const uint32_t *bins;
    TRT_HD uint32_t get(int k, int bi, float *lo, float *hi) const {
        const uint32_t *w = bins + (k * kOwnBins + bi) * kBinWords;
        if (w[0])
            for (int j = 0; j < 3; j++) lo[j] = kfloat(w[1 + j]), hi[j] = kfloat(w[4 + j]);
        return w[0];
    }
};
struct BinItems {  // straight from a small node's triangles, loaded once (the same min / max, so the same floats)
    GBox x[kSmallNode];
    uint8_t bin[3][kSmallNode];
    uint32_t n;
    TRT_HD BinItems(const GBox *box, const uint32_t *order, uint32_t b, uint32_t e, const Slot &s) : n(e - b) {
        for (uint32_t i = 0; i < n; i++) {
            x[i] = box[order[b + i]];
            for (int k = 0; k < 3; k++)
                bin[k][i] = (s.axes >> k & 1u) ? uint8_t(bin_of(centroid(x[i], k), s.clo[k], s.scale[k])) : 0;
        }
    }
    TRT_HD uint32_t get(int k, int bi, float *lo, float *hi) const {
        uint32_t c = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (bin[k][i] != bi) continue;
            for (int q = 0; q < 3; q++) {
                lo[q] = c == 0 || x[i].lo[q] < lo[q] ? x[i].lo[q] : lo[q];
                hi[q] = c == 0 || hi[q] < x[i].hi[q] ? x[i].hi[q] : hi[q];
            }
            c++;
        }
        return c;
    }
};

// The host's SAH sweep over the node's bins and its leaf rule; fills the slot's decision. Returns 1 if the node
// splits, 0 for a leaf, -1 if the host must build the mesh (a median split of distinct centroids, as when every
// split's cost overflows).
template <class Bins>
TRT_HD inline int split_decide(const TNode &n, const Bins &bins, Slot &s) {
    const uint32_t cnt = n.e - n.b;
    const float leaf_cost = float(cnt);
    int best_axis = -1, best_bin = -1;
    float best_cost = __builtin_inff();
    for (int k = 0; k < 3; k++) {
        if (!(s.axes >> k & 1u)) continue;
        float right_area[kOwnBins];
        uint32_t right_count[kOwnBins];
        float alo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float ahi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        uint32_t ac = 0;
        for (int bi = kOwnBins - 1; bi > 0; bi--) {
            float l[3], h[3];
            const uint32_t cnt_bi = bins.get(k, bi, l, h);
            if (cnt_bi) {
                for (int j = 0; j < 3; j++) {
                    alo[j] = l[j] < alo[j] ? l[j] : alo[j];
                    ahi[j] = ahi[j] < h[j] ? h[j] : ahi[j];
                }
            }
            ac += cnt_bi;
            right_area[bi] = ac ? area3(alo, ahi) : 0.0f;
            right_count[bi] = ac;
        }
        for (int j = 0; j < 3; j++) alo[j] = __builtin_inff(), ahi[j] = -__builtin_inff();
        ac = 0;
        for (int bi = 0; bi < kOwnBins - 1; bi++) {
            float l[3], h[3];
            const uint32_t cnt_bi = bins.get(k, bi, l, h);
            if (cnt_bi) {
                for (int j = 0; j < 3; j++) {
                    alo[j] = l[j] < alo[j] ? l[j] : alo[j];
                    ahi[j] = ahi[j] < h[j] ? h[j] : ahi[j];
                }
            }
            ac += cnt_bi;
            if (ac == 0 || right_count[bi + 1] == 0) continue;
            const float cost = area3(alo, ahi) * float(ac) + right_area[bi + 1] * float(right_count[bi + 1]);
            if (cost < best_cost) best_cost = cost, best_axis = k, best_bin = bi;
        }
    }
    const float na = area3(n.lo, n.hi);
    const float split_cost = best_axis >= 0 ? 1.0f + (na > 0.0f ? best_cost / na : float(cnt)) : __builtin_inff();
    if (cnt <= kOwnLeaf && (cnt == 1 || !(split_cost < leaf_cost))) {
        s.mode = kModeLeaf;
        return 0;
    }
    if (best_axis >= 0) {
        uint32_t nl = 0;
        for (int bi = 0; bi <= best_bin; bi++) {
            float l[3], h[3];
            nl += bins.get(best_axis, bi, l, h);
        }
        s.mode = kModeSah;
        s.axis = uint32_t(best_axis);
        s.bin = uint32_t(best_bin);
        s.mid = n.b + nl;
    } else {
        if (!(s.axes & kAllEqual) || n.depth >= kOwnDepth) return -1;
        s.mode = kModeMiddle;
        s.mid = n.b + cnt / 2;
    }
    return 1;
}

// Whether the triangle at position j goes to the left child.
TRT_HD inline bool goes_left(const Slot &s, const GBox &b, uint32_t j) {
    if (s.mode == kModeSah) return uint32_t(bin_of(centroid(b, s.axis), s.clo[s.axis], s.scale[s.axis])) <= s.bin;
    if (s.mode == kModeMiddle) return j < s.mid;
    return false;
}

// ---- the collapse into 4-wide nodes and their quantisation (wide.cpp Collapser and quantise) ---------------------
TRT_HD inline uint32_t wref_tris(uint32_t first, uint32_t count) {
    return (kRefTris << kRefKindShift) | ((count - 1) << kRefTriFirstBits) | first;
}

// A binary node's surface area, its extents clamped at 0 (the Collapser's measure for which child to open).
TRT_HD inline float open_area(const Node &n) {
    const float x0 = n.bmax[0] - n.bmin[0], y0 = n.bmax[1] - n.bmin[1], z0 = n.bmax[2] - n.bmin[2];
    const float ex = x0 < 0.0f ? 0.0f : x0, ey = y0 < 0.0f ? 0.0f : y0, ez = z0 < 0.0f ? 0.0f : z0;
    return 2.0f * (ex * ey + ey * ez + ez * ex);
}

// The children of the wide node made from binary interior node `b` (children at nodes[left_or_first] + {0, 1},
// relative to `off`): while fewer than four, the interior child of largest area (the first of equals) is replaced
// by its two children in place. Returns the count.
TRT_HD inline int open_kids(const Node *nodes, uint32_t off, uint32_t b, uint32_t *kid) {
    kid[0] = off + nodes[b].left_or_first;
    kid[1] = kid[0] + 1;
    kid[2] = kid[3] = 0;
    int nc = 2;
    while (nc < 4) {
        int best = -1;
        float best_area = -1.0f;
        for (int c = 0; c < nc; c++) {
            const Node &k = nodes[kid[c]];
            if (k.count == 0 && open_area(k) > best_area) best_area = open_area(k), best = c;
        }
        if (best < 0) break;
        const uint32_t l = off + nodes[kid[best]].left_or_first;
        for (int c = nc; c > best + 1; c--) kid[c] = kid[c - 1];
        kid[best] = l;
        kid[best + 1] = l + 1;
        nc++;
    }
    return nc;
}

// ceil(log2(y)) for a positive normal double, exactly (from its exponent and whether its mantissa is zero).
TRT_HD inline int ceil_log2(double y) {
    uint64_t u;
    __builtin_memcpy(&u, &y, 8);
    const int e = int((u >> 52) & 0x7ff) - 1023;
    return (u & ((uint64_t{1} << 52) - 1)) ? e + 1 : e;
}
TRT_HD inline double pow2(int e) {  // 2^e for -1022 <= e <= 1023
    const uint64_t u = uint64_t(e + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &u, 8);
    return d;
}
TRT_HD inline float next_down_d(double want) {  // the float at or below `want`
    float f = static_cast<float>(want);
    if (double(f) > want) f = next_down(f);
    return f;
}

// Conservative 8-bit child boxes against a float origin and a power-of-two step per axis; false if an axis cannot
// be held (wide.cpp quantise).
TRT_HD inline bool quantise_node(WideNode &w, const Node *const *kids, int nc) {
    for (int k = 0; k < 3; k++) {
        double lo = __builtin_inf(), hi = -__builtin_inf();
        for (int c = 0; c < nc; c++) {
            const double a = kids[c]->bmin[k], b = kids[c]->bmax[k];
            lo = a < lo ? a : lo;
            hi = hi < b ? b : hi;
        }
        const double ext = hi - lo;
        int e = -100;
        if (ext > 0) {
            const int c2 = ceil_log2(ext / 252.0);
            e = c2 < -100 ? -100 : c2;
        }
        for (;; e++) {
            if (e > 126) return false;
            const double s = pow2(e);
            const float of = next_down_d(lo - s);
            if (!(double(of) - double(of) == 0.0)) return false;
            bool ok = true;
            uint32_t qlo = 0, qhi = 0;
            for (int c = 0; c < nc && ok; c++) {
                const double clo = kids[c]->bmin[k], chi = kids[c]->bmax[k];
                double ql = __builtin_floor((clo - double(of)) / s) - 1.0;
                double qh = __builtin_ceil((chi - double(of)) / s) + 1.0;
                if (ql < 0) ql = 0;
                while (ql > 0 && double(of) + ql * s > clo) ql -= 1.0;
                while (double(of) + qh * s < chi) qh += 1.0;
                if (qh > 255.0 || double(of) + ql * s > clo) {
                    ok = false;
                    break;
                }
                qlo |= uint32_t(ql) << (8 * c);
                qhi |= uint32_t(qh) << (8 * c);
            }
            if (!ok) continue;
            for (int c = nc; c < 4; c++) qlo |= 0xffu << (8 * c);  // an empty slot: low 255, high 0
            w.origin[k] = of;
            w.exp[k] = static_cast<uint8_t>(e + 127);
            w.qlo[k] = qlo;
            w.qhi[k] = qhi;
            break;
        }
    }
    w.nchild = static_cast<uint8_t>(nc);
    return true;
}

}  // namespace gb
}  // namespace trt
