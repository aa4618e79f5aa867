// The library's own mesh trees built on the device, one level of the tree at a time over every mesh of the scene
// together. The trees are the host build's bytes (wide.cpp OwnTree): the same binned SAH over the same float
// arithmetic (gbuild_core.h), with the split's stable partition made from prefix sums, and the host's node
// numbering (children allocated in pairs as a right-first depth-first walk meets their parents) recovered at the
// end from subtree sizes.
//
// Each level: the bounds of every open node (min / max over its triangles' boxes and centroids, as ordered integer
// keys), its bins on the axes it splits on (count and box of 16 bins an axis), the host's SAH sweep and leaf rule
// one node a thread, then the prefix sums that move each triangle to its child's range. A workgroup takes 1024
// consecutive positions; the open nodes among them are consecutive, and the first kPriv of them gather in LDS
// before one set of global atomics, so the top levels (a few nodes over millions of triangles) do not queue on
// the same addresses.
//
// A mesh the device cannot finish as the host would (a split at the median of distinct centroids: past depth 64,
// or when every split's cost overflows) is marked and built on the host instead; so is a mesh with no finite
// triangle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gbuild_core.h"
#include "gbuild.h"

namespace trt {
namespace gb {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr uint32_t kChunk = kThreads * kPerThread;  // positions a workgroup takes
constexpr int kPriv = 8;                            // open nodes a workgroup gathers in LDS

struct MeshDev {
    uint32_t first;  // its first leaf in the scene's leaf buffer
    uint32_t base;   // its first item (and position) in the build
    uint32_t n;      // its triangles
    uint32_t live;   // its finite triangles (after the count)
};

__device__ inline uint32_t mesh_of(const MeshDev *m, uint32_t count, uint32_t g) {
    uint32_t lo = 0, hi = count;  // the last mesh whose base <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (m[mid].base <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---- prefix sums (exclusive, u32): tile sums, one scan of the tile sums, then each tile again ----------------------
constexpr uint32_t kScanTile = kThreads * 8;

__device__ inline uint32_t block_exclusive(uint32_t v, uint32_t *sh, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    uint32_t off = 0, all = 0;
    for (int w = 0; w < kThreads / 64; w++) {
        if (w < wave) off += sh[w];
        all += sh[w];
    }
    __syncthreads();
    *total = all;
    return off + x - v;
}

__global__ __launch_bounds__(kThreads) void k_scan_tiles(const uint32_t *in, uint64_t n, uint32_t *tile_sum) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * 8;
    uint32_t s = 0;
    for (int k = 0; k < 8; k++)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_exclusive(s, sh, &total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_scan_top(uint32_t *tile_sum, uint32_t tiles, uint32_t *grand) {
    __shared__ uint32_t sh[kThreads / 64];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kThreads) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < tiles ? tile_sum[i] : 0;
        uint32_t total;
        const uint32_t ex = block_exclusive(v, sh, &total);
        if (i < tiles) tile_sum[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t *in, uint64_t n, const uint32_t *tile_sum,
                                                         uint32_t *out) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * 8;
    uint32_t v[8], s = 0;
    for (int k = 0; k < 8; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive(s, sh, &total) + tile_sum[blockIdx.x];
    for (int k = 0; k < 8; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// ---- the build's kernels ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_boxes(const Leaf *leaves, const MeshDev *meshes, uint32_t mesh_count,
                                                    uint32_t n, GBox *box, uint32_t *finite) {
    const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
    if (g > n) return;
    if (g == n) {
        finite[g] = 0;  // the scan's last entry: the total
        return;
    }
    const MeshDev m = meshes[mesh_of(meshes, mesh_count, g)];
    GBox b;
    finite[g] = tri_box(leaves[uint64_t(m.first) + (g - m.base)], b) ? 1u : 0u;
    box[g] = b;
}

// Positions: each mesh's finite triangles first in index order, then the others (never in a node).
__global__ __launch_bounds__(kThreads) void k_place(const MeshDev *meshes, const int32_t *mesh_slot,
                                                    uint32_t mesh_count, uint32_t n, const uint32_t *finite,
                                                    const uint32_t *live_before, uint32_t *order, int32_t *pos) {
    const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
    if (g >= n) return;
    const uint32_t mi = mesh_of(meshes, mesh_count, g);
    const MeshDev m = meshes[mi];
    const uint32_t rank = live_before[g] - live_before[m.base];
    const bool live = finite[g] != 0;
    const uint32_t p = live ? m.base + rank : m.base + m.live + (g - m.base - rank);
    order[p] = g;
    pos[p] = live ? mesh_slot[mi] : -1;
}

__global__ __launch_bounds__(kThreads) void k_clear(uint32_t *bounds, uint32_t *bins, uint32_t m) {
    const uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < uint64_t(m) * kBoundWords) {
        const uint32_t w = uint32_t(i % kBoundWords);
        bounds[i] = (w < 3 || (w >= 6 && w < 9)) ? kKeyMin : kKeyMax;
    }
    if (i < uint64_t(m) * kSlotBinWords) {
        const uint32_t w = uint32_t(i % kBinWords);
        bins[i] = w == 0 ? 0u : (w < 4 ? kKeyMin : kKeyMax);
    }
}

// The first open node among a workgroup's positions (slots rise with positions).
__device__ inline int32_t first_slot(const int32_t *pos, uint32_t base, uint32_t n, int32_t *sh) {
    if (threadIdx.x == 0) *sh = 0x7fffffff;
    __syncthreads();
    int32_t mine = 0x7fffffff;
    for (int k = 0; k < kPerThread; k++) {
        const uint32_t j = base + threadIdx.x + k * kThreads;
        if (j < n && pos[j] >= 0) mine = min(mine, pos[j]);
    }
    if (mine != 0x7fffffff) atomicMin(sh, mine);
    __syncthreads();
    return *sh;
}

// Node bounds: min / max keys of the boxes and centroids of each open node's triangles.
__global__ __launch_bounds__(kThreads) void k_bounds(const uint32_t *order, const int32_t *pos, const GBox *box,
                                                     uint32_t n, uint32_t *bounds) {
    __shared__ uint32_t sh[kPriv * kBoundWords];
    __shared__ int32_t a0s;
    const uint32_t base = blockIdx.x * kChunk;
    for (int i = threadIdx.x; i < kPriv * kBoundWords; i += kThreads) {
        const int w = i % kBoundWords;
        sh[i] = (w < 3 || (w >= 6 && w < 9)) ? kKeyMin : kKeyMax;
    }
    const int32_t a0 = first_slot(pos, base, n, &a0s);  // (its barriers order the LDS fill above)
    const uint32_t lane = threadIdx.x & 63u;
    for (int k = 0; k < kPerThread; k++) {
        const uint32_t j = base + threadIdx.x + k * kThreads;
        const int32_t a = j < n ? pos[j] : -1;
        uint32_t key[kBoundWords];
        if (a >= 0) {
            const GBox b = box[order[j]];
            for (int q = 0; q < 3; q++) {
                key[q] = fkey(b.lo[q]);
                key[3 + q] = fkey(b.hi[q]);
                key[6 + q] = key[9 + q] = fkey(centroid(b, q));
            }
        } else {
            for (int q = 0; q < 3; q++) key[q] = key[6 + q] = kKeyMin, key[3 + q] = key[9 + q] = kKeyMax;
        }
        // a wave's lanes hold consecutive positions, so a node's are a run of lanes: fold each key down the run to
        // its first lane (after the step of distance d, a lane holds its run's next 2d lanes), which alone updates
        // the node
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t oa = __shfl_down(a, d, 64);
            const bool take = lane + d < 64 && oa == a;
#pragma unroll
            for (int q = 0; q < kBoundWords; q++) {
                const uint32_t o = __shfl_down(key[q], d, 64);
                const bool is_min = q < 3 || (q >= 6 && q < 9);
                if (take) key[q] = is_min ? min(key[q], o) : max(key[q], o);
            }
        }
        const int32_t prev = __shfl_up(a, 1, 64);
        if (a < 0 || (lane > 0 && prev == a)) continue;
        if (a - a0 < kPriv) {
            uint32_t *d = sh + (a - a0) * kBoundWords;
            for (int q = 0; q < 3; q++) {
                atomicMin(d + q, key[q]);
                atomicMax(d + 3 + q, key[3 + q]);
                atomicMin(d + 6 + q, key[6 + q]);
                atomicMax(d + 9 + q, key[9 + q]);
            }
        } else {
            uint32_t *d = bounds + uint64_t(a) * kBoundWords;
            for (int q = 0; q < 3; q++) {
                atomicMin(d + q, key[q]);
                atomicMax(d + 3 + q, key[3 + q]);
                atomicMin(d + 6 + q, key[6 + q]);
                atomicMax(d + 9 + q, key[9 + q]);
            }
        }
    }
    __syncthreads();
    if (a0 == 0x7fffffff) return;
    for (int i = threadIdx.x; i < kPriv * kBoundWords; i += kThreads) {
        const int w = i % kBoundWords;
        const bool is_min = w < 3 || (w >= 6 && w < 9);
        const uint32_t v = sh[i];
        if (v == (is_min ? kKeyMin : kKeyMax)) continue;  // untouched (only open nodes of this chunk touch)
        uint32_t *d = bounds + uint64_t(a0 + i / kBoundWords) * kBoundWords + w;
        if (is_min) atomicMin(d, v);
        else atomicMax(d, v);
    }
}

// Each open node's box, centroid box and binned axes; a node the device cannot split as the host would marks its
// mesh for the host.
__global__ __launch_bounds__(kThreads) void k_slots(TNode *tn, uint32_t level_base, uint32_t m,
                                                    const uint32_t *bounds, Slot *slots, uint32_t *host_mesh) {
    const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
    if (a >= m) return;
    TNode nd = tn[level_base + a];
    Slot s;
    if (!slot_setup(nd, bounds + uint64_t(a) * kBoundWords, s)) {
        host_mesh[nd.mesh] = 1;
        s.axes = 0;  // its subtree stops here: the mesh is built again on the host
        s.mode = kModeLeaf;
    }
    tn[level_base + a] = nd;
    slots[a] = s;
}

__global__ __launch_bounds__(kThreads) void k_bins(const uint32_t *order, const int32_t *pos, const GBox *box,
                                                   uint32_t n, const Slot *slots, uint32_t *bins) {
    __shared__ uint32_t sh[kPriv * kSlotBinWords];
    __shared__ int32_t a0s;
    const uint32_t base = blockIdx.x * kChunk;
    for (int i = threadIdx.x; i < kPriv * kSlotBinWords; i += kThreads) {
        const int w = i % kBinWords;
        sh[i] = w == 0 ? 0u : (w < 4 ? kKeyMin : kKeyMax);
    }
    const int32_t a0 = first_slot(pos, base, n, &a0s);
    for (int k = 0; k < kPerThread; k++) {
        const uint32_t j = base + threadIdx.x + k * kThreads;
        if (j >= n) continue;
        const int32_t a = pos[j];
        if (a < 0) continue;
        const Slot s = slots[a];
        if (!(s.axes & 7u) || (s.axes & kSmall)) continue;  // (a small node is binned in k_split)
        const GBox b = box[order[j]];
        uint32_t lo[3], hi[3];
        for (int q = 0; q < 3; q++) lo[q] = fkey(b.lo[q]), hi[q] = fkey(b.hi[q]);
        uint32_t *dst = a - a0 < kPriv ? sh + (a - a0) * kSlotBinWords : bins + uint64_t(a) * kSlotBinWords;
        for (int ax = 0; ax < 3; ax++) {
            if (!(s.axes >> ax & 1u)) continue;
            const int bi = bin_of(centroid(b, ax), s.clo[ax], s.scale[ax]);
            uint32_t *d = dst + (ax * kOwnBins + bi) * kBinWords;
            atomicAdd(d, 1u);
            for (int q = 0; q < 3; q++) {
                atomicMin(d + 1 + q, lo[q]);
                atomicMax(d + 4 + q, hi[q]);
            }
        }
    }
    __syncthreads();
    if (a0 == 0x7fffffff) return;
    for (int i = threadIdx.x; i < kPriv * kSlotBinWords; i += kThreads) {
        const int w = i % kBinWords;
        const uint32_t v = sh[i];
        if (v == (w == 0 ? 0u : (w < 4 ? kKeyMin : kKeyMax))) continue;
        uint32_t *d = bins + uint64_t(a0 + i / kSlotBinWords) * kSlotBinWords + (i % kSlotBinWords);
        if (w == 0) atomicAdd(d, v);
        else if (w < 4) atomicMin(d, v);
        else atomicMax(d, v);
    }
}

__global__ __launch_bounds__(kThreads) void k_split(const TNode *tn, uint32_t level_base, uint32_t m,
                                                    const uint32_t *bins, const GBox *box, const uint32_t *order,
                                                    Slot *slots, uint32_t *splits, uint32_t *host_mesh) {
    const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
    if (a > m) return;
    if (a == m) {
        splits[a] = 0;
        return;
    }
    const TNode nd = tn[level_base + a];
    Slot s = slots[a];
    int r = 0;
    if (s.axes != 0 || nd.e - nd.b <= kOwnLeaf) {
        if (s.axes & kSmall) r = split_decide(nd, BinItems(box, order, nd.b, nd.e, s), s);
        else r = split_decide(nd, BinWords{bins + uint64_t(a) * kSlotBinWords}, s);
    }
    if (r < 0) {
        host_mesh[nd.mesh] = 1;
        r = 0;
        s.mode = kModeLeaf;
    }
    slots[a] = s;
    splits[a] = r > 0 ? 1u : 0u;
}

__global__ __launch_bounds__(kThreads) void k_children(TNode *tn, uint32_t level_base, uint32_t next_base,
                                                       uint32_t m, Slot *slots, const uint32_t *split_before) {
    const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
    if (a >= m) return;
    TNode &nd = tn[level_base + a];
    Slot &s = slots[a];
    if (s.mode == kModeLeaf) {
        nd.left = kLeafMark;
        return;
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

__global__ __launch_bounds__(kThreads) void k_flags(const uint32_t *order, const int32_t *pos, const GBox *box,
                                                    uint32_t n, const Slot *slots, uint32_t *flag) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j > n) return;
    uint32_t f = 0;
    if (j < n && pos[j] >= 0) f = goes_left(slots[pos[j]], box[order[j]], j) ? 1u : 0u;
    flag[j] = f;
}

__global__ __launch_bounds__(kThreads) void k_move(const uint32_t *order, const int32_t *pos, uint32_t n,
                                                   const TNode *tn, uint32_t level_base, const Slot *slots,
                                                   const uint32_t *left_before, uint32_t *order2, int32_t *pos2) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= n) return;
    const int32_t a = pos[j];
    if (a < 0 || slots[a].mode == kModeLeaf) {
        order2[j] = order[j];
        pos2[j] = -1;
        return;
    }
    const Slot s = slots[a];
    const TNode nd = tn[level_base + a];
    const uint32_t lr = left_before[j] - left_before[nd.b];
    const uint32_t nl = left_before[nd.e] - left_before[nd.b];
    const bool left = left_before[j + 1] != left_before[j];
    const uint32_t p = left ? nd.b + lr : nd.b + nl + (j - nd.b - lr);
    order2[p] = order[j];
    pos2[p] = int32_t(s.child + (left ? 0u : 1u));
}

__global__ __launch_bounds__(kThreads) void k_inner(const TNode *tn, uint32_t t0, uint32_t cnt, uint32_t *inner) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t t = t0 + i;
    const uint32_t l = tn[t].left;
    inner[t] = l == kLeafMark ? 0u : 1u + inner[l] + inner[l + 1];
}

// The host's numbering: the k-th interior node a right-first depth-first walk meets has its children at 2k + 1
// and 2k + 2. A node of rank r: its right child has rank r + 1, its left child r + 1 + (interior nodes under the
// right child).
__global__ __launch_bounds__(kThreads) void k_rank(const TNode *tn, uint32_t t0, uint32_t cnt, const uint32_t *inner,
                                                   uint32_t *rank, uint32_t *index) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t t = t0 + i;
    const uint32_t l = tn[t].left;
    if (l == kLeafMark) return;
    const uint32_t r = rank[t];
    rank[l + 1] = r + 1;
    rank[l] = r + 1 + inner[l + 1];
    index[l] = 2 * r + 1;
    index[l + 1] = 2 * r + 2;
}

__global__ __launch_bounds__(kThreads) void k_emit(const TNode *tn, uint32_t count, const uint32_t *rank,
                                                   const uint32_t *index, const MeshDev *meshes,
                                                   const uint64_t *node_base, const uint32_t *host_mesh, Node *out) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= count) return;
    const TNode nd = tn[t];
    if (host_mesh[nd.mesh]) return;
    Node o;
    for (int k = 0; k < 3; k++) o.bmin[k] = nd.lo[k], o.bmax[k] = nd.hi[k];
    if (nd.left == kLeafMark) {
        o.left_or_first = nd.b - meshes[nd.mesh].base;
        o.count = nd.e - nd.b;
    } else {
        o.left_or_first = 2 * rank[t] + 1;
        o.count = 0;
    }
    out[node_base[nd.mesh] + index[t]] = o;
}

// ---- the collapse into 4-wide nodes, one wide level at a time -------------------------------------------------
// Wide node t (level order) stands for binary interior node wbin[t] of mesh wmesh[t]; its children are the binary
// nodes open_kids gives, and the interior ones among them become the next level's wide nodes.
__global__ __launch_bounds__(kThreads) void k_w_open(uint32_t t0, uint32_t cnt, const Node *own,
                                                     const uint64_t *node_base, const uint32_t *wmesh,
                                                     const uint32_t *wbin, uint32_t *wkid, uint32_t *wnc,
                                                     uint32_t *inner_kids) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i > cnt) return;
    if (i == cnt) {
        inner_kids[i] = 0;
        return;
    }
    const uint32_t t = t0 + i;
    const Node *nodes = own + node_base[wmesh[t]];
    uint32_t kid[4];
    const int nc = open_kids(nodes, 0, wbin[t], kid);
    uint32_t inner = 0;
    for (int c = 0; c < 4; c++) {
        wkid[4 * uint64_t(t) + c] = kid[c];
        if (c < nc && nodes[kid[c]].count == 0) inner++;
    }
    wnc[t] = uint32_t(nc);
    inner_kids[i] = inner;
}

__global__ __launch_bounds__(kThreads) void k_w_link(uint32_t t0, uint32_t cnt, uint32_t next_base, const Node *own,
                                                     const uint64_t *node_base, uint32_t *wmesh, uint32_t *wbin,
                                                     const uint32_t *wkid, const uint32_t *wnc,
                                                     const uint32_t *inner_before, uint32_t *wchild) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t t = t0 + i, mesh = wmesh[t];
    const Node *nodes = own + node_base[mesh];
    uint32_t next = next_base + inner_before[i];
    for (uint32_t c = 0; c < 4; c++) {
        const uint32_t k = wkid[4 * uint64_t(t) + c];
        if (c < wnc[t] && nodes[k].count == 0) {
            wchild[4 * uint64_t(t) + c] = next;
            wmesh[next] = mesh;
            wbin[next] = k;
            next++;
        } else {
            wchild[4 * uint64_t(t) + c] = kNone;
        }
    }
}

// Wide nodes under each (itself included) and the deepest a walk's stack gets under it (wide.cpp convert).
__global__ __launch_bounds__(kThreads) void k_w_size(uint32_t t0, uint32_t cnt, const uint32_t *wnc,
                                                     const uint32_t *wchild, uint32_t *wsize, uint32_t *wbound) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t t = t0 + i;
    uint32_t size = 1, deepest = 0;
    for (uint32_t c = 0; c < wnc[t]; c++) {
        const uint32_t ch = wchild[4 * uint64_t(t) + c];
        if (ch == kNone) continue;
        size += wsize[ch];
        deepest = max(deepest, wbound[ch]);
    }
    wsize[t] = size;
    wbound[t] = deepest + wnc[t] - 1;
}

// Preorder indices: a node's children follow it, each child's subtree before the next child.
__global__ __launch_bounds__(kThreads) void k_w_index(uint32_t t0, uint32_t cnt, const uint32_t *wnc,
                                                      const uint32_t *wchild, const uint32_t *wsize, uint32_t *widx) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t t = t0 + i;
    uint32_t run = widx[t] + 1;
    for (uint32_t c = 0; c < wnc[t]; c++) {
        const uint32_t ch = wchild[4 * uint64_t(t) + c];
        if (ch == kNone) continue;
        widx[ch] = run;
        run += wsize[ch];
    }
}

__global__ __launch_bounds__(kThreads) void k_w_emit(uint32_t count, const Node *own, const uint64_t *node_base,
                                                     const uint32_t *wmesh, const uint32_t *wkid,
                                                     const uint32_t *wnc, const uint32_t *wchild,
                                                     const uint32_t *widx, uint64_t total, WideNode *out,
                                                     uint32_t *err) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= count) return;
    if (widx[t] >= total) {  // (a numbering that does not fit: reported, never written)
        atomicOr(err, 4u);
        return;
    }
    const Node *nodes = own + node_base[wmesh[t]];
    const uint32_t nc = wnc[t];
    WideNode w{};
    const Node *kp[4] = {nodes, nodes, nodes, nodes};
    for (uint32_t c = 0; c < 4; c++) {
        if (c >= nc) {
            w.child[c] = kNone;
            continue;
        }
        const uint32_t k = wkid[4 * uint64_t(t) + c];
        kp[c] = nodes + k;
        const uint32_t ch = wchild[4 * uint64_t(t) + c];
        if (ch != kNone) {
            w.child[c] = widx[ch];
        } else {
            if (nodes[k].left_or_first > kRefTriFirstMask) atomicOr(err, 2u);
            w.child[c] = wref_tris(nodes[k].left_or_first, nodes[k].count);
        }
    }
    if (!quantise_node(w, kp, int(nc))) atomicOr(err, 1u);
    out[widx[t]] = w;
}

// Every leaf in the wide trees' order: a mesh's range laid out in its tree's order, each leaf carrying its
// reference index in e2.w; leaves of no mesh stay where they are.
__global__ __launch_bounds__(kThreads) void k_w_leaves(const Leaf *leaves, uint64_t leaf_count, const uint32_t *first,
                                                       const uint32_t *count, const uint32_t *base,
                                                       uint32_t meshes, const uint32_t *order, Leaf *out) {
    const uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= leaf_count) return;
    uint32_t lo = 0, hi = meshes;  // the last mesh (by first leaf) whose first <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (first[mid] <= i) lo = mid;
        else hi = mid;
    }
    uint32_t src = uint32_t(i);
    if (meshes && first[lo] <= i && i < uint64_t(first[lo]) + count[lo])
        src = first[lo] + (order[base[lo] + (uint32_t(i) - first[lo])] - base[lo]);
    Leaf l = leaves[src];
    __builtin_memcpy(&l.r[2][3], &src, 4);
    out[i] = l;
}

inline uint32_t blocks(uint64_t n) { return uint32_t((n + kThreads - 1) / kThreads); }

// The build's working memory, stream-ordered (hipMallocAsync / hipFreeAsync on the build's stream): plain
// hipFree waits for the whole device, which would hold up the renders other threads queue meanwhile.
struct Scratch {
    hipStream_t st;
    std::vector<void *> ptrs;
    explicit Scratch(hipStream_t s) : st(s) {}
    ~Scratch() {
        for (void *p : ptrs)
            if (p) (void)hipFreeAsync(p, st);
    }
    template <class T>
    void drop(T *p) {
        for (void *&q : ptrs)
            if (q && q == p) {
                (void)hipFreeAsync(q, st);
                q = nullptr;
            }
    }
    template <class T>
    hipError_t get(T **p, uint64_t count) {
        void *v = nullptr;
        const hipError_t e = hipMallocAsync(&v, std::max<uint64_t>(count * sizeof(T), 64), st);
        if (e != hipSuccess) return e;
        ptrs.push_back(v);
        *p = static_cast<T *>(v);
        return hipSuccess;
    }
};

#define GB_HIP(x)                                                                        \
    do {                                                                                 \
        const hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                          \
            *why = std::string("device build: ") + #x + ": " + hipGetErrorString(e_);   \
            return false;                                                                \
        }                                                                                \
    } while (0)

// Exclusive prefix sums of in[0 .. n) into out[0 .. n); out[n - 1] + in[n - 1] is not kept (callers append a 0).
bool scan(const uint32_t *in, uint64_t n, uint32_t *out, uint32_t *tile_sum, uint32_t *grand, hipStream_t st,
          std::string *why) {
    const uint32_t tiles = uint32_t((n + kScanTile - 1) / kScanTile);
    k_scan_tiles<<<tiles, kThreads, 0, st>>>(in, n, tile_sum);
    k_scan_top<<<1, kThreads, 0, st>>>(tile_sum, tiles, grand);
    k_scan_apply<<<tiles, kThreads, 0, st>>>(in, n, tile_sum, out);
    GB_HIP(hipGetLastError());
    return true;
}

// The meshes' wide nodes and the wide leaves from the own trees in `own` (each mesh's at node_base), all on the
// device; the caller takes the two buffers.
bool collapse(const Leaf *d_leaves, uint64_t leaf_count, const std::vector<MeshRef> &meshes,
              const std::vector<MeshDev> &hm, const std::vector<uint64_t> &node_base, const Node *own,
              const uint64_t *dbase, const uint32_t *order, hipStream_t st, Scratch &sc, hipEvent_t e0, hipEvent_t e1,
              size_t tree_levels, uint64_t own_nodes, DeviceWide &res, std::string *why) {
    const uint32_t mc = uint32_t(meshes.size());
    // each mesh's root: a leaf (no wide node: the mesh is one run of triangles) or the first wide node
    std::vector<Node> roots(mc);
    for (uint32_t i = 0; i < mc; i++)
        GB_HIP(hipMemcpyAsync(&roots[i], own + node_base[i], sizeof(Node), hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    std::vector<uint32_t> wroot_mesh;
    for (uint32_t i = 0; i < mc; i++)
        if (roots[i].count == 0) wroot_mesh.push_back(i);
    const uint64_t wcap = own_nodes / 2 + mc + 1;  // wide nodes <= interior binary nodes
    uint32_t *wmesh, *wbin, *wkid, *wnc, *wchild, *wsize, *wbound, *widx, *inner_kids, *inner_before, *err,
        *tile_sum, *grand;
    GB_HIP(sc.get(&wmesh, wcap));
    GB_HIP(sc.get(&wbin, wcap));
    GB_HIP(sc.get(&wkid, 4 * wcap));
    GB_HIP(sc.get(&wnc, wcap));
    GB_HIP(sc.get(&wchild, 4 * wcap));
    GB_HIP(sc.get(&wsize, wcap));
    GB_HIP(sc.get(&wbound, wcap));
    GB_HIP(sc.get(&widx, wcap));
    GB_HIP(sc.get(&inner_kids, wcap + 1));
    GB_HIP(sc.get(&inner_before, wcap + 1));
    GB_HIP(sc.get(&err, 1));
    GB_HIP(sc.get(&tile_sum, (wcap + 1) / kScanTile + 2));
    GB_HIP(sc.get(&grand, 1));
    GB_HIP(hipMemsetAsync(err, 0, 4, st));
    const uint32_t r0 = uint32_t(wroot_mesh.size());
    if (r0) {
        std::vector<uint32_t> zero(r0, 0);
        GB_HIP(hipMemcpyAsync(wmesh, wroot_mesh.data(), r0 * 4, hipMemcpyHostToDevice, st));
        GB_HIP(hipMemcpyAsync(wbin, zero.data(), r0 * 4, hipMemcpyHostToDevice, st));
    }
    std::vector<std::pair<uint32_t, uint32_t>> levels;
    uint32_t t0 = 0, cnt = r0;
    while (cnt) {
        levels.push_back({t0, cnt});
        k_w_open<<<blocks(uint64_t(cnt) + 1), kThreads, 0, st>>>(t0, cnt, own, dbase, wmesh, wbin, wkid, wnc,
                                                                  inner_kids);
        if (!scan(inner_kids, uint64_t(cnt) + 1, inner_before, tile_sum, grand, st, why)) return false;
        uint32_t next = 0;
        GB_HIP(hipMemcpyAsync(&next, inner_before + cnt, 4, hipMemcpyDeviceToHost, st));
        GB_HIP(hipStreamSynchronize(st));
        if (uint64_t(t0) + cnt + next > wcap) {
            *why = "device build: more wide nodes than interior nodes";
            return false;
        }
        k_w_link<<<blocks(cnt), kThreads, 0, st>>>(t0, cnt, t0 + cnt, own, dbase, wmesh, wbin, wkid, wnc,
                                                   inner_before, wchild);
        t0 += cnt;
        cnt = next;
    }
    const uint32_t wcount = t0;
    for (size_t k = levels.size(); k-- > 0;)
        k_w_size<<<blocks(levels[k].second), kThreads, 0, st>>>(levels[k].first, levels[k].second, wnc, wchild,
                                                                  wsize, wbound);
    std::vector<uint32_t> rsize(r0), rbound(r0);
    if (r0) {
        GB_HIP(hipMemcpyAsync(rsize.data(), wsize, r0 * 4, hipMemcpyDeviceToHost, st));
        GB_HIP(hipMemcpyAsync(rbound.data(), wbound, r0 * 4, hipMemcpyDeviceToHost, st));
    }
    GB_HIP(hipStreamSynchronize(st));
    // meshes' blocks in mesh order, as the host appends them
    res.mesh_ref.assign(mc, 0);
    res.mesh_bound.assign(mc, 0);
    std::vector<uint32_t> ridx(r0);
    uint64_t total = 0;
    for (uint32_t i = 0, k = 0; i < mc; i++) {
        if (roots[i].count > 0) {
            if (roots[i].left_or_first > kRefTriFirstMask) {
                *why = "a mesh has more triangles than a reference holds (2^26)";
                return false;
            }
            res.mesh_ref[i] = wref_tris(roots[i].left_or_first, roots[i].count);
            continue;
        }
        ridx[k] = uint32_t(total);
        res.mesh_ref[i] = uint32_t(total);
        res.mesh_bound[i] = rbound[k];
        total += rsize[k];
        k++;
    }
    if (total >= (uint64_t{1} << 30)) {
        *why = "device build: tree too large";
        return false;
    }
    if (r0) GB_HIP(hipMemcpyAsync(widx, ridx.data(), r0 * 4, hipMemcpyHostToDevice, st));
    for (const auto &lv : levels)
        k_w_index<<<blocks(lv.second), kThreads, 0, st>>>(lv.first, lv.second, wnc, wchild, wsize, widx);
    void *wn = nullptr, *wl = nullptr;
    GB_HIP(hipMallocAsync(&wn, std::max<uint64_t>(total * sizeof(WideNode), 64), st));  // (freed on `st` too)
    res.d_wnodes = wn;
    res.wnode_count = total;
    GB_HIP(hipMalloc(&wl, std::max<uint64_t>(leaf_count * sizeof(Leaf), 64)));
    res.d_wleaves = wl;
    if (wcount)
        k_w_emit<<<blocks(wcount), kThreads, 0, st>>>(wcount, own, dbase, wmesh, wkid, wnc, wchild, widx, total,
                                                      static_cast<WideNode *>(wn), err);
    // the leaves: meshes by first leaf
    std::vector<uint32_t> by(mc);
    for (uint32_t i = 0; i < mc; i++) by[i] = i;
    std::sort(by.begin(), by.end(), [&](uint32_t a, uint32_t b) { return meshes[a].first < meshes[b].first; });
    std::vector<uint32_t> hf(mc), hcnt(mc), hb(mc);
    for (uint32_t k = 0; k < mc; k++) hf[k] = meshes[by[k]].first, hcnt[k] = meshes[by[k]].n, hb[k] = hm[by[k]].base;
    uint32_t *df, *dc, *db;
    GB_HIP(sc.get(&df, mc));
    GB_HIP(sc.get(&dc, mc));
    GB_HIP(sc.get(&db, mc));
    GB_HIP(hipMemcpyAsync(df, hf.data(), mc * 4, hipMemcpyHostToDevice, st));
    GB_HIP(hipMemcpyAsync(dc, hcnt.data(), mc * 4, hipMemcpyHostToDevice, st));
    GB_HIP(hipMemcpyAsync(db, hb.data(), mc * 4, hipMemcpyHostToDevice, st));
    if (leaf_count)
        k_w_leaves<<<blocks(leaf_count), kThreads, 0, st>>>(d_leaves, leaf_count, df, dc, db, mc, order,
                                                            static_cast<Leaf *>(wl));
    GB_HIP(hipGetLastError());
    GB_HIP(hipEventRecord(e1, st));
    uint32_t herr = 0;
    GB_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    if (herr & 4u) {
        *why = "device build: wide node numbering out of range";
        return false;
    }
    if (herr & 2u) {
        *why = "a mesh has more triangles than a reference holds (2^26)";
        return false;
    }
    if (herr & 1u) {
        *why = "a node's box cannot be quantised";
        return false;
    }
    float ms = 0;
    GB_HIP(hipEventElapsedTime(&ms, e0, e1));
    res.times.device_ms = ms;
    res.times.levels = uint32_t(tree_levels + levels.size());
    res.times.nodes = own_nodes;
    res.complete = true;
    for (auto &o : res.trees) o = OwnBuilt{};
    return true;
}

}  // namespace

bool device_wide(const Leaf *d_leaves, uint64_t leaf_count, const std::vector<MeshRef> &meshes, hipStream_t st,
                 DeviceWide &res, std::string *why) {
    res = DeviceWide{};
    std::vector<OwnBuilt> &out = res.trees;
    BuildTimes *times = &res.times;
    out.assign(meshes.size(), OwnBuilt{});
    uint64_t n64 = 0;
    for (const MeshRef &m : meshes) n64 += m.n;
    if (n64 == 0) return true;
    if (n64 >= (uint64_t{1} << 31) - kScanTile) {
        *why = "device build: more than 2^31 triangles";
        return false;
    }
    const uint32_t n = uint32_t(n64), mc = uint32_t(meshes.size());
    struct Events {  // destroyed on every way out
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    GB_HIP(hipEventCreate(&ev.a));
    GB_HIP(hipEventCreate(&ev.b));
    hipEvent_t e0 = ev.a, e1 = ev.b;
    GB_HIP(hipEventRecord(e0, st));
    Scratch sc(st);
    std::vector<MeshDev> hm(mc);
    for (uint32_t i = 0, b = 0; i < mc; i++) {
        hm[i] = MeshDev{meshes[i].first, b, meshes[i].n, 0};
        b += meshes[i].n;
    }
    MeshDev *dm;
    GBox *box;
    uint32_t *finite, *scan_out, *tile_sum, *grand, *order, *order2, *host_mesh;
    int32_t *pos, *pos2, *mslot;
    const uint32_t tiles = uint32_t((uint64_t(n) + 1 + kScanTile - 1) / kScanTile) + 1;
    GB_HIP(sc.get(&dm, mc));
    GB_HIP(sc.get(&box, n));
    GB_HIP(sc.get(&finite, uint64_t(n) + 1));
    GB_HIP(sc.get(&scan_out, uint64_t(n) + 1));
    GB_HIP(sc.get(&tile_sum, tiles));
    GB_HIP(sc.get(&grand, 1));
    GB_HIP(sc.get(&order, n));
    GB_HIP(sc.get(&order2, n));
    GB_HIP(sc.get(&pos, n));
    GB_HIP(sc.get(&pos2, n));
    GB_HIP(sc.get(&mslot, mc));
    GB_HIP(sc.get(&host_mesh, mc));
    GB_HIP(hipMemcpyAsync(dm, hm.data(), mc * sizeof(MeshDev), hipMemcpyHostToDevice, st));
    GB_HIP(hipMemsetAsync(host_mesh, 0, mc * sizeof(uint32_t), st));

    // the triangles' boxes; the finite ones of each mesh
    k_boxes<<<blocks(uint64_t(n) + 1), kThreads, 0, st>>>(d_leaves, dm, mc, n, box, finite);
    if (!scan(finite, uint64_t(n) + 1, scan_out, tile_sum, grand, st, why)) return false;
    std::vector<uint32_t> before(uint64_t(n) + 1);
    GB_HIP(hipMemcpyAsync(before.data(), scan_out, (uint64_t(n) + 1) * 4, hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    // the count of the last entry: finite[n] is 0, so before[n] is the total
    std::vector<int32_t> hslot(mc, -1);
    std::vector<uint32_t> roots;  // level-0 nodes: one a mesh with a finite triangle
    for (uint32_t i = 0; i < mc; i++) {
        hm[i].live = before[hm[i].base + hm[i].n] - before[hm[i].base];
        if (hm[i].live) {
            hslot[i] = int32_t(roots.size());
            roots.push_back(i);
        }
    }
    GB_HIP(hipMemcpyAsync(dm, hm.data(), mc * sizeof(MeshDev), hipMemcpyHostToDevice, st));
    GB_HIP(hipMemcpyAsync(mslot, hslot.data(), mc * sizeof(int32_t), hipMemcpyHostToDevice, st));
    k_place<<<blocks(n), kThreads, 0, st>>>(dm, mslot, mc, n, finite, scan_out, order, pos);

    // nodes: at most 2 live - 1 a mesh
    uint64_t live_total = 0;
    for (auto &m : hm) live_total += m.live;
    const uint64_t tcap = 2 * live_total + mc;
    TNode *tn;
    uint32_t *inner, *rank, *index;
    GB_HIP(sc.get(&tn, tcap));
    GB_HIP(sc.get(&inner, tcap));
    GB_HIP(sc.get(&rank, tcap));
    GB_HIP(sc.get(&index, tcap));
    {
        std::vector<TNode> r0(roots.size());
        for (size_t k = 0; k < roots.size(); k++) {
            const MeshDev &m = hm[roots[k]];
            r0[k] = TNode{};
            r0[k].b = m.base, r0[k].e = m.base + m.live, r0[k].depth = 0, r0[k].left = kLeafMark;
            r0[k].mesh = roots[k];
        }
        GB_HIP(hipMemcpyAsync(tn, r0.data(), r0.size() * sizeof(TNode), hipMemcpyHostToDevice, st));
        std::vector<uint32_t> zero(roots.size(), 0);
        GB_HIP(hipMemcpyAsync(rank, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, st));
        GB_HIP(hipMemcpyAsync(index, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, st));
    }

    // level by level
    uint32_t *bounds = nullptr, *bins = nullptr, *splits = nullptr, *split_before = nullptr;
    Slot *slots = nullptr;
    uint64_t cap = 0;
    std::vector<std::pair<uint32_t, uint32_t>> levels;  // (first node, count)
    uint32_t level_base = 0, m = uint32_t(roots.size());
    uint64_t total_nodes = m;
    while (m) {
        if (m > cap) {  // (the stream is idle here: the last level's count was read back)
            cap = std::max<uint64_t>(m, 2 * cap);
            sc.drop(bounds), sc.drop(bins), sc.drop(splits), sc.drop(split_before), sc.drop(slots);
            GB_HIP(sc.get(&bounds, cap * kBoundWords));
            GB_HIP(sc.get(&bins, cap * kSlotBinWords));
            GB_HIP(sc.get(&splits, cap + 1));
            GB_HIP(sc.get(&split_before, cap + 1));
            GB_HIP(sc.get(&slots, cap));
        }
        levels.push_back({level_base, m});
        k_clear<<<blocks(uint64_t(m) * kSlotBinWords), kThreads, 0, st>>>(bounds, bins, m);
        const uint32_t chunks = uint32_t((uint64_t(n) + kChunk - 1) / kChunk);
        k_bounds<<<chunks, kThreads, 0, st>>>(order, pos, box, n, bounds);
        k_slots<<<blocks(m), kThreads, 0, st>>>(tn, level_base, m, bounds, slots, host_mesh);
        k_bins<<<chunks, kThreads, 0, st>>>(order, pos, box, n, slots, bins);
        k_split<<<blocks(uint64_t(m) + 1), kThreads, 0, st>>>(tn, level_base, m, bins, box, order, slots, splits,
                                                              host_mesh);
        if (!scan(splits, uint64_t(m) + 1, split_before, tile_sum, grand, st, why)) return false;
        uint32_t nsplit = 0;
        GB_HIP(hipMemcpyAsync(&nsplit, split_before + m, 4, hipMemcpyDeviceToHost, st));
        GB_HIP(hipStreamSynchronize(st));
        const uint32_t next_base = level_base + m;
        if (uint64_t(next_base) + 2 * uint64_t(nsplit) > tcap) {
            *why = "device build: more nodes than two a triangle";
            return false;
        }
        k_children<<<blocks(m), kThreads, 0, st>>>(tn, level_base, next_base, m, slots, split_before);
        if (nsplit) {
            k_flags<<<blocks(uint64_t(n) + 1), kThreads, 0, st>>>(order, pos, box, n, slots, finite);
            if (!scan(finite, uint64_t(n) + 1, scan_out, tile_sum, grand, st, why)) return false;
            k_move<<<blocks(n), kThreads, 0, st>>>(order, pos, n, tn, level_base, slots, scan_out, order2, pos2);
            std::swap(order, order2);
            std::swap(pos, pos2);
        }
        level_base = next_base;
        m = 2 * nsplit;
        total_nodes += m;
        GB_HIP(hipGetLastError());
    }

    // the host's numbering, each mesh's nodes in one buffer
    for (size_t k = levels.size(); k-- > 0;)
        k_inner<<<blocks(levels[k].second), kThreads, 0, st>>>(tn, levels[k].first, levels[k].second, inner);
    for (const auto &lv : levels)
        k_rank<<<blocks(lv.second), kThreads, 0, st>>>(tn, lv.first, lv.second, inner, rank, index);
    std::vector<uint32_t> root_inner(roots.size()), failed(mc);
    if (!roots.empty())
        GB_HIP(hipMemcpyAsync(root_inner.data(), inner, roots.size() * 4, hipMemcpyDeviceToHost, st));
    GB_HIP(hipMemcpyAsync(failed.data(), host_mesh, mc * 4, hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> node_base(mc, 0);
    uint64_t out_nodes = 0;
    for (size_t k = 0; k < roots.size(); k++) {
        node_base[roots[k]] = out_nodes;
        out_nodes += 2 * uint64_t(root_inner[k]) + 1;
    }
    uint64_t *dbase;
    Node *dout;
    GB_HIP(sc.get(&dbase, mc));
    GB_HIP(sc.get(&dout, out_nodes));
    GB_HIP(hipMemcpyAsync(dbase, node_base.data(), mc * 8, hipMemcpyHostToDevice, st));
    k_emit<<<blocks(total_nodes), kThreads, 0, st>>>(tn, uint32_t(total_nodes), rank, index, dm, dbase, host_mesh,
                                                     dout);
    GB_HIP(hipGetLastError());
    bool all_device = roots.size() == mc;
    for (uint32_t i = 0; i < mc; i++) all_device = all_device && !failed[i];
    if (all_device) return collapse(d_leaves, leaf_count, meshes, hm, node_base, dout, dbase, order, st, sc, e0, e1,
                                    levels.size(), out_nodes, res, why);
    GB_HIP(hipEventRecord(e1, st));
    std::vector<Node> hn(out_nodes);
    std::vector<uint32_t> ho(n);
    GB_HIP(hipMemcpyAsync(hn.data(), dout, out_nodes * sizeof(Node), hipMemcpyDeviceToHost, st));
    GB_HIP(hipMemcpyAsync(ho.data(), order, uint64_t(n) * 4, hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    if (times) {
        float ms = 0;
        GB_HIP(hipEventElapsedTime(&ms, e0, e1));
        times->device_ms = ms;
        times->levels = uint32_t(levels.size());
        times->nodes = out_nodes;
    }
    for (size_t k = 0; k < roots.size(); k++) {
        const uint32_t i = roots[k];
        if (failed[i]) continue;  // built on the host
        OwnBuilt &o = out[i];
        const uint64_t cnt = 2 * uint64_t(root_inner[k]) + 1;
        o.nodes.assign(hn.begin() + node_base[i], hn.begin() + node_base[i] + cnt);
        o.order.resize(hm[i].n);
        for (uint32_t q = 0; q < hm[i].n; q++) o.order[q] = ho[hm[i].base + q] - hm[i].base;
    }
    return true;
}

}  // namespace gb
}  // namespace trt
