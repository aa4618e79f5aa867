// Internal types shared by the host code and the HIP kernels of libthreert.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace trt {

// bvh.rs:6-15 GpuNode (32 bytes).
struct Node {
    float bmin[3];
    uint32_t left_or_first;
    float bmax[3];
    uint32_t count;
};
static_assert(sizeof(Node) == 32, "GpuNode is 32 bytes");

// indexed.rs:19-26 Leaf (48 bytes): (v0, section id), (e1, index in section), (e2, 0).
struct Leaf {
    float r[3][4];
};
static_assert(sizeof(Leaf) == 48, "Leaf is 48 bytes");

// scene.rs:229-234 GpuInstance (160 bytes).
struct Instance {
    float m[3][4];
    float n[3][4];
    uint32_t info[4];
    float p[3][4];
};
static_assert(sizeof(Instance) == 160, "GpuInstance is 160 bytes");

struct Tree {
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    uint32_t depth = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

constexpr uint32_t kBvhStack = 48;   // bvh.rs:56
constexpr uint32_t kStack = 64;      // units.rs:30
constexpr uint32_t kTopDepth = 19;   // units.rs:33
constexpr uint32_t kMeshDepth = 45;  // units.rs:35
constexpr uint32_t kMaxUnits = 1u << 18;  // units.rs:37
constexpr uint32_t kMaxLeaf = 16;    // units.rs:39
constexpr uint32_t kUnitShadow = 1;  // units.rs:337
constexpr uint32_t kUnitShifted = 2; // units.rs:338
constexpr uint32_t kInstCutout = 16, kInstGlass = 32;            // scene.rs INST_CUTOUT / INST_GLASS
constexpr uint32_t kInstBits = 4u | 16u | 32u | 64u | 128u;     // scene.rs INST_BITS
constexpr float kInf = 3.0e38f;      // pathtrace.wgsl INF
constexpr uint32_t kNone = 0xffffffffu;

inline uint32_t to_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float from_bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---- the library's 4-wide quantised tree (64-byte nodes) -------------------------------------------------------
// A child reference: bits 31:30 say what it is.
//   00  a wide node: its index
//   01  a run of triangles: bits 29:26 count - 1, bits 25:0 the first, counted from the unit's first triangle
//   10  a unit (top tree): bits 29:0 its record
//   11  no child (0xffffffff)
constexpr uint32_t kRefKindShift = 30;
constexpr uint32_t kRefNode = 0u, kRefTris = 1u, kRefUnit = 2u;
constexpr uint32_t kRefTriFirstBits = 26;
constexpr uint32_t kRefTriFirstMask = (1u << kRefTriFirstBits) - 1;
inline uint32_t ref_node(uint32_t i) { return i; }
inline uint32_t ref_tris(uint32_t first, uint32_t count) {
    return (kRefTris << kRefKindShift) | ((count - 1) << kRefTriFirstBits) | first;
}
inline uint32_t ref_unit(uint32_t rec) { return (kRefUnit << kRefKindShift) | rec; }

// Child c's box is origin + q * 2^(e - 127) per axis, q its byte of qlo/qhi; boxes are widened outward by one step
// beyond the exact children's boxes so a walk's rounding never loses a hit.
struct alignas(64) WideNode {
    float origin[3];
    uint8_t exp[3];   // biased exponent of the axis step (a power of two)
    uint8_t nchild;
    uint32_t qlo[3];  // byte c: child c's low bound on the axis
    uint32_t qhi[3];
    uint32_t child[4];
    uint32_t pad[2];
};
static_assert(sizeof(WideNode) == 64, "a wide node is 64 bytes");

// A unit as the wide walk reads it (64 bytes; the 160-byte record's translation, normal rows, flags, root, first).
struct alignas(64) WideUnit {
    float t[3];
    uint32_t flags;
    float n0[3];
    uint32_t root;   // the mesh tree's wide root reference
    float n1[3];
    uint32_t first;  // the mesh's first triangle
    float n2[3];
    uint32_t pad;
};
static_assert(sizeof(WideUnit) == 64, "a wide unit is 64 bytes");

struct WideScene {
    uint64_t node_base = 0;  // wide nodes made elsewhere (the device's mesh trees), ahead of `nodes`
    std::vector<WideNode> nodes;
    std::vector<WideUnit> units;
    // The leaves in the wide trees' own order, each mesh's within its range; e2.w holds the reference leaf index.
    std::vector<Leaf> wleaves;
    uint32_t top_ref = kNone;
    uint32_t max_stack = 0;
    float top_box[6] = {0, 0, 0, 0, 0, 0};
};

// Builders (build.cpp).
bool build_tris(const float *tris, uint64_t n, uint32_t max_depth, Tree &out);
void build_median(const float *boxes, uint64_t n, uint32_t max_leaf, Tree &out);
void build_sah(const float *boxes, uint64_t n, uint32_t max_depth, Tree &out);
int check_tree(const Node *nodes, uint64_t count, uint32_t max_depth, uint64_t items, uint32_t max_leaf,
               uint64_t records_n, std::string *why);
bool mesh_tree(const float *tris, uint64_t n, Tree &out, std::string *why);
void unit_box(const float *world, const float *lo, const float *hi, float *out);
bool top_tree(const float *boxes, const uint32_t *records, uint64_t n, Tree &out, std::string *why);
void unit_record(Instance &r, uint32_t root, bool cast_shadow);

// The library's own mesh trees (wide.cpp OwnTree on the host, gbuild.hip on the device: the same bytes). Binned SAH
// over the triangles' boxes, 16 bins an axis, leaves of at most kOwnLeaf triangles, split at the median past
// kOwnDepth.
constexpr uint32_t kOwnLeaf = 4;
constexpr int kOwnBins = 16;
constexpr uint32_t kOwnDepth = 64;
struct OwnBuilt {
    std::vector<Node> nodes;      // nodes[0] is the root; children at left_or_first and left_or_first + 1 (absolute)
    std::vector<uint32_t> order;  // tree position -> the triangle's index in its mesh; leaves name positions
};
struct MeshRef {
    uint32_t root, first, n;  // the reference mesh tree's root, its first leaf and its triangle count
};
void own_tree_host(const Leaf *leaves, uint32_t first, uint32_t n, OwnBuilt &out);

// The wide tree of an uploaded scene (wide.cpp). The plan: the top tree's records in walk order, the distinct mesh
// trees they name (in record order) and each record's mesh, after checking the trees (units.rs check_top /
// check_mesh on the buffers). `pre` (optional) holds own trees already built for the plan's meshes (an entry with no
// nodes is built on the host).
struct WidePlan {
    std::vector<uint32_t> records;
    std::vector<MeshRef> meshes;
    std::vector<uint32_t> mesh_of;
};
bool wide_plan(const Node *nodes, uint64_t node_count, const Instance *inst, uint64_t inst_count,
               uint64_t leaf_count, uint32_t top_first, uint32_t unit_count, WidePlan &plan, std::string *why);
// `mesh_wide` (optional): the meshes' wide nodes were made elsewhere (the device), mesh_wide->count of them ahead
// of the top tree's; out.node_base is then that count and out.nodes holds the top tree's only, out.wleaves is left
// empty, and each mesh's root reference and stack bound come from mesh_wide.
struct MeshWide {
    uint64_t count = 0;
    std::vector<uint32_t> ref, bound;
};
bool build_wide(const Node *nodes, uint64_t node_count, const Leaf *leaves, const Instance *inst, uint64_t inst_count,
                uint64_t leaf_count, uint32_t top_first, uint32_t unit_count, WideScene &out, std::string *why,
                const WidePlan *plan = nullptr, const std::vector<OwnBuilt> *pre = nullptr,
                const MeshWide *mesh_wide = nullptr);

}  // namespace trt
