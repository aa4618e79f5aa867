// The FP64 check of a batch's hits, on the GPU: each reported hit is intersected again in double precision. The ray
// is placed in the hit's unit (translation, then the record's normal rows, as the walk does) and the hit triangle's
// stored v0, e1, e2 are tested with the same Moller-Trumbore steps (pathtrace.wgsl:742-760), every operation in
// double. The result says how far the FP32 distance is from the FP64 one, and which hits the FP64 test would not
// make (disputed rays: the FP32 test accepted a ray that passes within rounding of the triangle's edge).
#include "trace_kernels.h"

namespace trt {
namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ D3 d3(float4 v) { return {double(v.x), double(v.y), double(v.z)}; }
__device__ __forceinline__ D3 drows(D3 n0, D3 n1, D3 n2, D3 d) {
    return {n0.x * d.x + n1.x * d.y + n2.x * d.z, n0.y * d.x + n1.y * d.y + n2.y * d.z,
            n0.z * d.x + n1.z * d.y + n2.z * d.z};
}

// t64[i]: the FP64 distance to the hit triangle; 0 for a ray that hit nothing; -1 where the FP64 test misses the
// reported triangle; -2 where the hit names a triangle or unit outside the scene.
__global__ __launch_bounds__(256) void k_check64(const float4 *__restrict__ rays, const uint32_t *__restrict__ hits,
                                                 const float4 *__restrict__ leaves, const float4 *__restrict__ units,
                                                 uint64_t n, uint64_t leaf_count, uint64_t unit_count,
                                                 double *__restrict__ t64) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t tri = hits[6 * i + 1], inst = hits[6 * i + 4];
        if (tri == kNone) {
            t64[i] = 0.0;
            continue;
        }
        if (tri >= leaf_count || inst >= unit_count) {
            t64[i] = -2.0;
            continue;
        }
        const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
        const float4 u0 = units[4 * uint64_t(inst)];
        D3 o = dsub(D3{double(r0.x), double(r0.y), double(r0.z)}, d3(u0)), d{double(r1.x), double(r1.y), double(r1.z)};
        if ((__float_as_uint(u0.w) & kUnitShifted) == 0) {
            const D3 c0 = d3(units[4 * uint64_t(inst) + 1]), c1 = d3(units[4 * uint64_t(inst) + 2]),
                     c2 = d3(units[4 * uint64_t(inst) + 3]);
            o = drows(c0, c1, c2, o);
            d = drows(c0, c1, c2, d);
        }
        const D3 v0 = d3(leaves[3 * uint64_t(tri)]), e1 = d3(leaves[3 * uint64_t(tri) + 1]),
                 e2 = d3(leaves[3 * uint64_t(tri) + 2]);
        const D3 pv = dcross(d, e2);
        const double det = ddot(e1, pv);
        double t = -1.0;
        if (det != 0.0) {
            const double inv = 1.0 / det;
            const D3 tv = dsub(o, v0);
            const double u = ddot(tv, pv) * inv;
            const D3 qv = dcross(tv, e1);
            const double v = ddot(d, qv) * inv;
            const double tt = ddot(e2, qv) * inv;
            if (u >= 0.0 && u <= 1.0 && v >= 0.0 && u + v <= 1.0 && tt > 0.0) t = tt;
        }
        t64[i] = t;
    }
}

}  // namespace

int launch_check64(const float4 *rays, const uint32_t *hits, const float4 *leaves, const float4 *units, uint64_t n,
                   uint64_t leaf_count, uint64_t unit_count, double *t64, int grid, hipStream_t s) {
    if (n == 0) return 0;
    const uint64_t need = (n + 255) / 256;
    const int g = int(need < uint64_t(grid) ? need : uint64_t(grid));
    hipLaunchKernelGGL(k_check64, dim3(g), dim3(256), 0, s, rays, hits, leaves, units, n, leaf_count, unit_count,
                       t64);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
