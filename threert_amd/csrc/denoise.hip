// The reference's edge-avoiding a-trous denoiser on the GPU: denoise.rs:10-74 (Dammertz et al. 2010; SVGF's 3 x 3
// variance prefilter, Schied et al. 2017), guided by first-hit albedo and normal. One thread a pixel; each
// iteration is two passes (the variance's 3 x 3 Gaussian, then the 5 x 5 wavelet step at stride 2^it), HBM-bound.
// exp is Moshier's Cephes expf (fmaf), powi(64) six squarings, everything else in the reference's order, so the
// result is the C restatement's to the bit.
#include <hip/hip_runtime.h>

#include "denoise_kernels.h"

namespace trt {
namespace {

__device__ __forceinline__ float expf_c(float x) {
    if (x < -87.3365447505f) return 0.0f;
    if (x > 88.7228391117f) return __builtin_inff();
    const float LOG2EF = 1.44269504088896341f, C1 = 0.693359375f, C2 = -2.12194440e-4f;
    const float fx = floorf(fmaf(x, LOG2EF, 0.5f));
    float r = fmaf(-fx, C1, x);
    r = fmaf(-fx, C2, r);
    const float z = r * r;
    float y = fmaf(1.9875691500e-4f, r, 1.3981999507e-3f);
    y = fmaf(y, r, 8.3334519073e-3f);
    y = fmaf(y, r, 4.1665795894e-2f);
    y = fmaf(y, r, 1.6666665459e-1f);
    y = fmaf(y, r, 5.0000001201e-1f);
    y = fmaf(y, z, r) + 1.0f;
    return ldexpf(y, static_cast<int>(fx));
}
__device__ __forceinline__ float lum(const float *c) { return c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f; }
__device__ __forceinline__ float pow64(float x) {
    const float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16;
    return x32 * x32;
}

__global__ void k_demod(uint32_t n, const float *color, const float *albedo, const float *variance, float *cur,
                        float *var) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float am[3];
    for (int k = 0; k < 3; k++) {
        am[k] = fmaxf(albedo[3 * i + k], 1e-3f);
        cur[3 * i + k] = color[3 * i + k] / am[k];
    }
    const float l = am[0] * 0.2126f + am[1] * 0.7152f + am[2] * 0.0722f;
    var[i] = variance[i] / (l * l);
}

__global__ void k_var3(uint32_t w, uint32_t h, const float *var, float *varf) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(w) * h) return;
    const int64_t x = int64_t(i % w), y = int64_t(i / w);
    float s = 0.0f, ws = 0.0f;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int64_t qx = x + dx, qy = y + dy;
            if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
            const float k = (dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f);
            s += k * var[qy * w + qx];
            ws += k;
        }
    varf[i] = s / ws;
}

__global__ void k_atrous(uint32_t w, uint32_t h, int64_t step, const float *cur, const float *var,
                         const float *varf, const float *normal, const float *albedo, float *nxt, float *nvar) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(w) * h) return;
    const float K[3] = {3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const int64_t x = int64_t(i % w), y = int64_t(i / w);
    const float *cp = cur + 3 * i, *np = normal + 3 * i, *ap = albedo + 3 * i;
    const float lp = lum(cp);
    const float sigma_l = 4.0f * sqrtf(fmaxf(varf[i], 0.0f)) + 1e-4f;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, wsum = 0.0f, vsum = 0.0f;
    for (int dy = -2; dy <= 2; dy++)
        for (int dx = -2; dx <= 2; dx++) {
            const int64_t qx = x + dx * step, qy = y + dy * step;
            if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
            const int64_t q = qy * w + qx;
            const float k = K[dx < 0 ? -dx : dx] * K[dy < 0 ? -dy : dy];
            const float *cq = cur + 3 * q, *nq = normal + 3 * q, *aq = albedo + 3 * q;
            const float wl = expf_c(-fabsf(lp - lum(cq)) / sigma_l);
            const float nd = np[0] * nq[0] + np[1] * nq[1] + np[2] * nq[2];
            const float wn = pow64(fmaxf(nd, 0.0f));
            const float d0 = ap[0] - aq[0], d1 = ap[1] - aq[1], d2 = ap[2] - aq[2];
            const float wa = expf_c(-(d0 * d0 + d1 * d1 + d2 * d2) / 0.01f);
            const float wt = k * wl * wn * wa;
            s0 += cq[0] * wt;
            s1 += cq[1] * wt;
            s2 += cq[2] * wt;
            vsum += wt * wt * var[q];
            wsum += wt;
        }
    if (wsum > 0.0f) {
        nxt[3 * i] = s0 / wsum;
        nxt[3 * i + 1] = s1 / wsum;
        nxt[3 * i + 2] = s2 / wsum;
        nvar[i] = vsum / (wsum * wsum);
    } else {
        nxt[3 * i] = cp[0];
        nxt[3 * i + 1] = cp[1];
        nxt[3 * i + 2] = cp[2];
        nvar[i] = var[i];
    }
}

__global__ void k_remod(uint32_t n, const float *cur, const float *albedo, float *out) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; k++) out[3 * i + k] = cur[3 * i + k] * fmaxf(albedo[3 * i + k], 1e-3f);
}

}  // namespace

int launch_atrous(uint32_t w, uint32_t h, const float *color, const float *albedo, const float *normal,
                  const float *variance, uint32_t iterations, float *out, float *scratch, hipStream_t s) {
    const uint32_t n = w * h;
    float *cur = scratch, *nxt = scratch + 3 * uint64_t(n), *var = scratch + 6 * uint64_t(n),
          *nvar = scratch + 7 * uint64_t(n), *varf = scratch + 8 * uint64_t(n);
    const dim3 grid((n + 255) / 256), block(256);
    hipLaunchKernelGGL(k_demod, grid, block, 0, s, n, color, albedo, variance, cur, var);
    for (uint32_t it = 0; it < iterations; it++) {
        hipLaunchKernelGGL(k_var3, grid, block, 0, s, w, h, var, varf);
        hipLaunchKernelGGL(k_atrous, grid, block, 0, s, w, h, int64_t(1) << it, cur, var, varf, normal, albedo, nxt,
                           nvar);
        float *t = cur;
        cur = nxt, nxt = t;
        t = var;
        var = nvar, nvar = t;
    }
    hipLaunchKernelGGL(k_remod, grid, block, 0, s, n, cur, albedo, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
