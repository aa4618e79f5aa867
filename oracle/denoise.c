/*
 * Oracle (test infrastructure only): the reference's CPU denoiser restated, denoise.rs:10-74 `denoise`
 * (edge-avoiding a-trous wavelet, Dammertz et al. 2010, guided by albedo and normal, luminance edge-stopping width
 * from the 3 x 3-filtered variance). glam Vec3 is scalar f32: dot = (x x' + y y') + z z'. exp by or_expf (Moshier's
 * Cephes expf: range reduction by ln 2 in two parts, degree-6 polynomial, fmaf), powi(64) by six squarings.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_render.h"

float or_expf(float x) {
    if (x < -87.3365447505f) return 0.0f;
    if (x > 88.7228391117f) return INFINITY;
    const float LOG2EF = 1.44269504088896341f, C1 = 0.693359375f, C2 = -2.12194440e-4f;
    float fx = floorf(fmaf(x, LOG2EF, 0.5f));
    float r = fmaf(-fx, C1, x);
    r = fmaf(-fx, C2, r);
    float z = r * r;
    float y = fmaf(1.9875691500e-4f, r, 1.3981999507e-3f);
    y = fmaf(y, r, 8.3334519073e-3f);
    y = fmaf(y, r, 4.1665795894e-2f);
    y = fmaf(y, r, 1.6666665459e-1f);
    y = fmaf(y, r, 5.0000001201e-1f);
    y = fmaf(y, z, r) + 1.0f;
    return ldexpf(y, (int)fx);
}

static inline float lum(const float *c) { return c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f; }
static inline float pow64(float x) {
    float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16;
    return x32 * x32;
}

void or_denoise_atrous(uint32_t w, uint32_t h, const float *color, const float *albedo, const float *normal,
                       const float *variance, uint32_t iterations, float *out) {
    const uint64_t n = (uint64_t)w * h;
    const float eps = 1e-3f;
    float *cur = (float *)malloc(n * 3 * sizeof(float)), *nxt = (float *)malloc(n * 3 * sizeof(float));
    float *var = (float *)malloc(n * sizeof(float)), *nvar = (float *)malloc(n * sizeof(float));
    float *varf = (float *)malloc(n * sizeof(float));
    for (uint64_t i = 0; i < n; i++) {
        float am[3];
        for (int k = 0; k < 3; k++) {
            am[k] = fmaxf(albedo[3 * i + k], eps);
            cur[3 * i + k] = color[3 * i + k] / am[k];
        }
        float l = am[0] * 0.2126f + am[1] * 0.7152f + am[2] * 0.0722f;
        var[i] = variance[i] / (l * l);
    }
    static const float K[3] = {3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    for (uint32_t it = 0; it < iterations; it++) {
        const int64_t step = (int64_t)1 << it;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            int64_t x = i % w, y = i / w;
            float s = 0.0f, ws = 0.0f;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int64_t qx = x + dx, qy = y + dy;
                    if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
                    float k = (dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f);
                    s += k * var[qy * w + qx];
                    ws += k;
                }
            varf[i] = s / ws;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            int64_t x = i % w, y = i / w;
            const float *cp = cur + 3 * i, *np = normal + 3 * i, *ap = albedo + 3 * i;
            float lp = lum(cp);
            float sigma_l = 4.0f * sqrtf(fmaxf(varf[i], 0.0f)) + 1e-4f;
            float sum[3] = {0, 0, 0}, wsum = 0.0f, vsum = 0.0f;
            for (int dy = -2; dy <= 2; dy++)
                for (int dx = -2; dx <= 2; dx++) {
                    int64_t qx = x + dx * step, qy = y + dy * step;
                    if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
                    int64_t q = qy * w + qx;
                    float k = K[abs(dx)] * K[abs(dy)];
                    const float *cq = cur + 3 * q, *nq = normal + 3 * q, *aq = albedo + 3 * q;
                    float wl = or_expf(-fabsf(lp - lum(cq)) / sigma_l);
                    float nd = np[0] * nq[0] + np[1] * nq[1] + np[2] * nq[2];
                    float wn = pow64(fmaxf(nd, 0.0f));
                    float da[3] = {ap[0] - aq[0], ap[1] - aq[1], ap[2] - aq[2]};
                    float wa = or_expf(-(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]) / 0.01f);
                    float wt = k * wl * wn * wa;
                    for (int c = 0; c < 3; c++) sum[c] += cq[c] * wt;
                    vsum += wt * wt * var[q];
                    wsum += wt;
                }
            if (wsum > 0.0f) {
                for (int c = 0; c < 3; c++) nxt[3 * i + c] = sum[c] / wsum;
                nvar[i] = vsum / (wsum * wsum);
            } else {
                for (int c = 0; c < 3; c++) nxt[3 * i + c] = cp[c];
                nvar[i] = var[i];
            }
        }
        float *t = cur;
        cur = nxt;
        nxt = t;
        t = var;
        var = nvar;
        nvar = t;
    }
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) out[3 * i + k] = cur[3 * i + k] * fmaxf(albedo[3 * i + k], eps);
    free(cur);
    free(nxt);
    free(var);
    free(nvar);
    free(varf);
}
