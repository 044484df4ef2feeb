// Particle-to-grid assignment shared by the TILD kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace mft {

// spline_get_weights (ps_grid_utils.cu:17-67), FP32, with the p = 4 term
// s^3 -> s^4 of ps_grid_utils.cu:57 kept as the reference computes it.
__device__ __forceinline__ void spline_weights(float sx, int p, float* W)
{
    float sx2 = sx * sx;
    if (p == 3) {
        float sx3 = sx2 * sx;
        W[0] = (1.f - 6.f * sx + 12.f * sx2 - 8.f * sx3) * (1.f / 48.f);
        W[1] = (23.f - 30.f * sx - 12.f * sx2 + 24.f * sx3) * (1.f / 48.f);
        W[2] = (23.f + 30.f * sx - 12.f * sx2 - 24.f * sx3) * (1.f / 48.f);
        W[3] = (1.f + 6.f * sx + 12.f * sx2 + 8.f * sx3) * (1.f / 48.f);
    } else if (p == 1) {
        W[0] = 0.5f - sx;
        W[1] = 0.5f + sx;
    } else if (p == 2) {
        W[0] = 0.125f - 0.5f * sx + 0.5f * sx2;
        W[1] = 0.75f - sx2;
        W[2] = 0.125f + 0.5f * sx + 0.5f * sx2;
    } else if (p == 4) {
        float sx3 = sx2 * sx2;
        float sx4 = sx2 * sx2;
        W[0] = (1.f - 8.f * sx + 24.f * sx2 - 32.f * sx3 + 16.f * sx4) * (1.f / 384.f);
        W[1] = (19.f - 44.f * sx + 24.f * sx2 + 16.f * sx3 - 16.f * sx4) * (1.f / 96.f);
        W[2] = (115.f - 120.f * sx2 + 48.f * sx4) * (1.f / 192.f);
        W[3] = (19.f + 44.f * sx + 24.f * sx2 - 16.f * sx3 - 16.f * sx4) * (1.f / 96.f);
        W[4] = (1.f + 8.f * sx + 24.f * sx2 + 32.f * sx3 + 16.f * sx4) * (1.f / 384.f);
    } else {
        W[0] = 1.f;
    }
}

// Base grid index of one coordinate exactly as ps_grid_utils.cu:105-113 forms
// it (FP32, correctly rounded division, no contraction) and the spline
// weights of the p+1 points g - p/2 .. g - p/2 + p.
__device__ __forceinline__ int base_index_weights(float x, float dx, int p, float* W)
{
    int g;
    float gdx;
    if ((p & 1) == 0) {
        float s = __fadd_rn(x, __fmul_rn(0.5f, dx));
        g = (int)__fdiv_rn(s, dx);
        gdx = fmaf(-(float)g, dx, x);
    } else {
        g = (int)__fdiv_rn(x, dx);
        gdx = fmaf(-((float)g + 0.5f), dx, x);
    }
    spline_weights(gdx / dx, p, W);
    return g;
}

// Same index, weights from s = gdx * (1/dx): only the index needs the
// correctly rounded division; the weights are compared within tolerance.
__device__ __forceinline__ int base_index_weights_fast(float x, float dx, float inv_dx, int p, float* W)
{
    int g;
    float gdx;
    if ((p & 1) == 0) {
        float s = __fadd_rn(x, __fmul_rn(0.5f, dx));
        g = (int)__fdiv_rn(s, dx);
        gdx = fmaf(-(float)g, dx, x);
    } else {
        g = (int)__fdiv_rn(x, dx);
        gdx = fmaf(-((float)g + 0.5f), dx, x);
    }
    spline_weights(gdx * inv_dx, p, W);
    return g;
}

__device__ __forceinline__ int base_index(float x, float dx, int p)
{
    if ((p & 1) == 0) return (int)__fdiv_rn(__fadd_rn(x, __fmul_rn(0.5f, dx)), dx);
    return (int)__fdiv_rn(x, dx);
}

// i mod n in [0, n) (n > 0).  Indices within one period of [0, n) -- every
// base index of a coordinate inside the box, and the stencil and window
// offsets from one -- take a compare and an add; only lanes further out
// (a branch the wave skips when it has none) divide.
__device__ __forceinline__ int wrap(int i, int n)
{
    int r = i < 0 ? i + n : (i >= n ? i - n : i);
    if ((unsigned)r >= (unsigned)n) {
        r = i % n;
        r = r < 0 ? r + n : r;
    }
    return r;
}

} // namespace mft
