// Compact-kernel stencils of the two CK-MPM sub-grids (mpm_kernel.cuh:91-114):
// base node, fractional offset t and node-0 weight w0 = 1 - t + sin(2 pi t)/(2 pi)
// per axis; node 1 of an axis has weight 1 - w0.
#pragma once
#include <hip/hip_runtime.h>

namespace ckmpm {

constexpr float kInv2Pi = 0.15915494309189535f;

// sin(2 pi t) / (2 pi) for t in [0, 1): with u = t - 1/2, sin(2 pi t) = -sin(2 pi u);
// |u| > 1/4 folds to 1/2 sgn(u) - u (sin(pi - a) = sin a), so x = 2 pi u' lies in
// [-pi/2, pi/2], where the Taylor series of sin x / x through x^10 is within 4e-8
// (the next term, x^12 / 13!).  Result = -u' (1 - z/3! + z^2/5! - ... - z^5/11!), z = x^2.
__device__ __forceinline__ float SinTwoPiFrac(float t) {
    float u = t - 0.5f;
    if (u > 0.25f) u = 0.5f - u;
    else if (u < -0.25f) u = -0.5f - u;
    const float z = 39.47841760435743f * u * u;  // (2 pi u)^2
    float p = -2.505210838544172e-08f;           // -1 / 11!
    p = p * z + 2.755731922398589e-06f;          //  1 / 9!
    p = p * z - 1.984126984126984e-04f;          // -1 / 7!
    p = p * z + 8.333333333333333e-03f;          //  1 / 5!
    p = p * z - 1.666666666666667e-01f;          // -1 / 3!
    p = p * z + 1.0f;
    return -u * p;
}

// Stencil of one sub-grid at an FP32 position (xs = x * invDx): base cell c,
// fractional offset t and node-0 weight w0 per axis.  mpm_algorithm.cuh:1555-1569,
// mpm_kernel.cuh:91-114 (w0 = 1 - t + sin(2 pi t) / (2 pi)).
struct Stencil32 {
    int c[3];
    float t[3];
    float w0[3];
};
template <int g>
__device__ __forceinline__ void MakeStencil(const float (&xs)[3], Stencil32& s) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float shift = g == 0 ? 0.25f : -0.25f;
        const int c = static_cast<int>(floorf(__fadd_rn(xs[a], shift)));
        const float t = __fadd_rn(xs[a] - static_cast<float>(c), shift);
        s.c[a] = c;
        s.t[a] = t;
        s.w0[a] = 1.0f - t + SinTwoPiFrac(t);
    }
}
// Both sub-grid stencils from one sine per axis.  With x/dx < 2^22 every step is exact
// in FP32: grid 1 has t1 = t0 - 1/2 (same base cell) when t0 >= 1/2, else t1 = t0 + 1/2
// (base cell one lower), which equals MakeStencil<1>'s c and t bit for bit; and
// sin(2 pi t1) = -sin(2 pi t0).
__device__ __forceinline__ void MakeStencilPair(const float (&xs)[3], Stencil32& s0, Stencil32& s1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int c = static_cast<int>(floorf(__fadd_rn(xs[a], 0.25f)));
        const float t = __fadd_rn(xs[a] - static_cast<float>(c), 0.25f);
        const float sn = SinTwoPiFrac(t);
        const bool hi = t >= 0.5f;
        s0.c[a] = c;
        s0.t[a] = t;
        s0.w0[a] = 1.0f - t + sn;
        const float t1 = hi ? t - 0.5f : t + 0.5f;
        s1.c[a] = hi ? c : c - 1;
        s1.t[a] = t1;
        s1.w0[a] = 1.0f - t1 - sn;
    }
}
__device__ __forceinline__ float NodeWeight(const Stencil32& s, int dx, int dy, int dz) {
    const float wx = dx ? 1.0f - s.w0[0] : s.w0[0];
    const float wy = dy ? 1.0f - s.w0[1] : s.w0[1];
    const float wz = dz ? 1.0f - s.w0[2] : s.w0[2];
    return wx * wy * wz;
}

}  // namespace ckmpm
