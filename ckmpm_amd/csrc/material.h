// Deformation update and Kirchhoff stress of the four constitutive models (FP32).
#pragma once

#include "common.h"
#include "svd.h"

namespace ckmpm {

__device__ __forceinline__ void MatMul3(const float (&A)[9], const float (&B)[9], float (&C)[9]) {
    float T[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) T[i * 3 + k] = A[i * 3 + 0] * B[0 * 3 + k] + A[i * 3 + 1] * B[1 * 3 + k] + A[i * 3 + 2] * B[2 * 3 + k];
#pragma unroll
    for (int i = 0; i < 9; ++i) C[i] = T[i];
}

__device__ __forceinline__ float Det3(const float (&a)[9]) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// Inverse of a symmetric 3x3 matrix (adjugate / determinant), data_type.cuh:369-387.
__device__ __forceinline__ void InverseSym3(const float (&a)[9], float (&r)[9]) {
    const float c00 = a[4] * a[8] - a[5] * a[7];
    const float c01 = a[2] * a[7] - a[1] * a[8];
    const float c02 = a[1] * a[5] - a[2] * a[4];
    const float c11 = a[0] * a[8] - a[2] * a[6];
    const float c12 = a[2] * a[3] - a[0] * a[5];
    const float c22 = a[0] * a[4] - a[1] * a[3];
    const float id = 1.0f / (a[0] * c00 + a[1] * (a[5] * a[6] - a[3] * a[8]) + a[2] * (a[3] * a[7] - a[4] * a[6]));
    r[0] = id * c00;
    r[1] = r[3] = id * c01;
    r[2] = r[6] = id * c02;
    r[4] = id * c11;
    r[5] = r[7] = id * c12;
    r[8] = id * c22;
}

// Fixed corotated, mpm_material.cuh:318-357: tau = U diag(2 mu s (s - 1) + lambda J (J - 1)) U^T.
__device__ __forceinline__ void StressFixedCorotated(float lambda, float mu, const float (&F)[9], float (&tau)[9]) {
    Svd3 d;
    SignedSvd3(F, d);
    const float J = d.s[0] * d.s[1] * d.s[2];
    const float vol = lambda * J * (J - 1.0f);
    float k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k[a] = 2.0f * mu * d.s[a] * (d.s[a] - 1.0f) + vol;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = d.U[i * 3 + 0] * k[0] * d.U[j * 3 + 0] + d.U[i * 3 + 1] * k[1] * d.U[j * 3 + 1] +
                            d.U[i * 3 + 2] * k[2] * d.U[j * 3 + 2];
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// Neo-Hookean (closed form): tau = mu (F F^T - I) + lambda ln(J) I.
__device__ __forceinline__ void StressNeoHookean(float lambda, float mu, const float (&F)[9], float (&tau)[9]) {
    const float lnJ = logf(Det3(F));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float b = F[i * 3 + 0] * F[j * 3 + 0] + F[i * 3 + 1] * F[j * 3 + 1] + F[i * 3 + 2] * F[j * 3 + 2];
            const float v = mu * (b - (i == j ? 1.0f : 0.0f)) + (i == j ? lambda * lnJ : 0.0f);
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// Drucker-Prager StVK-Hencky return map, mpm_material.cuh:360-464 (cohesion 0, beta 1,
// volume correction).  Projects F in place and updates logJp.
__device__ __forceinline__ void StressDruckerPrager(float lambda, float mu, float sin_phi, float (&F)[9], float& logJp,
                                                    float (&tau)[9]) {
    const float alpha = 0.816496580927726f * 2.0f * sin_phi / (3.0f - sin_phi);  // sqrt(2/3) 2 sin / (3 - sin)
    const float smu = 2.0f * mu;
    Svd3 d;
    SignedSvd3(F, d);
    float eps[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) eps[a] = logf(fmaxf(fabsf(d.s[a]), 1e-4f));
    const float eps_sum = eps[0] + eps[1] + eps[2];
    const float eps_tr = eps_sum + logJp;
    float eh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) eh[a] = eps[a] - eps_tr * (1.0f / 3.0f);
    const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    float H[3] = {eps[0], eps[1], eps[2]};
    bool tension = eps_tr >= 0.0f;
    if (tension) {
        H[0] = H[1] = H[2] = 0.0f;
        logJp = eps_sum + logJp;
    } else if (mu != 0.0f) {
        logJp = 0.0f;
        const float dg = ehn + (3.0f * lambda + smu) / smu * eps_tr * alpha;
        if (dg > 0.0f) {
            const float r = dg / ehn;
#pragma unroll
            for (int a = 0; a < 3; ++a) H[a] = eps[a] - r * eh[a];
        }
    }
    float ns[3], ph[3];
    const float htr = H[0] + H[1] + H[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ns[a] = expf(H[a]);
        ph[a] = (smu * H[a] + lambda * htr) / ns[a];  // log(ns) = H
    }
    // F = U diag(ns) V^T ; tau = U diag(ph) V^T F^T = U diag(ph * ns) U^T
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F[i * 3 + j] = d.U[i * 3 + 0] * ns[0] * d.V[j * 3 + 0] + d.U[i * 3 + 1] * ns[1] * d.V[j * 3 + 1] +
                           d.U[i * 3 + 2] * ns[2] * d.V[j * 3 + 2];
    float k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k[a] = ph[a] * ns[a];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = d.U[i * 3 + 0] * k[0] * d.U[j * 3 + 0] + d.U[i * 3 + 1] * k[1] * d.U[j * 3 + 1] +
                            d.U[i * 3 + 2] * k[2] * d.U[j * 3 + 2];
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// Weakly compressible fluid, mpm_algorithm.cuh:539-547 and mpm_material.cuh:466-477:
// J <- J (1 + dt tr C); p = K (J^-gamma - 1); tau = J (nu (C + C^T) - p I).
// The state holds q = J - 1 (the reference stores J): near J = 1 an FP32 J keeps only
// ~6e-8 absolute resolution, which is a large relative error of J - 1 and hence of the
// pressure; q keeps full relative precision, and p = K expm1(-gamma log1p(q)).
__device__ __forceinline__ void StressFluid(const MaterialDev& m, const float (&C)[9], float dt, float& q, float (&tau)[9]) {
    q += (C[0] + C[4] + C[8]) * dt * (1.0f + q);
    const float J = 1.0f + q;
    const float p = m.bulk * expm1f(-m.gamma * log1pf(q));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = J * (m.viscosity * (C[i * 3 + j] + C[j * 3 + i]) - (i == j ? p : 0.0f));
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

}  // namespace ckmpm
