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
    // v_rcp_f32 (1 ulp) instead of an IEEE division sequence
    const float id = __builtin_amdgcn_rcpf(a[0] * c00 + a[1] * (a[5] * a[6] - a[3] * a[8]) + a[2] * (a[3] * a[7] - a[4] * a[6]));
    r[0] = id * c00;
    r[1] = r[3] = id * c01;
    r[2] = r[6] = id * c02;
    r[4] = id * c11;
    r[5] = r[7] = id * c12;
    r[8] = id * c22;
}

// The solid state holds E = F - I (the reference stores F): an FP32 F near the identity
// keeps only ~6e-8 absolute resolution, a large relative error of small strains; E keeps
// full relative precision and the stresses below are evaluated from it without forming
// differences of numbers near 1.

// det(I + E) - 1 = tr E + ((tr E)^2 - tr(E^2)) / 2 + det E
__device__ __forceinline__ float DetIPlusEMinus1(const float (&E)[9]) {
    const float tr = E[0] + E[4] + E[8];
    const float tr2 = E[0] * E[0] + E[4] * E[4] + E[8] * E[8] + 2.0f * (E[1] * E[3] + E[2] * E[6] + E[5] * E[7]);
    return tr + 0.5f * (tr * tr - tr2) + Det3(E);
}

// Fixed corotated, mpm_material.cuh:318-357: tau = U diag(2 mu s (s - 1) + lambda J (J - 1)) U^T.
__device__ __forceinline__ void StressFixedCorotated(float lambda, float mu, const float (&E)[9], float (&tau)[9]) {
    Svd3 sv;
    float d[3];
    SignedSvdStrain(E, sv, d);
    const float jm1 = d[0] + d[1] + d[2] + d[0] * d[1] + d[0] * d[2] + d[1] * d[2] + d[0] * d[1] * d[2];
    const float vol = lambda * (1.0f + jm1) * jm1;
    float k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k[a] = 2.0f * mu * sv.s[a] * d[a] + vol;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = sv.U[i * 3 + 0] * k[0] * sv.U[j * 3 + 0] + sv.U[i * 3 + 1] * k[1] * sv.U[j * 3 + 1] +
                            sv.U[i * 3 + 2] * k[2] * sv.U[j * 3 + 2];
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// log(1 + d) and exp(x) of the small principal strains: for |d|, |x| <= 1/32 their
// Taylor series (the terms fall by 1/32 or faster each, so 7 terms leave under 2^-35
// relative), otherwise the library functions, which cost several times as many
// instructions.  The lanes of a wavefront take the series together whenever their
// strains are all small.
__device__ __forceinline__ float Log1pSmall(float d) {
    if (fabsf(d) <= 0.03125f) {
        float p = 1.0f / 7.0f;  // d (1 - d/2 + d^2/3 - ... + d^6/7)
        p = fmaf(p, d, -1.0f / 6.0f);
        p = fmaf(p, d, 0.2f);
        p = fmaf(p, d, -0.25f);
        p = fmaf(p, d, 1.0f / 3.0f);
        p = fmaf(p, d, -0.5f);
        p = fmaf(p, d, 1.0f);
        return d * p;
    }
    return log1pf(d);
}
__device__ __forceinline__ float ExpSmall(float x) {
    if (fabsf(x) <= 0.03125f) {
        float p = 1.0f / 720.0f;  // 1 + x + x^2/2 + ... + x^6/720
        p = fmaf(p, x, 1.0f / 120.0f);
        p = fmaf(p, x, 1.0f / 24.0f);
        p = fmaf(p, x, 1.0f / 6.0f);
        p = fmaf(p, x, 0.5f);
        p = fmaf(p, x, 1.0f);
        return fmaf(p, x, 1.0f);
    }
    return expf(x);
}

// Neo-Hookean (closed form): tau = mu (F F^T - I) + lambda ln(J) I,
// F F^T - I = E + E^T + E E^T.
__device__ __forceinline__ void StressNeoHookean(float lambda, float mu, const float (&E)[9], float (&tau)[9]) {
    const float lnJ = Log1pSmall(DetIPlusEMinus1(E));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float b = E[i * 3 + j] + E[j * 3 + i] + E[i * 3 + 0] * E[j * 3 + 0] + E[i * 3 + 1] * E[j * 3 + 1] +
                            E[i * 3 + 2] * E[j * 3 + 2];
            const float v = mu * b + (i == j ? lambda * lnJ : 0.0f);
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// Drucker-Prager StVK-Hencky return map, mpm_material.cuh:360-464 (cohesion 0, beta 1,
// volume correction).  Projects E = F - I in place and updates logJp.  In the elastic
// branch (deltaGamma <= 0) the reference rebuilds F = U diag(exp eps) V^T with
// eps = log max(|s|, 1e-4) (:385,417-443): that is F itself only when every singular
// value is at least 1e-4 (s_2 > 0); otherwise the rebuild clamps |s| at 1e-4 and
// un-inverts an inverted F, so E is rebuilt then and kept as is only in the identity
// case.  With mu == 0 and eps_tr < 0 no branch runs and F stays as it is (:412).
__device__ __forceinline__ void StressDruckerPrager(float lambda, float mu, float sin_phi, float (&E)[9], float& logJp,
                                                    float (&tau)[9]) {
    const float alpha = 0.816496580927726f * 2.0f * sin_phi / (3.0f - sin_phi);  // sqrt(2/3) 2 sin / (3 - sin)
    const float smu = 2.0f * mu;
    Svd3 sv;
    float d[3];
    SignedSvdStrain(E, sv, d);
    float eps[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) eps[a] = sv.s[a] >= 1e-4f ? Log1pSmall(d[a]) : logf(fmaxf(fabsf(sv.s[a]), 1e-4f));
    const float eps_sum = eps[0] + eps[1] + eps[2];
    const float eps_tr = eps_sum + logJp;
    float eh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) eh[a] = eps[a] - eps_tr * (1.0f / 3.0f);
    const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    float H[3] = {eps[0], eps[1], eps[2]};
    bool rebuild = true, keep = false;  // keep: mu == 0 branch, F unchanged
    if (eps_tr >= 0.0f) {
        H[0] = H[1] = H[2] = 0.0f;
        logJp = eps_sum + logJp;
    } else if (mu != 0.0f) {
        logJp = 0.0f;
        const float dg = ehn + (3.0f * lambda + smu) / smu * eps_tr * alpha;
        if (dg > 0.0f) {
            const float r = dg / ehn;
#pragma unroll
            for (int a = 0; a < 3; ++a) H[a] = eps[a] - r * eh[a];
        } else {
            // H = eps: the rebuild is an identity unless a singular value is clamped
            rebuild = !(sv.s[0] >= 1e-4f && sv.s[1] >= 1e-4f && sv.s[2] >= 1e-4f);
        }
    } else {
        rebuild = false;
        keep = true;
    }
    const float htr = H[0] + H[1] + H[2];
    float k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k[a] = smu * H[a] + lambda * htr;  // (2 mu log ns + lambda tr log ns) / ns * ns
    if (rebuild || keep) {  // the new singular values ns = exp H are needed
        float ns[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ns[a] = ExpSmall(H[a]);
            // F kept with singular values s: tau = U diag(phat s) U^T
            if (keep) k[a] *= sv.s[a] / ns[a];
        }
        if (rebuild) {
            // E = U diag(ns) V^T - I
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    E[i * 3 + j] = sv.U[i * 3 + 0] * ns[0] * sv.V[j * 3 + 0] + sv.U[i * 3 + 1] * ns[1] * sv.V[j * 3 + 1] +
                                   sv.U[i * 3 + 2] * ns[2] * sv.V[j * 3 + 2] - (i == j ? 1.0f : 0.0f);
        }
    }
    // tau = U diag(phat) V^T F^T = U diag(phat * ns) U^T
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = sv.U[i * 3 + 0] * k[0] * sv.U[j * 3 + 0] + sv.U[i * 3 + 1] * k[1] * sv.U[j * 3 + 1] +
                            sv.U[i * 3 + 2] * k[2] * sv.U[j * 3 + 2];
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// Hencky strain of F = I + E from its signed SVD: eps_i = log s_i (log1p of the
// precise s_i - 1), with |s_i| clamped at 1e-4 as in the Drucker-Prager model (the
// reference's logf of a non-positive singular value is NaN).
__device__ __forceinline__ void HenckyStrain(const Svd3& sv, const float (&d)[3], float (&eps)[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) eps[a] = sv.s[a] >= 1e-4f ? Log1pSmall(d[a]) : logf(fmaxf(fabsf(sv.s[a]), 1e-4f));
}

// E = U diag(ns) V^T - I
__device__ __forceinline__ void RebuildStrain(const Svd3& sv, const float (&ns)[3], float (&E)[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            E[i * 3 + j] = sv.U[i * 3 + 0] * ns[0] * sv.V[j * 3 + 0] + sv.U[i * 3 + 1] * ns[1] * sv.V[j * 3 + 1] +
                           sv.U[i * 3 + 2] * ns[2] * sv.V[j * 3 + 2] - (i == j ? 1.0f : 0.0f);
}

// tau = U diag(k) U^T (symmetric)
__device__ __forceinline__ void PrincipalToTau(const Svd3& sv, const float (&k)[3], float (&tau)[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = sv.U[i * 3 + 0] * k[0] * sv.U[j * 3 + 0] + sv.U[i * 3 + 1] * k[1] * sv.U[j * 3 + 1] +
                            sv.U[i * 3 + 2] * k[2] * sv.U[j * 3 + 2];
            tau[i * 3 + j] = v;
            tau[j * 3 + i] = v;
        }
}

// von Mises plasticity over the StVK-Hencky law: return map mpm_algorithm.cuh:450-487,
// stress mpm_material.cuh:479-519.  Yield when |dev eps| exceeds yield / (2 mu); the
// return moves dev eps radially back onto that radius (trace kept) and rebuilds
// F = U exp(eps) V^T.  The reference then takes the SVD of the projected F again; its
// singular values are exp(eps) with the same U, so tau = U diag(2 mu eps + lambda tr eps)
// U^T directly.
__device__ __forceinline__ void StressVonMises(float lambda, float mu, float yield, float (&E)[9], float (&tau)[9]) {
    Svd3 sv;
    float d[3], eps[3];
    SignedSvdStrain(E, sv, d);
    HenckyStrain(sv, d, eps);
    const float tr = eps[0] + eps[1] + eps[2];
    float eh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) eh[a] = eps[a] - tr * (1.0f / 3.0f);
    const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    const float over = ehn - yield / (2.0f * mu);
    if (over > 0.0f && ehn > 0.0f) {
        const float r = over / ehn;
        float ns[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            eps[a] -= r * eh[a];
            ns[a] = ExpSmall(eps[a]);
        }
        RebuildStrain(sv, ns, E);
    }
    float k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k[a] = 2.0f * mu * eps[a] + lambda * tr;
    PrincipalToTau(sv, k, tau);
}

// Non-associated Cam-clay, mpm_material.cuh:521-671 (called from
// mpm_algorithm.cuh:329-400).  kappa = 2 mu / 3 + lambda; the yield surface in (p, q)
// is (1 + 2 beta) q^2 + M^2 (p - p_min)(p - p0) = 0 with p0 = kappa (1e-5 +
// sinh(xi max(-logJp, 0))), p_min = -beta p0.
//  * trial state from the SVD of F: J, s_hat = mu J^(-2/3) dev(s^2), p = -kappa/2 (J^2 - 1);
//  * p outside [p_min, p0]: F is projected to the isotropic stretch of the cap pressure,
//    J_new = sqrt(1 - 2 p / kappa), and logJp += log(J / J_new) when hardening;
//  * else, when y >= 1e-4, s_hat is scaled onto the yield surface at the trial p
//    (s_i^2 = s_hat_i c + tr(s^2)/3) and, when hardening and p lies inside the band, logJp
//    grows by log(J / J_fake), J_fake from the point where the ray from the ellipse
//    centre through (p, q) meets the surface.  The reference performs this update inside
//    a 9-fold loop over matrix entries (mpm_material.cuh:612-640), so it adds 9 times;
//    kept, for identical results;
//  * tau = mu J_trial^(-2/3) dev(F F^T) + kappa/2 (J^2 - 1) I (the deviatoric factor
//    uses the trial J, as the reference does).
// The reference rebuilds a projected F as U diag(s) V (mpm_material.cuh:574-581,
// 599-609); this uses U diag(s) V^T.  The two differ by a rotation on the right, which
// changes neither F F^T nor the singular values, hence neither the stress nor any later
// state the stress depends on.
// Small strains are formed from d = s - 1 (J - 1 and s^2 - mean s^2 without cancellation).
__device__ __forceinline__ void StressCamClay(const MaterialDev& m, float lambda, float mu, float (&E)[9], float& logJp,
                                              float (&tau)[9]) {
    const float kappa = 2.0f * mu / 3.0f + lambda;
    const float beta = m.beta, msq = m.msq;
    Svd3 sv;
    float d[3];
    SignedSvdStrain(E, sv, d);
    const float p0 = kappa * (1e-5f + sinhf(m.xi * fmaxf(-logJp, 0.0f)));
    const float pmin = -beta * p0;
    // J - 1 and s_i^2 - 1 from d
    const float jm1 = d[0] + d[1] + d[2] + d[0] * d[1] + d[0] * d[2] + d[1] * d[2] + d[0] * d[1] * d[2];
    const float Je = 1.0f + jm1;
    float bd[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) bd[a] = d[a] * (2.0f + d[a]);
    const float bmean = (bd[0] + bd[1] + bd[2]) * (1.0f / 3.0f);
    const float devf = mu * expf(-2.0f / 3.0f * log1pf(jm1));  // mu J^(-2/3)
    float sh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) sh[a] = devf * (bd[a] - bmean);
    const float ptr = -0.5f * kappa * jm1 * (2.0f + jm1);  // -kappa/2 (J^2 - 1)
    const float ysh = 1.5f * (1.0f + 2.0f * beta);
    const float yph = msq * (ptr - pmin) * (ptr - p0);
    const float sn2 = sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2];
    const float y = ysh * sn2 + yph;
    // stretches after the return (b = s^2 - 1 kept precise when nothing is projected)
    float b1[3] = {bd[0], bd[1], bd[2]};
    float j2m1 = jm1 * (2.0f + jm1);  // J^2 - 1
    bool rebuild = false;
    float ns[3];
    if (ptr > p0 || ptr < pmin) {
        const float p = ptr > p0 ? p0 : pmin;
        const float jn2 = 1.0f - 2.0f * p / kappa;
        const float jn = sqrtf(jn2);
        const float sn = cbrtf(jn);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ns[a] = sn;
            b1[a] = sn * sn - 1.0f;
        }
        j2m1 = jn2 - 1.0f;
        rebuild = true;
        if (m.hardening) logJp += logf(Je / jn);
    } else if (y >= 1e-4f) {
        const float c = expf(2.0f / 3.0f * log1pf(jm1)) / mu * sqrtf(-yph / ysh) / sqrtf(sn2);
        const float tr3 = 1.0f + bmean;  // tr(s^2) / 3
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ns[a] = sqrtf(sh[a] * c + tr3);
            b1[a] = sh[a] * c + bmean;
        }
        const float jn = ns[0] * ns[1] * ns[2];
        j2m1 = jn * jn - 1.0f;
        rebuild = true;
        if (m.hardening && p0 > 1e-4f && ptr < p0 - 1e-4f && ptr > 1e-4f + pmin) {
            const float pc = (1.0f - beta) * p0 * 0.5f;
            const float qt = sqrtf(1.5f * sn2);
            float d0 = pc - ptr, d1 = -qt;
            const float dn = sqrtf(d0 * d0 + d1 * d1);
            d0 /= dn;
            d1 /= dn;
            const float A = msq * d0 * d0 + (1.0f + 2.0f * beta) * d1 * d1;
            const float B = msq * d0 * (2.0f * pc - p0 - pmin);
            const float C = msq * (pc - pmin) * (pc - p0);
            const float disc = sqrtf(B * B - 4.0f * A * C);
            const float p1 = pc + (-B + disc) / (2.0f * A) * d0;
            const float p2 = pc + (-B - disc) / (2.0f * A) * d0;
            const float pf = (ptr - pc) * (p1 - pc) > 0.0f ? p1 : p2;
            const float jf = sqrtf(fabsf(1.0f - 2.0f * pf / kappa));
            if (jf > 1e-4f) logJp += 9.0f * logf(Je / jf);
        }
    }
    if (rebuild) RebuildStrain(sv, ns, E);
    const float b1m = (b1[0] + b1[1] + b1[2]) * (1.0f / 3.0f);
    const float hyd = 0.5f * kappa * j2m1;
    float k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k[a] = devf * (b1[a] - b1m) + hyd;
    PrincipalToTau(sv, k, tau);
}

// Weakly compressible fluid, mpm_algorithm.cuh:539-547 and mpm_material.cuh:466-477:
// J <- J (1 + dt tr C); p = K (J^-gamma - 1); tau = J (nu (C + C^T) - p I).
// The state holds q = J - 1 (the reference stores J): near J = 1 an FP32 J keeps only
// ~6e-8 absolute resolution, which is a large relative error of J - 1 and hence of the
// pressure; q keeps full relative precision.  For an integer gamma g (the scenes use 7),
// with r = 1/J: J^-g - 1 = (r - 1)(1 + r + ... + r^(g-1)) and r - 1 = -q r, a sum of
// positive terms that keeps full relative precision near J = 1 (a rcp and g - 1 fmas);
// otherwise p = K expm1(-gamma log1p(q)), which costs two library calls.
__device__ __forceinline__ void StressFluid(const MaterialDev& m, const float (&C)[9], float dt, float& q, float (&tau)[9]) {
    q += (C[0] + C[4] + C[8]) * dt * (1.0f + q);
    const float J = 1.0f + q;
    float p;
    if (m.gamma_int > 0) {
        const float r = __builtin_amdgcn_rcpf(J);
        float sum = 1.0f;
        for (int k = 1; k < m.gamma_int; ++k) sum = fmaf(sum, r, 1.0f);
        p = m.bulk * (-q * r) * sum;
    } else {
        p = m.bulk * expm1f(-m.gamma * log1pf(q));
    }
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
