// FP32 3x3 signed SVD for the constitutive updates.
//
// Method: the structure of McAdams, Selle, Tamstorf, Teran, Sifakis, "Computing the
// Singular Value Decomposition of 3x3 matrices with minimal branching and elementary
// floating point operations", UW-Madison CS TR1690 (2011):
//   1. up to four sweeps of cyclic Jacobi on S = A^T A, accumulating V.  The rotations
//      are exact Jacobi rotations (Golub & Van Loan, Matrix Computations, sec. 8.5)
//      rather than the paper's approximate half-angle ones: in FP32 those leave relative
//      off-diagonal residues up to ~1e-3 after four sweeps when two singular values
//      are close, while exact rotations converge fully in four.  Jacobi converges
//      quadratically, so after two sweeps a wavefront whose every lane is diagonal to
//      FP32 resolution skips the rest (deformation gradients near a rotation, the
//      common case, need two);
//   2. B = A V, columns sorted by decreasing norm (swaps negate one column so that
//      det V stays +1);
//   3. Givens QR of B (rows (0,1), (0,2), (1,2)); U is the product of the transposed
//      rotations and the diagonal of R gives the singular values.
// Convention (the reference's, mpm_math.cuh:37-1072): U, V rotations, s0 >= s1 >= |s2|,
// the sign of det A carried by s2.  This is an independent implementation written from
// the published method; it does not reproduce any vendor math routine.
#pragma once

#include <hip/hip_runtime.h>

namespace ckmpm {

struct Svd3 {
    float U[9];
    float s[3];
    float V[9];
};

namespace svd_detail {

constexpr float kQrEps = 1e-6f;

// One Jacobi conjugation S <- Q^T S Q in the (p, q) plane (r the third index),
// V <- V Q, with Q = [[c, -s], [s, c]] the exact rotation that zeroes S[p][q]:
// t = tan(theta) is the smaller root of b t^2 - (d - a) t - b = 0,
// t = -2 b sgn(d - a) / (|d - a| + sqrt((d - a)^2 + 4 b^2)).
template <int p, int q, int r>
__device__ __forceinline__ void JacobiConjugate(float (&S)[9], float (&V)[9]) {
    const float a = S[p * 3 + p], d = S[q * 3 + q], b = S[p * 3 + q];
    const float dm = d - a;
    // single-instruction sqrt / reciprocal / rsqrt (1 ulp): the rotation only has to
    // be orthogonal, which c = rsq(1 + t^2), s = t c gives whatever t is
    const float den = fabsf(dm) + __builtin_amdgcn_sqrtf(dm * dm + 4.0f * b * b);
    const float t = den > 0.0f ? (dm >= 0.0f ? -2.0f : 2.0f) * b * __builtin_amdgcn_rcpf(den) : 0.0f;
    const float c = __builtin_amdgcn_rsqf(1.0f + t * t);
    const float s = t * c;
    const float cc = c * c, ss = s * s, cs = c * s;
    const float spp = a * cc + 2.0f * b * cs + d * ss;
    const float sqq = a * ss - 2.0f * b * cs + d * cc;
    const float spr = c * S[p * 3 + r] + s * S[q * 3 + r];
    const float sqr = -s * S[p * 3 + r] + c * S[q * 3 + r];
    S[p * 3 + p] = spp;
    S[q * 3 + q] = sqq;
    S[p * 3 + q] = S[q * 3 + p] = 0.0f;
    S[p * 3 + r] = S[r * 3 + p] = spr;
    S[q * 3 + r] = S[r * 3 + q] = sqr;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float vp = V[k * 3 + p], vq = V[k * 3 + q];
        V[k * 3 + p] = c * vp + s * vq;
        V[k * 3 + q] = -s * vp + c * vq;
    }
}

// Every lane of the wavefront has |S_pq| <= 2^-24 min(|S_pp|, |S_qq|) in all three
// planes: a further rotation would move the eigenvalues by at most an ulp (and rotate
// eigenvectors only within a plane of equal eigenvalues, to which the stresses are
// insensitive).  Wave-uniform, so the whole wavefront skips the remaining sweeps.
__device__ __forceinline__ bool JacobiConverged(const float (&S)[9]) {
    const float tol = 5.9604645e-8f;  // 2^-24
    const float a = fabsf(S[0]), b = fabsf(S[4]), c = fabsf(S[8]);
    const bool ok = fabsf(S[1]) <= tol * fminf(a, b) && fabsf(S[2]) <= tol * fminf(a, c) &&
                    fabsf(S[5]) <= tol * fminf(b, c);
    return __all(ok) != 0;
}

__device__ __forceinline__ void JacobiSweep(float (&S)[9], float (&V)[9]) {
    JacobiConjugate<0, 1, 2>(S, V);
    JacobiConjugate<0, 2, 1>(S, V);
    JacobiConjugate<1, 2, 0>(S, V);
}

// Two sweeps, then up to two more while some lane of the wavefront is not diagonal.
__device__ __forceinline__ void JacobiDiagonalise(float (&S)[9], float (&V)[9]) {
    JacobiSweep(S, V);
    JacobiSweep(S, V);
    if (!JacobiConverged(S)) {
        JacobiSweep(S, V);
        if (!JacobiConverged(S)) JacobiSweep(S, V);
    }
}

// Swap columns i, j of B and V, negating the new column j of both (keeps B = A V and
// det V = +1), when column i is shorter than column j.
template <int i, int j>
__device__ __forceinline__ void SortPair(float (&B)[9], float (&V)[9], float (&rho)[3]) {
    const bool sw = rho[i] < rho[j];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float bi = B[k * 3 + i], bj = B[k * 3 + j];
        const float vi = V[k * 3 + i], vj = V[k * 3 + j];
        B[k * 3 + i] = sw ? bj : bi;
        B[k * 3 + j] = sw ? -bi : bj;
        V[k * 3 + i] = sw ? vj : vi;
        V[k * 3 + j] = sw ? -vi : vj;
    }
    const float ri = rho[i], rj = rho[j];
    rho[i] = sw ? rj : ri;
    rho[j] = sw ? ri : rj;
}

// Givens rotation on rows (p, q) of B zeroing B[q][col]; U <- U G^T.
template <int p, int q, int col>
__device__ __forceinline__ void QrGivens(float (&B)[9], float (&U)[9]) {
    const float a1 = B[p * 3 + col], a2 = B[q * 3 + col];
    const float rho = __builtin_amdgcn_sqrtf(a1 * a1 + a2 * a2);
    float sh = rho > kQrEps ? a2 : 0.0f;
    float ch = fabsf(a1) + fmaxf(rho, kQrEps);
    if (a1 < 0.0f) {
        const float tmp = sh;
        sh = ch;
        ch = tmp;
    }
    const float w = __builtin_amdgcn_rsqf(ch * ch + sh * sh);
    ch *= w;
    sh *= w;
    const float c = ch * ch - sh * sh, s = 2.0f * ch * sh;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float bp = B[p * 3 + k], bq = B[q * 3 + k];
        B[p * 3 + k] = c * bp + s * bq;
        B[q * 3 + k] = -s * bp + c * bq;
        const float up = U[k * 3 + p], uq = U[k * 3 + q];
        U[k * 3 + p] = c * up + s * uq;
        U[k * 3 + q] = -s * up + c * uq;
    }
}

}  // namespace svd_detail

__device__ __forceinline__ void SignedSvd3(const float (&A)[9], Svd3& out) {
    using namespace svd_detail;
    float S[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = A[0 * 3 + i] * A[0 * 3 + j] + A[1 * 3 + i] * A[1 * 3 + j] + A[2 * 3 + i] * A[2 * 3 + j];
            S[i * 3 + j] = v;
            S[j * 3 + i] = v;
        }
    float V[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    JacobiDiagonalise(S, V);
    float B[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) B[i * 3 + j] = A[i * 3 + 0] * V[0 * 3 + j] + A[i * 3 + 1] * V[1 * 3 + j] + A[i * 3 + 2] * V[2 * 3 + j];
    float rho[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) rho[j] = B[0 * 3 + j] * B[0 * 3 + j] + B[1 * 3 + j] * B[1 * 3 + j] + B[2 * 3 + j] * B[2 * 3 + j];
    SortPair<0, 1>(B, V, rho);
    SortPair<0, 2>(B, V, rho);
    SortPair<1, 2>(B, V, rho);
    float U[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    QrGivens<0, 1, 0>(B, U);
    QrGivens<0, 2, 0>(B, U);
    QrGivens<1, 2, 1>(B, U);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        out.U[k] = U[k];
        out.V[k] = V[k];
    }
    out.s[0] = B[0];
    out.s[1] = B[4];
    out.s[2] = B[8];
}

// Signed SVD of F = I + E that also returns d_i = s_i - 1 to full relative precision.
// The Jacobi stage diagonalises M = F^T F - I = E + E^T + E^T E, whose entries are of
// strain size, so its eigenvalues m_i (s_i^2 = 1 + m_i) keep the strain's relative
// precision; d_i = m_i / (1 + sqrt(1 + m_i)) for s_i >= 1/2, and s_i = R_ii (the Givens QR
// of F V, which also gives U as above) for a smaller or negative s_i.
__device__ __forceinline__ void SignedSvdStrain(const float (&E)[9], Svd3& out, float (&d)[3]) {
    using namespace svd_detail;
    float S[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const float v = E[i * 3 + j] + E[j * 3 + i] + E[0 * 3 + i] * E[0 * 3 + j] + E[1 * 3 + i] * E[1 * 3 + j] +
                            E[2 * 3 + i] * E[2 * 3 + j];
            S[i * 3 + j] = v;
            S[j * 3 + i] = v;
        }
    float V[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    JacobiDiagonalise(S, V);
    float B[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            B[i * 3 + j] = V[i * 3 + j] + E[i * 3 + 0] * V[0 * 3 + j] + E[i * 3 + 1] * V[1 * 3 + j] + E[i * 3 + 2] * V[2 * 3 + j];
    float m[3] = {S[0], S[4], S[8]};
    SortPair<0, 1>(B, V, m);
    SortPair<0, 2>(B, V, m);
    SortPair<1, 2>(B, V, m);
    float U[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    QrGivens<0, 1, 0>(B, U);
    QrGivens<0, 2, 0>(B, U);
    QrGivens<1, 2, 1>(B, U);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        out.U[k] = U[k];
        out.V[k] = V[k];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // near 1, s_i = 1 + m_i / (1 + sqrt(1 + m_i)) keeps the relative precision of a
        // small strain; below 1/2 (a compressed or inverted F) the diagonal of R is the
        // accurate value (1 + m_i carries an absolute error ~1e-7, which is the whole of
        // a singular value near 1e-5)
        const float root = sqrtf(fmaxf(1.0f + m[i], 0.0f));
        const float r = B[i * 4];
        const bool direct = r < 0.5f;
        d[i] = direct ? r - 1.0f : m[i] / (1.0f + root);
        out.s[i] = direct ? r : 1.0f + d[i];
    }
}

}  // namespace ckmpm
