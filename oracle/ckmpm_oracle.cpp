// FP64 host restatement of the CK-MPM substep (binning, block activation,
// fused G2P/constitutive/P2G, grid update with sticky boundary, rebin).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is compared
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it.  The product library (libckmpm.so) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference tree).  The reference leaves three things unspecified or
// nondeterministic (atomic block insertion order, atomic particle order within a
// block, float atomic summation order); this restatement fixes them with the
// canonical rules that the HIP engine implements as well:
//   * block keys are computed from FP32 positions with FP32 arithmetic: the engine keeps
//     each position in FP32 relative to the origin 4 b dx of the block b that processed
//     it, so a particle's next block is b + floor((floor(u/dx + 1/4) - 2) / 4) on its
//     FP32 local coordinate u (identical to the reference's rule on global coordinates
//     whenever the global FP32 value is exact, and about 100x finer otherwise);
//   * a block list is "previous order with empty blocks removed, then newly
//     occupied blocks in list order", and the neighbour / exterior extensions are
//     appended in ascending linear key;
//   * particle order inside a block is by the move that brought the particle in, then
//     by the previous slot (a stable counting sort by (new block index, move)); an
//     ordering word is slot | move << 27, move = (dx+1)*9 + (dy+1)*3 + (dz+1) of the
//     block change.  Each move into a block comes from one source block, so the
//     particles that stay (move 13) keep their previous order.
// Physics (weights, G2P, F update, stress, P2G, grid update) runs in FP64.
//
// Parity pinning: see the header of oracle/pyoracle.py and tests/test_oracle_*.py.
// Every function below is checked there against closed forms, invariants,
// brute force or numpy (SVD); none is left unpinned.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr double kPi = 3.14159265358979323846;
constexpr int kMoveShift = 27;
constexpr uint32_t kSlotMask = (1u << kMoveShift) - 1u;
constexpr uint32_t kNoMove = 13u << kMoveShift;

enum Kind { kFixedCorotated = 0, kNeoHookean = 1, kDruckerPrager = 2, kFluid = 3, kVonMises = 4, kCamClay = 5 };

enum Flags : uint32_t {
    kFlagOutOfWindow = 1u << 0,   // particle left its 8^3-node P2G window
    kFlagOutOfDomain = 1u << 1,   // particle block outside the domain
    kFlagNonFinite = 1u << 2,     // NaN grid velocity (mapped to +inf)
    kFlagCapacity = 1u << 3,      // block list exceeds capacity
    kFlagMissingBlock = 1u << 4,  // a stencil node's block not in the neighbour set
};

struct Domain {
    int dim[3];
    int off[3];
    uint64_t nblocks;
};

inline uint32_t KeyOf(const Domain& d, int bx, int by, int bz) {
    // mpm_partition.cuh:154-166 (Index: GetFlattenedIndex<Dy,Dz>(x,y,z))
    return static_cast<uint32_t>((static_cast<int64_t>(bx) * d.dim[1] + by) * d.dim[2] + bz);
}
inline void CoordsOf(const Domain& d, uint32_t key, int c[3]) {
    c[2] = static_cast<int>(key % static_cast<uint32_t>(d.dim[2]));
    const uint32_t r = key / static_cast<uint32_t>(d.dim[2]);
    c[1] = static_cast<int>(r % static_cast<uint32_t>(d.dim[1]));
    c[0] = static_cast<int>(r / static_cast<uint32_t>(d.dim[1]));
}
inline bool InDomain(const Domain& d, int bx, int by, int bz) {
    return bx >= 0 && by >= 0 && bz >= 0 && bx < d.dim[0] && by < d.dim[1] && bz < d.dim[2];
}
inline int FloorDiv4(int v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }

// ---------------------------------------------------------------- FP32 binning
// mpm_algorithm.cuh:59-67 (GetCellIndex<-1>): cell = int(x*invDx + 0.25f) in FP32.
// Built with -ffp-contract=off so the multiply and the add round separately, exactly
// as the two FP32 operations of the reference.
inline int Cell0F(float x, float inv_dx) {
    volatile float t = x * inv_dx;
    float u = t + 0.25f;
    return static_cast<int>(std::floor(u));
}

// Particle block of a position: (cell-2)/4 per axis, mpm_algorithm.cuh:595-600 and
// mpm_particle.cuh:337-340 (floor division; identical to the reference's C division
// for every cell >= 2, which the sticky boundary band guarantees).
inline void BlockOfF(const float x[3], float inv_dx, int b[3], int c[3]) {
    for (int a = 0; a < 3; ++a) {
        c[a] = Cell0F(x[a], inv_dx);
        b[a] = FloorDiv4(c[a] - 2);
    }
}

// ------------------------------------------------------------------- 3x3 math
using M3 = double[9];

inline void MatMul(const double* A, const double* B, double* C) {
    double T[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int j = 0; j < 3; ++j) s += A[i * 3 + j] * B[j * 3 + k];
            T[i * 3 + k] = s;
        }
    std::memcpy(C, T, sizeof(T));
}
inline void Transpose(const double* A, double* T) {
    double R[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = A[j * 3 + i];
    std::memcpy(T, R, sizeof(R));
}
// data_type.cuh:402-414
inline double Det(const double* a) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// data_type.cuh:369-387 (adjugate / determinant)
inline void Inverse(const double* a, double* r) {
    const double id = 1.0 / Det(a);
    double t[9];
    t[0] = id * (a[4] * a[8] - a[5] * a[7]);
    t[1] = id * (a[2] * a[7] - a[1] * a[8]);
    t[2] = id * (a[1] * a[5] - a[2] * a[4]);
    t[3] = id * (a[5] * a[6] - a[3] * a[8]);
    t[4] = id * (a[0] * a[8] - a[2] * a[6]);
    t[5] = id * (a[2] * a[3] - a[0] * a[5]);
    t[6] = id * (a[3] * a[7] - a[4] * a[6]);
    t[7] = id * (a[1] * a[6] - a[0] * a[7]);
    t[8] = id * (a[0] * a[4] - a[1] * a[3]);
    std::memcpy(r, t, sizeof(t));
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi, run to convergence.
// Columns of Q are eigenvectors, lam the eigenvalues.
void SymEig(const double* S, double* Q, double* lam) {
    double A[9];
    std::memcpy(A, S, sizeof(A));
    for (int i = 0; i < 9; ++i) Q[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        const double diag = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = A[p * 3 + q];
                if (apq == 0.0) continue;
                const double app = A[p * 3 + p], aqq = A[q * 3 + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                // A <- J^T A J with J the rotation in the (p,q) plane
                for (int k = 0; k < 3; ++k) {
                    const double akp = A[k * 3 + p], akq = A[k * 3 + q];
                    A[k * 3 + p] = c * akp - s * akq;
                    A[k * 3 + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = A[p * 3 + k], aqk = A[q * 3 + k];
                    A[p * 3 + k] = c * apk - s * aqk;
                    A[q * 3 + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double qkp = Q[k * 3 + p], qkq = Q[k * 3 + q];
                    Q[k * 3 + p] = c * qkp - s * qkq;
                    Q[k * 3 + q] = s * qkp + c * qkq;
                }
            }
    }
    lam[0] = A[0];
    lam[1] = A[4];
    lam[2] = A[8];
}

// Signed SVD F = U diag(s) V^T with the sign and order convention of
// mpm_math.cuh:37-1072 (McAdams et al. 2011): U, V rotations (det +1),
// s0 >= s1 >= |s2|, the sign of det F carried by s2.  The FP64 algorithm itself
// (Jacobi on F^T F to convergence, then orthonormalisation of F V) is this
// oracle's own; only the convention follows the reference.
void SignedSvd(const double* F, double* U, double* s, double* V) {
    double FtF[9], Ft[9];
    Transpose(F, Ft);
    MatMul(Ft, F, FtF);
    double lam[3];
    SymEig(FtF, V, lam);
    // sort eigenpairs by eigenvalue descending
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int a, int b) { return lam[a] > lam[b]; });
    double Vs[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Vs[r * 3 + c] = V[r * 3 + idx[c]];
    std::memcpy(V, Vs, sizeof(Vs));
    if (Det(V) < 0)
        for (int r = 0; r < 3; ++r) V[r * 3 + 2] = -V[r * 3 + 2];
    double B[9];
    MatMul(F, V, B);
    // U columns: normalise b0, orthonormalise b1, u2 = u0 x u1
    double u0[3] = {B[0], B[3], B[6]}, u1[3] = {B[1], B[4], B[7]};
    double n0 = std::sqrt(u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
    if (n0 < 1e-300) {
        u0[0] = 1;
        u0[1] = 0;
        u0[2] = 0;
        n0 = 0;
    } else
        for (double& v : u0) v /= n0;
    double d = u0[0] * u1[0] + u0[1] * u1[1] + u0[2] * u1[2];
    for (int a = 0; a < 3; ++a) u1[a] -= d * u0[a];
    double n1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    if (n1 < 1e-150 * (n0 + 1e-300) || n1 < 1e-300) {
        // pick any unit vector orthogonal to u0
        double e[3] = {0, 0, 0};
        int m = 0;
        for (int a = 1; a < 3; ++a)
            if (std::fabs(u0[a]) < std::fabs(u0[m])) m = a;
        e[m] = 1;
        d = u0[m];
        for (int a = 0; a < 3; ++a) u1[a] = e[a] - d * u0[a];
        const double nn = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
        for (double& v : u1) v /= nn;
    } else {
        for (double& v : u1) v /= n1;
    }
    // true singular value of column 1 is |B1| (columns of F V are orthogonal)
    const double s1 = B[1] * u1[0] + B[4] * u1[1] + B[7] * u1[2];
    double u2[3] = {u0[1] * u1[2] - u0[2] * u1[1], u0[2] * u1[0] - u0[0] * u1[2], u0[0] * u1[1] - u0[1] * u1[0]};
    const double s2 = B[2] * u2[0] + B[5] * u2[1] + B[8] * u2[2];
    for (int r = 0; r < 3; ++r) {
        U[r * 3 + 0] = u0[r];
        U[r * 3 + 1] = u1[r];
        U[r * 3 + 2] = u2[r];
    }
    s[0] = n0;
    s[1] = s1;
    s[2] = s2;
}

struct Material {
    int kind = kFixedCorotated;
    double lambda = 0, mu = 0, bulk = 0, gamma = 0, viscosity = 0, J0 = 1, sin_phi = 0.5;
    double yield = 0;                                  // von Mises
    double alpha0 = 0, beta = 0, xi = 0, M = 0;        // Cam-clay
    bool hardening = true;
};
inline void MaterialFromParams(int kind, const double* params, Material& m) {
    m.kind = kind;
    m.lambda = params[0];
    m.mu = params[1];
    m.bulk = params[2];
    m.gamma = params[3];
    m.viscosity = params[4];
    m.J0 = params[5];
    m.sin_phi = std::sin(params[6] * kPi / 180.0);
    m.yield = params[7];
    m.alpha0 = params[8];
    m.beta = params[9];
    m.xi = params[10];
    m.M = params[11];
    m.hardening = params[12] != 0.0;
}

// FC: mpm_material.cuh:318-357.  tau = U diag(2 mu s(s-1) + lambda J(J-1)) U^T.
void StressFixedCorotated(double lambda, double mu, const double* F, double* tau) {
    double U[9], s[3], V[9];
    SignedSvd(F, U, s, V);
    const double J = s[0] * s[1] * s[2];
    double sc[3];
    for (int a = 0; a < 3; ++a) sc[a] = 2.0 * mu * s[a] * (s[a] - 1.0) + lambda * J * (J - 1.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += U[i * 3 + k] * sc[k] * U[j * 3 + k];
            tau[i * 3 + j] = v;
        }
}

// Neo-Hookean (absent from the reference, north-star addition; closed form):
// tau = mu (F F^T - I) + lambda ln(J) I.
void StressNeoHookean(double lambda, double mu, const double* F, double* tau) {
    double Ft[9], B[9];
    Transpose(F, Ft);
    MatMul(F, Ft, B);
    const double lnJ = std::log(Det(F));
    for (int i = 0; i < 9; ++i) tau[i] = mu * (B[i] - ((i % 4 == 0) ? 1.0 : 0.0));
    for (int i = 0; i < 3; ++i) tau[i * 4] += lambda * lnJ;
}

// Drucker-Prager StVK-Hencky return map: mpm_material.cuh:360-464 (cohesion 0,
// beta 1, volume correction on; friction angle a parameter, 30 degrees in the
// reference's hard-coded constant at :369).  Projects F in place, updates logJp.
void StressDruckerPrager(double lambda, double mu, double sin_phi, double* F, double& logJp, double* tau) {
    const double alpha = std::sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi);
    const double smu = 2.0 * mu;
    double U[9], s[3], V[9];
    SignedSvd(F, U, s, V);
    double eps[3];
    for (int a = 0; a < 3; ++a) eps[a] = std::log(std::max(std::fabs(s[a]), 1e-4));
    const double eps_sum = eps[0] + eps[1] + eps[2];
    const double eps_tr = eps_sum + logJp;
    double eh[3];
    for (int a = 0; a < 3; ++a) eh[a] = eps[a] - eps_tr / 3.0;
    const double ehn = std::sqrt(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    double ns[3];
    bool rebuild = true;
    if (eps_tr >= 0.0) {
        ns[0] = ns[1] = ns[2] = 1.0;
        logJp = eps_sum + logJp;
    } else {
        double H[3] = {eps[0], eps[1], eps[2]};
        if (mu != 0.0) {
            logJp = 0.0;
            const double dg = ehn + (3.0 * lambda + smu) / smu * eps_tr * alpha;
            if (dg > 0.0)
                for (int a = 0; a < 3; ++a) H[a] = eps[a] - (dg / ehn) * eh[a];
        } else {
            // :412: with mu == 0 neither branch runs and F is left as it is (the
            // reference's newS is then uninitialised; exp(eps) stands in for it here)
            rebuild = false;
        }
        for (int a = 0; a < 3; ++a) ns[a] = std::exp(H[a]);
    }
    // F = U diag(ns) V^T: in the elastic branch (:417-443) too, where it clamps |s| at
    // 1e-4 and un-inverts an inverted F
    if (rebuild)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int k = 0; k < 3; ++k) v += U[i * 3 + k] * ns[k] * V[j * 3 + k];
                F[i * 3 + j] = v;
            }
    // tau = U diag(phat) V^T F^T with phat = (2 mu log ns + lambda tr log ns) / ns
    double lg[3] = {std::log(ns[0]), std::log(ns[1]), std::log(ns[2])};
    const double ltr = lg[0] + lg[1] + lg[2];
    double ph[3];
    for (int a = 0; a < 3; ++a) ph[a] = (smu * lg[a] + lambda * ltr) / ns[a];
    double P[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += U[i * 3 + k] * ph[k] * V[j * 3 + k];
            P[i * 3 + j] = v;
        }
    double Ft[9];
    Transpose(F, Ft);
    MatMul(P, Ft, tau);
}

// tau = U diag(k) U^T
inline void PrincipalTau(const double* U, const double* k, double* tau) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0;
            for (int c = 0; c < 3; ++c) v += U[i * 3 + c] * k[c] * U[j * 3 + c];
            tau[i * 3 + j] = v;
        }
}
// F = U diag(s) V^T
inline void Compose(const double* U, const double* s, const double* V, double* F) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0;
            for (int c = 0; c < 3; ++c) v += U[i * 3 + c] * s[c] * V[j * 3 + c];
            F[i * 3 + j] = v;
        }
}

// von Mises over StVK-Hencky: return map mpm_algorithm.cuh:450-487 (eps = log s,
// yield when |dev eps| > yield / (2 mu), radial return of dev eps, F = U exp(eps) V^T),
// stress mpm_material.cuh:479-519 on the projected F: tau = U diag(phat) V^T F^T with
// phat = (2 mu log s + lambda tr log s) / s, from a second SVD as the reference takes it.
// log |s| clamped at 1e-4 as in the Drucker-Prager model (the reference's logf of a
// non-positive singular value is NaN).
void StressVonMises(double lambda, double mu, double yield, double* F, double* tau) {
    double U[9], s[3], V[9];
    SignedSvd(F, U, s, V);
    double eps[3];
    for (int a = 0; a < 3; ++a) eps[a] = std::log(std::max(std::fabs(s[a]), 1e-4));
    const double tr = eps[0] + eps[1] + eps[2];
    double eh[3];
    for (int a = 0; a < 3; ++a) eh[a] = eps[a] - tr / 3.0;
    const double ehn = std::sqrt(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    const double over = ehn - yield / (2.0 * mu);
    if (over > 0.0) {
        double ns[3];
        for (int a = 0; a < 3; ++a) ns[a] = std::exp(eps[a] - over * eh[a] / ehn);
        Compose(U, ns, V, F);
    }
    SignedSvd(F, U, s, V);
    double lg[3];
    for (int a = 0; a < 3; ++a) lg[a] = std::log(std::max(std::fabs(s[a]), 1e-4));
    const double ltr = lg[0] + lg[1] + lg[2];
    double ph[3];
    for (int a = 0; a < 3; ++a) ph[a] = (2.0 * mu * lg[a] + lambda * ltr) / s[a];
    double P[9], Ft[9];
    Compose(U, ph, V, P);
    Transpose(F, Ft);
    MatMul(P, Ft, tau);
}

// Non-associated Cam-clay: mpm_material.cuh:521-671 (UpdateForce mpm_algorithm.cuh:329-400).
// kappa = 2 mu / 3 + lambda, p0 = kappa (1e-5 + sinh(xi max(-logJp, 0))), p_min = -beta p0.
// Trial: J = s0 s1 s2, s_hat = mu J^(-2/3) (s^2 - tr(s^2)/3), p = -kappa/2 (J - 1/J) J,
// y = 3/2 (1 + 2 beta) |s_hat|^2 + M^2 (p - p_min)(p - p0).
//  * p > p0 or p < p_min: s = J_new^(1/3) with J_new = sqrt(1 - 2 p_cap / kappa);
//    logJp += log(J / J_new) when hardening (:555-583);
//  * else if y >= 1e-4: s_i = sqrt(s_hat_i c + tr(s^2)/3), c = J^(2/3) / mu
//    sqrt(-M^2 (p - p_min)(p - p0) / (3/2 (1 + 2 beta))) / |s_hat| (:585-610); with
//    hardening and p inside (p_min + 1e-4, p0 - 1e-4), logJp += log(J / J_fake) nine times
//    (the update sits in the reference's 9-iteration loop, :612-640), J_fake from the
//    point where the ray from (p_c, 0), p_c = (1 - beta) p0 / 2, through (p, -q),
//    q = sqrt(3/2) |s_hat|, meets the yield ellipse;
//  * tau = mu J_trial^(-2/3) dev(F F^T) + kappa/2 (J^2 - 1) I (:645-669).
// A projected F is U diag(s) V^T (the reference writes U diag(s) V; the two differ by a
// right rotation, which leaves F F^T, the singular values and so every stress unchanged).
void StressCamClay(const Material& m, double lambda, double mu, double* F, double& logJp, double* tau) {
    const double kappa = 2.0 * mu / 3.0 + lambda;
    const double beta = m.beta, msq = m.M * m.M;
    double U[9], s[3], V[9];
    SignedSvd(F, U, s, V);
    const double p0 = kappa * (1e-5 + std::sinh(m.xi * std::max(-logJp, 0.0)));
    const double pmin = -beta * p0;
    const double Je = s[0] * s[1] * s[2];
    const double bh[3] = {s[0] * s[0], s[1] * s[1], s[2] * s[2]};
    const double tr3 = (bh[0] + bh[1] + bh[2]) / 3.0;
    const double devf = mu * std::pow(Je, -2.0 / 3.0);
    double sh[3];
    for (int a = 0; a < 3; ++a) sh[a] = devf * (bh[a] - tr3);
    const double ptr = -(kappa * 0.5 * (Je - 1.0 / Je)) * Je;
    const double ysh = 1.5 * (1.0 + 2.0 * beta);
    const double yph = msq * (ptr - pmin) * (ptr - p0);
    const double sn2 = sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2];
    const double y = ysh * sn2 + yph;
    double ns[3] = {s[0], s[1], s[2]};
    if (ptr > p0 || ptr < pmin) {
        const double p = ptr > p0 ? p0 : pmin;
        const double jn = std::sqrt(-2.0 * p / kappa + 1.0);
        ns[0] = ns[1] = ns[2] = std::cbrt(jn);
        Compose(U, ns, V, F);
        if (m.hardening) logJp += std::log(Je / jn);
    } else if (y >= 1e-4) {
        const double c = std::pow(Je, 2.0 / 3.0) / mu * std::sqrt(-yph / ysh) / std::sqrt(sn2);
        for (int a = 0; a < 3; ++a) ns[a] = std::sqrt(sh[a] * c + tr3);
        Compose(U, ns, V, F);
        for (int rep = 0; rep < 9; ++rep)
            if (m.hardening && p0 > 1e-4 && ptr < p0 - 1e-4 && ptr > 1e-4 + pmin) {
                const double pc = (1.0 - beta) * p0 / 2.0;
                const double qt = std::sqrt(1.5 * sn2);
                double d0 = pc - ptr, d1 = -qt;
                const double dn = std::sqrt(d0 * d0 + d1 * d1);
                d0 /= dn;
                d1 /= dn;
                const double A = msq * d0 * d0 + (1.0 + 2.0 * beta) * d1 * d1;
                const double B = msq * d0 * (2.0 * pc - p0 - pmin);
                const double C = msq * (pc - pmin) * (pc - p0);
                const double disc = B * B - 4.0 * A * C;
                const double p1 = pc + (-B + std::sqrt(disc)) / (2.0 * A) * d0;
                const double p2 = pc + (-B - std::sqrt(disc)) / (2.0 * A) * d0;
                const double pf = (ptr - pc) * (p1 - pc) > 0 ? p1 : p2;
                const double jsq = -2.0 * pf / kappa + 1.0;
                const double jf = std::sqrt(std::fabs(jsq));
                if (jf > 1e-4) logJp += std::log(Je / jf);
            }
    }
    const double J = ns[0] * ns[1] * ns[2];
    double Ft[9], b[9];
    Transpose(F, Ft);
    MatMul(F, Ft, b);
    const double btr = (b[0] + b[4] + b[8]) / 3.0;
    const double hyd = kappa * 0.5 * (J * J - 1.0);
    for (int i = 0; i < 9; ++i) tau[i] = devf * (b[i] - ((i % 4 == 0) ? btr : 0.0));
    for (int i = 0; i < 3; ++i) tau[i * 4] += hyd;
}

// Weakly compressible fluid: mpm_algorithm.cuh:527-569 and mpm_material.cuh:466-477.
// J <- J (1 + dt tr C); p = K (J^-gamma - 1); tau = J (nu (C + C^T) - p I).
void StressFluid(const Material& m, const double* C, double J, double* tau) {
    const double p = m.bulk * (std::pow(J, -m.gamma) - 1.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau[i * 3 + j] = J * (m.viscosity * (C[i * 3 + j] + C[j * 3 + i]) - (i == j ? p : 0.0));
}

// Compact-kernel 1-D weight of node 0 at fractional offset t in [0,1):
// mpm_kernel.cuh:91-114 (w0 = 1 - t + sin(2 pi t)/(2 pi); node 1 gets 1 - w0).
inline double W0(double t) { return 1.0 - t + std::sin(2.0 * kPi * t) / (2.0 * kPi); }

struct Layout {
    std::vector<uint32_t> keys;
    uint32_t A = 0, Nb = 0, Ex = 0;
    std::vector<uint32_t> table;  // dense key -> list index (kEmpty if absent)
};

struct Model {
    Material mat;
    double mass = 0, volume = 0;
    uint32_t n = 0;
    int nattr = 0;                     // persisted attributes (x,y,z + F / J / logJp)
    std::vector<double> st;            // SoA [attr][n] in storage (slot) order
    std::vector<double> st_next;       // written by the physics step in processing order
    std::vector<float> vel0;           // initial velocities (3 per particle, or 3 total)
    uint32_t nvel = 0;
    std::vector<float> x0;             // initial FP32 positions (3n, AoS)
    std::vector<uint32_t> perm;        // processing position t -> slot | move << 27
    std::vector<uint32_t> setup_order; // processing position at setup -> input index
    std::vector<uint32_t> off;         // A+1 block offsets into processing order
    int pp_attr = 0;                   // state attribute of per-particle (lambda, mu); 0: none
    int rigid_attr = 0;                // state attribute of the rigid flag (1: rigid); 0: none
    double rigid_v[3] = {0, 0, 0};     // velocity of the rigid particles
};

inline int AttrCount(int kind) {
    switch (kind) {
        case kFixedCorotated:
        case kNeoHookean: return 12;
        case kDruckerPrager:
        case kCamClay: return 13;
        case kVonMises: return 12;
        default: return 4;
    }
}

struct Engine {
    Domain dom{};
    float dx_f = 0, inv_dx_f = 0;
    double dx = 0, inv_dx = 0;
    double gravity = 0, mass_clamp = 0;
    uint32_t capacity = 0;
    std::vector<Model> models;
    Layout cur, prev;             // prev: layout in which M is indexed
    std::vector<double> M;        // prev.Nb * 512  [blk][ch 4][node 128]
    std::vector<double> Mnext;    // cur.Nb * 512, produced by the physics step
    std::vector<double> V;        // cur.Nb * 384  [blk][ch 3][node 128]
    uint32_t flags = 0;
    double max_v2 = 0;
    bool physics_pending = false;
    // boundary description (the scene hook ProcessGridCellVelocity as data); custom =
    // false keeps the default sticky 2-block margin
    struct Region {
        int lo[3], hi[3], kind;
        double value[3], center[3];
    };
    bool bc_custom = false;
    int walls = 0;  // 0 sticky, 1 slip, 2 none
    std::vector<Region> regions;
};

// ------------------------------------------------------------------ layouts
// Neighbour (b + {0,1}^3) and exterior (b + {-1,0,1}^3) activation of the first A
// keys, domain-clipped: mpm_algorithm.cuh:657-695; appended in ascending key.
void BuildLayout(Engine& e, Layout& L, const std::vector<uint32_t>& akeys) {
    const Domain& d = e.dom;
    if (L.table.size() != d.nblocks) L.table.assign(d.nblocks, kEmpty);
    for (uint32_t k : L.keys) L.table[k] = kEmpty;
    L.keys = akeys;
    L.A = static_cast<uint32_t>(akeys.size());
    for (uint32_t i = 0; i < L.A; ++i) L.table[akeys[i]] = i;
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<uint32_t> cand;
        for (uint32_t i = 0; i < L.A; ++i) {
            int c[3];
            CoordsOf(d, akeys[i], c);
            const int lo = pass == 0 ? 0 : -1;
            for (int ox = lo; ox <= 1; ++ox)
                for (int oy = lo; oy <= 1; ++oy)
                    for (int oz = lo; oz <= 1; ++oz) {
                        const int bx = c[0] + ox, by = c[1] + oy, bz = c[2] + oz;
                        if (!InDomain(d, bx, by, bz)) continue;
                        const uint32_t k = KeyOf(d, bx, by, bz);
                        if (L.table[k] == kEmpty) cand.push_back(k);
                    }
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (uint32_t k : cand) {
            L.table[k] = static_cast<uint32_t>(L.keys.size());
            L.keys.push_back(k);
        }
        if (pass == 0) L.Nb = static_cast<uint32_t>(L.keys.size());
    }
    L.Ex = static_cast<uint32_t>(L.keys.size());
    if (L.Ex > e.capacity) e.flags |= kFlagCapacity;
}

inline uint32_t NodeBlock(const Engine& e, const Layout& L, int cx, int cy, int cz, int& local) {
    const int bx = FloorDiv4(cx), by = FloorDiv4(cy), bz = FloorDiv4(cz);
    local = (cx - 4 * bx) * 16 + (cy - 4 * by) * 4 + (cz - 4 * bz);
    if (!InDomain(e.dom, bx, by, bz)) return kEmpty;
    const uint32_t idx = L.table[KeyOf(e.dom, bx, by, bz)];
    if (idx == kEmpty || idx >= L.Nb) return kEmpty;
    return idx;
}

// Stencil of one sub-grid at a position: base node c, fractional offsets t, 1-D
// weights w[axis][d].  Grid g = 0 sits at offset -dx/4, g = 1 at +dx/4:
// mpm_algorithm.cuh:1555-1606 (cell = int(x/dx - g*0.25), t = x/dx - cell - 0.25 g).
struct Stencil {
    int c[3];
    double t[3];
    double w[3][2];
};
inline void MakeStencil(const Engine& e, const double* x, int g, Stencil& s) {
    const double sg = g == 0 ? -1.0 : 1.0;
    for (int a = 0; a < 3; ++a) {
        const double u = x[a] * e.inv_dx - 0.25 * sg;
        s.c[a] = static_cast<int>(std::floor(u));
        s.t[a] = u - s.c[a];
        s.w[a][0] = W0(s.t[a]);
        s.w[a][1] = 1.0 - s.w[a][0];
    }
}

// ---------------------------------------------------------------- rasterize
// mpm_algorithm.cuh:871-947: initial P2G of m and m v0 on both grids (no APIC, no
// stress), into grid blocks indexed by the setup layout.
void Rasterize(Engine& e) {
    const Layout& L = e.cur;
    e.M.assign(static_cast<size_t>(L.Nb) * 512, 0.0);
    for (const Model& m : e.models) {
        for (uint32_t p = 0; p < m.n; ++p) {
            double x[3] = {m.x0[3 * p], m.x0[3 * p + 1], m.x0[3 * p + 2]};
            const float* v = m.nvel == 1 ? &m.vel0[0] : &m.vel0[3 * p];
            for (int g = 0; g < 2; ++g) {
                Stencil s;
                MakeStencil(e, x, g, s);
                for (int o = 0; o < 8; ++o) {
                    const int dx_ = (o >> 2) & 1, dy_ = (o >> 1) & 1, dz_ = o & 1;
                    const double w = s.w[0][dx_] * s.w[1][dy_] * s.w[2][dz_];
                    int local;
                    const uint32_t bi = NodeBlock(e, L, s.c[0] + dx_, s.c[1] + dy_, s.c[2] + dz_, local);
                    if (bi == kEmpty) {
                        e.flags |= kFlagMissingBlock;
                        continue;
                    }
                    double* blk = &e.M[static_cast<size_t>(bi) * 512];
                    const int n = g * 64 + local;
                    const double wm = w * m.mass;
                    blk[n] += wm;
                    blk[128 + n] += wm * v[0];
                    blk[256 + n] += wm * v[1];
                    blk[384 + n] += wm * v[2];
                }
            }
        }
    }
    e.prev = e.cur;
}

// ---------------------------------------------------------------- grid update
// mpm_algorithm.cuh:738-869 (UpdateGridVelocityAndQueryMax<true>): v = p/m when
// m > massClamp, gravity on y, whole block zeroed within 2 blocks of the domain
// edge, NaN -> inf, max |v|^2.  Restated as a read of the previous layout's M
// through its key table (this also performs CopySelectedGridBlocks,
// mpm_algorithm.cuh:1920-1945, which only moves blocks between layouts).
// The scene hook of the reference's grid update (mpm_algorithm.cuh:795-799) as data: the
// regions in order on the node's integer cell (fixed: v = 0, velocity: v = value,
// rotation: v = value x (cell dx - center)), then the walls within 8 cells of a face
// (sticky: v = 0; slip: v -= (v . n) n with n the normalized sum of the inward face
// normals), as the fire-hydrant, two-longs and twisting-bar hooks do
// (tests/fire_hydrant/test_fire_hydrant.cuh:148-185,
// tests/fluid_with_two_longs/test_fluid_with_two_longs.cuh:133-161,
// tests/twisting_bar/test_twisting_bar.cuh:126-168).
void ApplyBoundary(const Engine& e, const int* cell, double* v) {
    for (const Engine::Region& g : e.regions) {
        bool in = true;
        for (int k = 0; k < 3; ++k) in = in && cell[k] >= g.lo[k] && cell[k] < g.hi[k];
        if (!in) continue;
        if (g.kind == 0) {
            v[0] = v[1] = v[2] = 0.0;
        } else if (g.kind == 1) {
            for (int k = 0; k < 3; ++k) v[k] = g.value[k];
        } else {
            double q[3];
            for (int k = 0; k < 3; ++k) q[k] = cell[k] * e.dx - g.center[k];
            const double w[3] = {g.value[0], g.value[1], g.value[2]};
            v[0] = w[1] * q[2] - w[2] * q[1];
            v[1] = w[2] * q[0] - w[0] * q[2];
            v[2] = w[0] * q[1] - w[1] * q[0];
        }
    }
    if (e.walls == 2) return;
    double n[3];
    bool on = false;
    for (int k = 0; k < 3; ++k) {
        n[k] = cell[k] < 8 ? 1.0 : (cell[k] >= 4 * e.dom.dim[k] - 8 ? -1.0 : 0.0);
        on = on || n[k] != 0.0;
    }
    if (!on) return;
    if (e.walls == 0) {
        v[0] = v[1] = v[2] = 0.0;
        return;
    }
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double vn = 0.0;
    for (int k = 0; k < 3; ++k) vn += v[k] * n[k] / nn;
    for (int k = 0; k < 3; ++k) v[k] -= vn * n[k] / nn;
}

// The max speed leaves out nodes holding less than 2^-20 of the lightest particle's
// mass, as the engine does (its fixed-point P2G gives such a node's momentum as a few
// units, so p / m there can be far off; particles see the node only with that weight).
// The velocity field itself is the reference's at every node.
void GridUpdate(Engine& e, double dt) {
    const Layout& L = e.cur;
    e.V.assign(static_cast<size_t>(L.Nb) * 384, 0.0);
    double vmax = 0.0;
    double mmin = 0.0;
    for (const Model& m : e.models) mmin = (mmin == 0.0 || m.mass < mmin) ? m.mass : mmin;
    const double floor_mass = std::ldexp(static_cast<double>(static_cast<float>(mmin)), -20);
    uint32_t flags = 0;
#pragma omp parallel for schedule(static) reduction(max : vmax) reduction(| : flags)
    for (uint32_t i = 0; i < L.Nb; ++i) {
        const uint32_t key = L.keys[i];
        int b[3];
        CoordsOf(e.dom, key, b);
        bool oob = false;
        for (int a = 0; a < 3; ++a) oob |= (b[a] + e.dom.off[a] < 2) || (b[a] + e.dom.off[a] + 2 >= e.dom.dim[a]);
        const uint32_t io = e.prev.table.empty() ? kEmpty : e.prev.table[key];
        const double* src = (io != kEmpty && io < e.prev.Nb) ? &e.M[static_cast<size_t>(io) * 512] : nullptr;
        double* dst = &e.V[static_cast<size_t>(i) * 384];
        for (int n = 0; n < 128; ++n) {
            double v2 = 0.0;
            const double mass = src ? src[n] : 0.0;
            double v[3] = {0, 0, 0};
            if (mass > e.mass_clamp) {
                for (int a = 0; a < 3; ++a) v[a] = src[128 * (a + 1) + n] / mass;
                v[1] += e.gravity * dt;
                if (e.bc_custom) {
                    const int cell[3] = {4 * b[0] + ((n >> 4) & 3), 4 * b[1] + ((n >> 2) & 3), 4 * b[2] + (n & 3)};
                    ApplyBoundary(e, cell, v);
                } else if (oob) {
                    v[0] = v[1] = v[2] = 0.0;
                }
                v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            }
            for (int a = 0; a < 3; ++a) dst[128 * a + n] = v[a];
            if (std::isnan(v2)) {
                v2 = INFINITY;
                flags |= kFlagNonFinite;
            }
            if (mass >= floor_mass || std::isinf(v2)) vmax = std::max(vmax, v2);
        }
    }
    e.flags |= flags;
    e.max_v2 = vmax;
}

inline void LoadState(const Model& m, uint32_t s, double* x, double* F, double& J, double& logJp) {
    const size_t n = m.n;
    for (int a = 0; a < 3; ++a) x[a] = m.st[a * n + s];
    if (m.mat.kind == kFluid) {
        J = m.st[3 * n + s];
    } else {
        for (int k = 0; k < 9; ++k) F[k] = m.st[(3 + k) * n + s];
        if (m.mat.kind == kDruckerPrager || m.mat.kind == kCamClay) logJp = m.st[12 * n + s];
    }
}
inline void StoreState(Model& m, uint32_t t, const double* x, const double* F, double J, double logJp) {
    const size_t n = m.n;
    for (int a = 0; a < 3; ++a) m.st_next[a * n + t] = x[a];
    if (m.mat.kind == kFluid) {
        m.st_next[3 * n + t] = J;
    } else {
        for (int k = 0; k < 9; ++k) m.st_next[(3 + k) * n + t] = F[k];
        if (m.mat.kind == kDruckerPrager || m.mat.kind == kCamClay) m.st_next[12 * n + t] = logJp;
    }
}

// ---------------------------------------------------------------- conservation metric
// CollectConservationMetric (mpm_algorithm.cuh:1083-1357), launched after the grid
// update of each substep (mpm_engine.cuh:570-603): for each particle of model mi at its
// current position x, the G2P velocity v = 0.5 sum_g sum_i w v_i on the updated grid
// (:1270-1331), summed as linear sum v and angular sum x cross v (:1334-1348), per unit
// mass.  out = {linear xyz, angular xyz}; the Eulerian entries the reference leaves
// zero are not part of it.
void Conservation(Engine& e, int mi, double* out) {
    const Layout& L = e.cur;
    const Model& m = e.models[mi];
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    for (uint32_t j = 0; j < L.A; ++j)
        for (uint32_t t = m.off[j]; t < m.off[j + 1]; ++t) {
            const uint32_t s = m.perm[t] & kSlotMask;
            double x[3], v[3] = {0, 0, 0};
            for (int a = 0; a < 3; ++a) x[a] = m.st[a * static_cast<size_t>(m.n) + s];
            for (int g = 0; g < 2; ++g) {
                Stencil st;
                MakeStencil(e, x, g, st);
                for (int o = 0; o < 8; ++o) {
                    const int d3[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
                    const double w = st.w[0][d3[0]] * st.w[1][d3[1]] * st.w[2][d3[2]];
                    int local;
                    const uint32_t bi = NodeBlock(e, L, st.c[0] + d3[0], st.c[1] + d3[1], st.c[2] + d3[2], local);
                    if (bi == kEmpty) continue;
                    for (int a = 0; a < 3; ++a) v[a] += w * e.V[static_cast<size_t>(bi) * 384 + 128 * a + g * 64 + local];
                }
            }
            for (int a = 0; a < 3; ++a) v[a] *= 0.5;
            out[0] += v[0];
            out[1] += v[1];
            out[2] += v[2];
            out[3] += x[1] * v[2] - x[2] * v[1];
            out[4] += x[2] * v[0] - x[0] * v[2];
            out[5] += x[0] * v[1] - x[1] * v[0];
        }
}

// ---------------------------------------------------------------- G2P2G
// mpm_algorithm.cuh:1365-1845 per particle block, in list order, particles in
// processing order (perm), FP64 throughout:
//   G2P (:1555-1606): v = 0.5 sum w v_i, B = sum v_i (w (x_i - x_p))^T,
//                     D = sum (x_i - x_p)(w (x_i - x_p))^T over both grids;
//   C = B D^-1, x += dt v (:1636-1638);
//   F / J / logJp update and Kirchhoff stress tau (UpdateForce :192-327, :527-569);
//   window check on the new cell (:1669-1683) -> flag and skip P2G;
//   D' at the new x, D^ = (0.5 D')^-1, K = m C - nextDt V D^ tau (:1687-1736);
//   P2G m w, w (m v + K (x_i - x_p)) on both grids (:1739-1792).
// Parallel form (OpenMP): the blocks are split into 8 colours by the parity of their
// coordinates and the colours run one after the other.  Two blocks of one colour are at
// least two blocks apart, so their P2G windows (the block and its +1 neighbours) are
// disjoint and the blocks of a colour run concurrently without sharing a node; each
// node receives its contributions in a fixed order (by colour, then list order), so the
// result does not depend on the thread count.
void Physics(Engine& e, double dt, double next_dt) {
    const Layout& L = e.cur;
    e.Mnext.assign(static_cast<size_t>(L.Nb) * 512, 0.0);
    std::vector<uint32_t> colour[8];
    for (uint32_t j = 0; j < L.A; ++j) {
        int bb[3];
        CoordsOf(e.dom, L.keys[j], bb);
        colour[(bb[0] & 1) * 4 + (bb[1] & 1) * 2 + (bb[2] & 1)].push_back(j);
    }
    uint32_t flags = 0;
    for (Model& m : e.models) {
        m.st_next.assign(m.st.size(), 0.0);
        for (int col = 0; col < 8; ++col) {
            const std::vector<uint32_t>& list = colour[col];
#pragma omp parallel for schedule(dynamic, 4) reduction(| : flags)
            for (size_t q = 0; q < list.size(); ++q) {
            const uint32_t j = list[q];
            int bb[3];
            CoordsOf(e.dom, L.keys[j], bb);
            for (uint32_t t = m.off[j]; t < m.off[j + 1]; ++t) {
                const uint32_t s = m.perm[t] & kSlotMask;
                double x[3], F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, J = 1.0, logJp = 0.0;
                LoadState(m, s, x, F, J, logJp);
                double v[3] = {0, 0, 0}, B[9] = {0}, D[9] = {0};
                for (int g = 0; g < 2; ++g) {
                    Stencil st;
                    MakeStencil(e, x, g, st);
                    for (int o = 0; o < 8; ++o) {
                        const int d3[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
                        const double w = st.w[0][d3[0]] * st.w[1][d3[1]] * st.w[2][d3[2]];
                        int local;
                        const uint32_t bi = NodeBlock(e, L, st.c[0] + d3[0], st.c[1] + d3[1], st.c[2] + d3[2], local);
                        double vi[3] = {0, 0, 0};
                        if (bi == kEmpty)
                            flags |= kFlagMissingBlock;
                        else
                            for (int a = 0; a < 3; ++a) vi[a] = e.V[static_cast<size_t>(bi) * 384 + 128 * a + g * 64 + local];
                        double xx[3];
                        for (int a = 0; a < 3; ++a) xx[a] = (d3[a] - st.t[a]) * e.dx;
                        for (int a = 0; a < 3; ++a) {
                            v[a] += w * vi[a];
                            for (int c = 0; c < 3; ++c) {
                                B[a * 3 + c] += vi[a] * w * xx[c];
                                D[a * 3 + c] += xx[a] * w * xx[c];
                            }
                        }
                    }
                }
                double Dinv[9], C[9];
                Inverse(D, Dinv);
                MatMul(B, Dinv, C);
                const size_t Nn = m.n;
                // rigid particles (GetExistRigidParticle, mpm_algorithm.cuh:1609-1624): no
                // velocity or velocity gradient from the grid, motion at the prescribed
                // velocity; the stress of their (unchanged) F is still deposited
                const bool rigid = m.rigid_attr && m.st[m.rigid_attr * Nn + s] != 0.0;
                if (m.rigid_attr) m.st_next[m.rigid_attr * Nn + t] = m.st[m.rigid_attr * Nn + s];
                for (int a = 0; a < 3; ++a) {
                    v[a] *= 0.5;
                    if (rigid) v[a] = 0.0;
                    x[a] += dt * (rigid ? m.rigid_v[a] : v[a]);
                }
                if (rigid)
                    for (double& c : C) c = 0.0;
                const double lam = m.pp_attr ? m.st[m.pp_attr * Nn + s] : m.mat.lambda;
                const double mu = m.pp_attr ? m.st[(m.pp_attr + 1) * Nn + s] : m.mat.mu;
                if (m.pp_attr) {
                    m.st_next[m.pp_attr * Nn + t] = lam;
                    m.st_next[(m.pp_attr + 1) * Nn + t] = mu;
                }
                double tau[9];
                if (m.mat.kind == kFluid) {
                    J += (C[0] + C[4] + C[8]) * dt * J;
                    StressFluid(m.mat, C, J, tau);
                } else {
                    double dW[9];
                    for (int k = 0; k < 9; ++k) dW[k] = C[k] * dt + ((k % 4 == 0) ? 1.0 : 0.0);
                    MatMul(dW, F, F);
                    if (m.mat.kind == kFixedCorotated)
                        StressFixedCorotated(lam, mu, F, tau);
                    else if (m.mat.kind == kNeoHookean)
                        StressNeoHookean(lam, mu, F, tau);
                    else if (m.mat.kind == kDruckerPrager)
                        StressDruckerPrager(lam, mu, m.mat.sin_phi, F, logJp, tau);
                    else if (m.mat.kind == kVonMises)
                        StressVonMises(lam, mu, m.mat.yield, F, tau);
                    else
                        StressCamClay(m.mat, lam, mu, F, logJp, tau);
                }
                StoreState(m, t, x, F, J, logJp);

                // window check on the new grid-0 cell, relative to this block's origin
                bool out = false;
                for (int a = 0; a < 3; ++a) {
                    const int c0 = static_cast<int>(std::floor(x[a] * e.inv_dx + 0.25));
                    const int loc = c0 - 4 * bb[a];
                    out |= (loc <= 0) || (loc >= 7);
                }
                if (out) {
                    flags |= kFlagOutOfWindow;
                    continue;
                }
                double Dn[9] = {0};
                Stencil sn[2];
                for (int g = 0; g < 2; ++g) {
                    MakeStencil(e, x, g, sn[g]);
                    for (int o = 0; o < 8; ++o) {
                        const int d3[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
                        const double w = sn[g].w[0][d3[0]] * sn[g].w[1][d3[1]] * sn[g].w[2][d3[2]];
                        double xx[3];
                        for (int a = 0; a < 3; ++a) xx[a] = (d3[a] - sn[g].t[a]) * e.dx;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) Dn[a * 3 + c] += xx[a] * w * xx[c];
                    }
                }
                for (double& q : Dn) q *= 0.5;
                double Dh[9], DT[9], K[9];
                Inverse(Dn, Dh);
                MatMul(Dh, tau, DT);
                for (int k = 0; k < 9; ++k) K[k] = C[k] * m.mass - (next_dt * m.volume) * DT[k];
                for (int g = 0; g < 2; ++g) {
                    const Stencil& st = sn[g];
                    for (int o = 0; o < 8; ++o) {
                        const int d3[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
                        const double w = st.w[0][d3[0]] * st.w[1][d3[1]] * st.w[2][d3[2]];
                        int local;
                        const uint32_t bi = NodeBlock(e, L, st.c[0] + d3[0], st.c[1] + d3[1], st.c[2] + d3[2], local);
                        if (bi == kEmpty) {
                            flags |= kFlagMissingBlock;
                            continue;
                        }
                        double xx[3];
                        for (int a = 0; a < 3; ++a) xx[a] = (d3[a] - st.t[a]) * e.dx;
                        double* blk = &e.Mnext[static_cast<size_t>(bi) * 512];
                        const int n = g * 64 + local;
                        const double wm = w * m.mass;
                        blk[n] += wm;
                        for (int a = 0; a < 3; ++a) {
                            const double km = K[a * 3 + 0] * xx[0] + K[a * 3 + 1] * xx[1] + K[a * 3 + 2] * xx[2];
                            blk[128 * (a + 1) + n] += wm * v[a] + w * km;
                        }
                    }
                }
            }
            }
        }
    }
    e.flags |= flags;
    e.physics_pending = true;
}

// ---------------------------------------------------------------- rebin
// Replaces AddAdvection / ParticleCellBucketToBlock / MarkActive* / ExclusiveScan /
// UpdatePartition / UpdateBucket / ComputeBinCount (mpm_particle.cuh:331-371,
// mpm_algorithm.cuh:572-649,1847-1918, mpm_engine.cuh:659-866) by the canonical
// order: each particle's new block is its FP32 grid-0 block, clamped to the 27
// neighbours of its current block and to the domain; the new particle-block list is
// the current exterior list with unoccupied blocks removed; particles are stably
// sorted by (new block index, move).  xf[m] holds each model's FP32 positions relative to the
// origin of the block that processed them, in processing order (3 per particle, AoS).
void Rebin(Engine& e, const std::vector<const float*>& xf) {
    const Layout& L = e.cur;
    const Domain& d = e.dom;
    std::vector<uint32_t> occ(L.Ex, 0);
    std::vector<std::vector<uint32_t>> target(e.models.size()), move(e.models.size());
    for (size_t mi = 0; mi < e.models.size(); ++mi) {
        Model& m = e.models[mi];
        target[mi].resize(m.n);
        move[mi].resize(m.n);
        for (uint32_t j = 0; j < L.A; ++j) {
            int bb[3];
            CoordsOf(d, L.keys[j], bb);
            for (uint32_t t = m.off[j]; t < m.off[j + 1]; ++t) {
                int b[3], c[3];
                BlockOfF(&xf[mi][3 * static_cast<size_t>(t)], e.inv_dx_f, b, c);  // local block
                int nb[3];
                bool clamped = false;
                for (int a = 0; a < 3; ++a) {
                    int delta = b[a];
                    if (delta < -1 || delta > 1) clamped = true;
                    delta = std::max(-1, std::min(1, delta));
                    nb[a] = bb[a] + delta;
                }
                if (clamped) e.flags |= kFlagOutOfWindow;
                if (!InDomain(d, nb[0], nb[1], nb[2])) {
                    e.flags |= kFlagOutOfDomain;
                    nb[0] = bb[0];
                    nb[1] = bb[1];
                    nb[2] = bb[2];
                }
                const uint32_t ei = L.table[KeyOf(d, nb[0], nb[1], nb[2])];
                target[mi][t] = ei;
                move[mi][t] = static_cast<uint32_t>((nb[0] - bb[0] + 1) * 9 + (nb[1] - bb[1] + 1) * 3 + (nb[2] - bb[2] + 1));
                occ[ei]++;
            }
        }
    }
    // new particle-block list: exterior list order, unoccupied removed
    std::vector<uint32_t> newidx(L.Ex, kEmpty), akeys;
    for (uint32_t i = 0; i < L.Ex; ++i)
        if (occ[i]) {
            newidx[i] = static_cast<uint32_t>(akeys.size());
            akeys.push_back(L.keys[i]);
        }
    const uint32_t A = static_cast<uint32_t>(akeys.size());
    for (size_t mi = 0; mi < e.models.size(); ++mi) {
        Model& m = e.models[mi];
        // counting sort by (new block, move): bucket j * 27 + move
        std::vector<uint32_t> cnt(static_cast<size_t>(A) * 27 + 1, 0);
        for (uint32_t t = 0; t < m.n; ++t) cnt[static_cast<size_t>(newidx[target[mi][t]]) * 27 + move[mi][t]]++;
        std::vector<uint32_t> fill(static_cast<size_t>(A) * 27 + 1, 0);
        for (size_t q = 0; q < static_cast<size_t>(A) * 27; ++q) fill[q + 1] = fill[q] + cnt[q];
        std::vector<uint32_t> off(A + 1, 0);
        for (uint32_t j = 0; j <= A; ++j) off[j] = fill[static_cast<size_t>(j) * 27];
        std::vector<uint32_t> perm(m.n);
        for (uint32_t t = 0; t < m.n; ++t)
            perm[fill[static_cast<size_t>(newidx[target[mi][t]]) * 27 + move[mi][t]]++] = t | (move[mi][t] << kMoveShift);
        m.perm.swap(perm);
        m.off.swap(off);
    }
    Layout next;
    next.table.swap(e.prev.table);  // reuse the oldest table's storage
    next.keys.swap(e.prev.keys);
    BuildLayout(e, next, akeys);
    e.prev = std::move(e.cur);
    e.cur = std::move(next);
}

}  // namespace

// =========================================================================== C ABI
extern "C" {

struct orc_engine;

orc_engine* orc_create(const int dim[3], const int offset[3], float dx, double gravity, double mass_clamp,
                       uint32_t capacity) {
    Engine* e = new Engine();
    for (int a = 0; a < 3; ++a) {
        e->dom.dim[a] = dim[a];
        e->dom.off[a] = offset ? offset[a] : 0;
    }
    e->dom.nblocks = static_cast<uint64_t>(dim[0]) * dim[1] * dim[2];
    e->dx_f = dx;
    e->inv_dx_f = static_cast<float>(1.0 / static_cast<double>(dx));  // mpm_config.h:22-25
    e->dx = dx;
    e->inv_dx = 1.0 / static_cast<double>(dx);
    e->gravity = gravity;
    e->mass_clamp = mass_clamp;
    e->capacity = capacity;
    return reinterpret_cast<orc_engine*>(e);
}

void orc_destroy(orc_engine* h) { delete reinterpret_cast<Engine*>(h); }

// params: {lambda, mu, bulk, gamma, viscosity, J0, friction_angle_deg}
int orc_add_model(orc_engine* h, int kind, const double* params, double mass, double volume, const float* xyz,
                  uint32_t n, const float* vel, uint32_t nvel, const float* lam_p, const float* mu_p) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Model m;
    MaterialFromParams(kind, params, m.mat);
    m.mass = mass;
    m.volume = volume;
    m.n = n;
    m.nattr = AttrCount(kind);
    m.x0.assign(xyz, xyz + 3 * static_cast<size_t>(n));
    m.nvel = nvel;
    m.vel0.assign(vel, vel + 3 * static_cast<size_t>(nvel));
    if (lam_p && mu_p) {
        m.pp_attr = m.nattr;
        m.nattr += 2;
    }
    // initial state in input order: x, F = I (solids), logJp = 0 (DP) or alpha0 (Cam-clay),
    // J = J0 (fluid) (CopyParticleArrayToParticleBuffer, mpm_algorithm.cuh:973-1081)
    m.st.assign(static_cast<size_t>(m.nattr) * n, 0.0);
    for (uint32_t p = 0; p < n; ++p)
        for (int a = 0; a < 3; ++a) m.st[a * static_cast<size_t>(n) + p] = xyz[3 * static_cast<size_t>(p) + a];
    if (m.pp_attr)
        for (uint32_t p = 0; p < n; ++p) {
            m.st[m.pp_attr * static_cast<size_t>(n) + p] = lam_p[p];
            m.st[(m.pp_attr + 1) * static_cast<size_t>(n) + p] = mu_p[p];
        }
    if (kind == kFluid) {
        for (uint32_t p = 0; p < n; ++p) m.st[3 * static_cast<size_t>(n) + p] = m.mat.J0;
    } else {
        for (int k = 0; k < 9; k += 4)
            for (uint32_t p = 0; p < n; ++p) m.st[(3 + k) * static_cast<size_t>(n) + p] = 1.0;
        if (kind == kCamClay)
            for (uint32_t p = 0; p < n; ++p) m.st[12 * static_cast<size_t>(n) + p] = m.mat.alpha0;
    }
    e->models.push_back(std::move(m));
    return static_cast<int>(e->models.size()) - 1;
}

// InitialSetup (mpm_engine.cuh:215-383): particle blocks from FP32 positions,
// ascending key; particles stably sorted by block; neighbour / exterior layout;
// rasterize.  Returns the flags.
// Rigid particles: the first `count` particles of model mi (input order) carry the
// rigid flag; before orc_initial_setup (the flag becomes one more state attribute), and
// the velocity may change at any time.
void orc_set_rigid(orc_engine* h, int mi, uint32_t count, const double* v) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Model& m = e->models[mi];
    if (count && !m.rigid_attr) {
        m.rigid_attr = m.nattr;
        m.nattr += 1;
        m.st.resize(static_cast<size_t>(m.nattr) * m.n, 0.0);
        for (uint32_t p = 0; p < std::min(count, m.n); ++p) m.st[m.rigid_attr * static_cast<size_t>(m.n) + p] = 1.0;
    }
    for (int a = 0; a < 3; ++a) m.rigid_v[a] = v[a];
}

uint32_t orc_initial_setup(orc_engine* h) {
    Engine* e = reinterpret_cast<Engine*>(h);
    std::vector<uint32_t> keys;
    std::vector<std::vector<uint32_t>> pkey(e->models.size()), pcell(e->models.size());
    for (size_t mi = 0; mi < e->models.size(); ++mi) {
        Model& m = e->models[mi];
        pkey[mi].resize(m.n);
        pcell[mi].resize(m.n);
        for (uint32_t p = 0; p < m.n; ++p) {
            int b[3], c[3];
            BlockOfF(&m.x0[3 * static_cast<size_t>(p)], e->inv_dx_f, b, c);
            if (!InDomain(e->dom, b[0], b[1], b[2])) {
                e->flags |= kFlagOutOfDomain;
                for (int a = 0; a < 3; ++a) b[a] = std::max(0, std::min(e->dom.dim[a] - 1, b[a]));
            }
            pkey[mi][p] = KeyOf(e->dom, b[0], b[1], b[2]);
            // binning cell inside the block from the unclamped cell, (c - 2) mod 4 per axis
            pcell[mi][p] = static_cast<uint32_t>(((c[0] - 2) & 3) * 16 + ((c[1] - 2) & 3) * 4 + ((c[2] - 2) & 3));
            keys.push_back(pkey[mi][p]);
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    BuildLayout(*e, e->cur, keys);
    const uint32_t A = e->cur.A;
    for (size_t mi = 0; mi < e->models.size(); ++mi) {
        Model& m = e->models[mi];
        std::vector<uint32_t> cnt(A, 0);
        for (uint32_t p = 0; p < m.n; ++p) cnt[e->cur.table[pkey[mi][p]]]++;
        m.off.assign(A + 1, 0);
        for (uint32_t j = 0; j < A; ++j) m.off[j + 1] = m.off[j] + cnt[j];
        // Setup order (this engine's canonical choice; the reference's is atomic): by
        // block, then by the particle's rank among the earlier input particles of its
        // binning cell (clamped to 2^26 - 1), then by the cell; ties keep input order.
        std::unordered_map<uint64_t, uint32_t> seen;
        std::vector<uint64_t> sk(m.n);
        for (uint32_t p = 0; p < m.n; ++p) {
            const uint64_t ck = (static_cast<uint64_t>(pkey[mi][p]) << 6) | pcell[mi][p];
            const uint32_t rank = std::min<uint32_t>(seen[ck]++, (1u << 26) - 1u);
            sk[p] = (static_cast<uint64_t>(e->cur.table[pkey[mi][p]]) << 32) | (static_cast<uint64_t>(rank) << 6) |
                    pcell[mi][p];
        }
        m.setup_order.resize(m.n);
        for (uint32_t p = 0; p < m.n; ++p) m.setup_order[p] = p;
        std::stable_sort(m.setup_order.begin(), m.setup_order.end(),
                         [&](uint32_t a, uint32_t b) { return sk[a] < sk[b]; });
        // state in processing order; the ordering becomes the identity with no move
        // (CopyParticleArrayToParticleBuffer + InitializeAdvectionBucket,
        // mpm_algorithm.cuh:949-1081)
        std::vector<double> st(m.st.size());
        for (int a = 0; a < m.nattr; ++a)
            for (uint32_t t = 0; t < m.n; ++t) st[a * static_cast<size_t>(m.n) + t] = m.st[a * static_cast<size_t>(m.n) + m.setup_order[t]];
        m.st.swap(st);
        m.perm.resize(m.n);
        for (uint32_t t = 0; t < m.n; ++t) m.perm[t] = t | kNoMove;
    }
    Rasterize(*e);
    return e->flags;
}

// Grid update with dt, then G2P2G(dt, next_dt); the rebin is separate.
uint32_t orc_physics(orc_engine* h, double dt, double next_dt) {
    Engine* e = reinterpret_cast<Engine*>(h);
    GridUpdate(*e, dt);
    Physics(*e, dt, next_dt);
    return e->flags;
}
// The two halves of orc_physics, for the frame loop (mpm_engine.cuh:517-590), which
// chooses the next dt from the grid update's max |v| before the particles move.
uint32_t orc_grid_update(orc_engine* h, double dt) {
    Engine* e = reinterpret_cast<Engine*>(h);
    GridUpdate(*e, dt);
    return e->flags;
}
uint32_t orc_particles(orc_engine* h, double dt, double next_dt) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Physics(*e, dt, next_dt);
    return e->flags;
}

// Rebin using the given FP32 positions (per model, AoS, processing order), or the
// oracle's own FP64 positions rounded to FP32 when xyz_per_model is NULL.  Commits
// the physics step's state and grid.
uint32_t orc_rebin(orc_engine* h, const float* const* xyz_per_model) {
    Engine* e = reinterpret_cast<Engine*>(h);
    std::vector<std::vector<float>> own;
    own.reserve(e->models.size());
    std::vector<const float*> ptr(e->models.size());
    for (size_t mi = 0; mi < e->models.size(); ++mi) {
        Model& m = e->models[mi];
        if (xyz_per_model && xyz_per_model[mi]) {
            ptr[mi] = xyz_per_model[mi];
        } else {
            // FP32 position relative to the origin of the processing block
            own.emplace_back(3 * static_cast<size_t>(m.n));
            const Layout& L = e->cur;
            for (uint32_t j = 0; j < L.A; ++j) {
                int bb[3];
                CoordsOf(e->dom, L.keys[j], bb);
                for (uint32_t t = m.off[j]; t < m.off[j + 1]; ++t)
                    for (int a = 0; a < 3; ++a)
                        own.back()[3 * static_cast<size_t>(t) + a] =
                            static_cast<float>(m.st_next[a * static_cast<size_t>(m.n) + t] - 4.0 * bb[a] * e->dx);
            }
            ptr[mi] = own.back().data();
        }
    }
    if (e->physics_pending) {
        for (Model& m : e->models) m.st.swap(m.st_next);
        e->M.swap(e->Mnext);
        e->physics_pending = false;
    }
    Rebin(*e, ptr);
    return e->flags;
}

uint32_t orc_substep(orc_engine* h, double dt, double next_dt) {
    orc_physics(h, dt, next_dt);
    return orc_rebin(h, nullptr);
}

// Boundary description: custom = 0 restores the default; region r has cells box[6 r ..
// 6 r + 5] = lo xyz, hi xyz, kind kinds[r], value vals[6 r .. 6 r + 2], center
// vals[6 r + 3 .. 6 r + 5].
void orc_set_boundary(orc_engine* h, int custom, int walls, uint32_t nregions, const int* box, const int* kinds,
                      const double* vals) {
    Engine* e = reinterpret_cast<Engine*>(h);
    e->bc_custom = custom != 0;
    e->walls = walls;
    e->regions.clear();
    for (uint32_t r = 0; r < nregions; ++r) {
        Engine::Region g;
        for (int k = 0; k < 3; ++k) {
            g.lo[k] = box[6 * r + k];
            g.hi[k] = box[6 * r + 3 + k];
            g.value[k] = vals[6 * r + k];
            g.center[k] = vals[6 * r + 3 + k];
        }
        g.kind = kinds[r];
        e->regions.push_back(g);
    }
}

// Conservation metric of model mi on the current grid velocities and particle state:
// call after orc_physics (which leaves both unchanged) and before orc_rebin.
void orc_conservation(orc_engine* h, int mi, double* out) {
    Conservation(*reinterpret_cast<Engine*>(h), mi, out);
}
// Replace the grid velocities of the current layout: Nb blocks x [vx,vy,vz][128]
void orc_set_grid_v(orc_engine* h, const double* V) {
    Engine* e = reinterpret_cast<Engine*>(h);
    e->V.assign(V, V + static_cast<size_t>(e->cur.Nb) * 384);
}

uint32_t orc_flags(orc_engine* h) { return reinterpret_cast<Engine*>(h)->flags; }
void orc_clear_flags(orc_engine* h) { reinterpret_cast<Engine*>(h)->flags = 0; }
// OpenMP threads of the physics and grid update (the results do not depend on it)
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads() { return omp_get_max_threads(); }
double orc_max_v2(orc_engine* h) { return reinterpret_cast<Engine*>(h)->max_v2; }

// which: 0 = current layout, 1 = previous layout (the one M is indexed by)
void orc_layout_counts(orc_engine* h, int which, uint32_t out[3]) {
    Engine* e = reinterpret_cast<Engine*>(h);
    const Layout& L = which ? e->prev : e->cur;
    out[0] = L.A;
    out[1] = L.Nb;
    out[2] = L.Ex;
}
void orc_layout_keys(orc_engine* h, int which, uint32_t* keys) {
    Engine* e = reinterpret_cast<Engine*>(h);
    const Layout& L = which ? e->prev : e->cur;
    std::memcpy(keys, L.keys.data(), L.keys.size() * sizeof(uint32_t));
}
uint32_t orc_model_size(orc_engine* h, int mi) { return reinterpret_cast<Engine*>(h)->models[mi].n; }
int orc_model_attrs(orc_engine* h, int mi) { return reinterpret_cast<Engine*>(h)->models[mi].nattr; }
void orc_get_perm(orc_engine* h, int mi, uint32_t* out) {
    const Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    std::memcpy(out, m.perm.data(), m.perm.size() * sizeof(uint32_t));
}
void orc_get_offsets(orc_engine* h, int mi, uint32_t* out) {
    const Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    std::memcpy(out, m.off.data(), m.off.size() * sizeof(uint32_t));
}
// committed state (SoA [attr][n], storage order)
void orc_get_state(orc_engine* h, int mi, double* out) {
    const Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    std::memcpy(out, m.st.data(), m.st.size() * sizeof(double));
}
// state produced by the last physics step (processing order), before the rebin
void orc_get_state_next(orc_engine* h, int mi, double* out) {
    const Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    std::memcpy(out, m.st_next.data(), m.st_next.size() * sizeof(double));
}
// M: previous-layout grid (after a rebin) ; Mnext: grid produced by the physics step
void orc_get_grid_m(orc_engine* h, double* out) {
    Engine* e = reinterpret_cast<Engine*>(h);
    std::memcpy(out, e->M.data(), e->M.size() * sizeof(double));
}
void orc_get_grid_mnext(orc_engine* h, double* out) {
    Engine* e = reinterpret_cast<Engine*>(h);
    std::memcpy(out, e->Mnext.data(), e->Mnext.size() * sizeof(double));
}
void orc_get_grid_v(orc_engine* h, double* out) {
    Engine* e = reinterpret_cast<Engine*>(h);
    std::memcpy(out, e->V.data(), e->V.size() * sizeof(double));
}

// ---- lockstep import: replace the engine's layouts, orderings, state and grid
// with externally supplied ones (e.g. the GPU engine's), so that one physics step
// and one rebin can be compared on identical inputs.
void orc_set_layout(orc_engine* h, int which, const uint32_t* keys, uint32_t A, uint32_t Nb, uint32_t Ex) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Layout& L = which ? e->prev : e->cur;
    if (L.table.size() != e->dom.nblocks) L.table.assign(e->dom.nblocks, kEmpty);
    for (uint32_t k : L.keys) L.table[k] = kEmpty;
    L.keys.assign(keys, keys + Ex);
    L.A = A;
    L.Nb = Nb;
    L.Ex = Ex;
    for (uint32_t i = 0; i < Ex; ++i) L.table[keys[i]] = i;
}
void orc_set_ordering(orc_engine* h, int mi, const uint32_t* perm, const uint32_t* off) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Model& m = e->models[mi];
    m.perm.assign(perm, perm + m.n);
    m.off.assign(off, off + e->cur.A + 1);
}
void orc_set_state_f32(orc_engine* h, int mi, const float* soa) {
    Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    for (size_t i = 0; i < m.st.size(); ++i) m.st[i] = soa[i];
}
// global positions (FP64) and the remaining attributes, SoA, storage order
void orc_set_state_f64(orc_engine* h, int mi, const double* soa) {
    Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    std::memcpy(m.st.data(), soa, m.st.size() * sizeof(double));
}
void orc_get_setup_order(orc_engine* h, int mi, uint32_t* out) {
    const Model& m = reinterpret_cast<Engine*>(h)->models[mi];
    std::memcpy(out, m.setup_order.data(), m.setup_order.size() * sizeof(uint32_t));
}
void orc_set_grid_m_f32(orc_engine* h, const float* M) {
    Engine* e = reinterpret_cast<Engine*>(h);
    e->M.resize(static_cast<size_t>(e->prev.Nb) * 512);
    for (size_t i = 0; i < e->M.size(); ++i) e->M[i] = M[i];
}

// ---- unit-level entry points used to pin the restatement in tests
void orc_svd(const double* F, double* U, double* s, double* V) { SignedSvd(F, U, s, V); }
void orc_stress(int kind, const double* params, const double* F_in, double logJp_in, const double* C, double J,
                double* F_out, double* logJp_out, double* tau) {
    Material m;
    MaterialFromParams(kind, params, m);
    double F[9];
    std::memcpy(F, F_in, sizeof(F));
    double lj = logJp_in;
    if (kind == kFixedCorotated) StressFixedCorotated(m.lambda, m.mu, F, tau);
    else if (kind == kNeoHookean) StressNeoHookean(m.lambda, m.mu, F, tau);
    else if (kind == kDruckerPrager) StressDruckerPrager(m.lambda, m.mu, m.sin_phi, F, lj, tau);
    else if (kind == kVonMises) StressVonMises(m.lambda, m.mu, m.yield, F, tau);
    else if (kind == kCamClay) StressCamClay(m, m.lambda, m.mu, F, lj, tau);
    else StressFluid(m, C, J, tau);
    std::memcpy(F_out, F, sizeof(F));
    *logJp_out = lj;
}
// 1-D weight, stencil of a position (grid g), binning of FP32 positions
double orc_weight0(double t) { return W0(t); }
void orc_stencil(orc_engine* h, const double* x, int g, int* c, double* t, double* w) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Stencil s;
    MakeStencil(*e, x, g, s);
    for (int a = 0; a < 3; ++a) {
        c[a] = s.c[a];
        t[a] = s.t[a];
        w[2 * a] = s.w[a][0];
        w[2 * a + 1] = s.w[a][1];
    }
}
void orc_bin_f32(orc_engine* h, const float* xyz, uint32_t n, int* block_xyz, int* cell_xyz) {
    Engine* e = reinterpret_cast<Engine*>(h);
    for (uint32_t p = 0; p < n; ++p) BlockOfF(&xyz[3 * static_cast<size_t>(p)], e->inv_dx_f, &block_xyz[3 * p], &cell_xyz[3 * p]);
}
// Neighbour / exterior extension of an ordered particle-block key list.
uint32_t orc_build_layout(orc_engine* h, const uint32_t* akeys, uint32_t A, uint32_t* out_keys, uint32_t* out_counts) {
    Engine* e = reinterpret_cast<Engine*>(h);
    Layout L;
    BuildLayout(*e, L, std::vector<uint32_t>(akeys, akeys + A));
    std::memcpy(out_keys, L.keys.data(), L.keys.size() * sizeof(uint32_t));
    out_counts[0] = L.A;
    out_counts[1] = L.Nb;
    out_counts[2] = L.Ex;
    return L.Ex;
}

// ------------------------------------------------------------ sparse block table
// MurmurHash3_x86_32 of a byte string (Appleby's published algorithm; the reference
// hashes its 12-byte int3 keys with it, gpu_hash_table.cuh:39-118).
uint32_t orc_murmur3_32(const uint8_t* data, uint32_t len, uint32_t seed) {
    auto rotl = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t h = seed;
    const uint32_t nblocks = len / 4;
    for (uint32_t i = 0; i < nblocks; ++i) {
        uint32_t k = static_cast<uint32_t>(data[4 * i]) | static_cast<uint32_t>(data[4 * i + 1]) << 8 |
                     static_cast<uint32_t>(data[4 * i + 2]) << 16 | static_cast<uint32_t>(data[4 * i + 3]) << 24;
        k *= c1;
        k = rotl(k, 15);
        k *= c2;
        h ^= k;
        h = rotl(h, 13);
        h = h * 5u + 0xe6546b64u;
    }
    const uint8_t* tail = data + 4 * nblocks;
    uint32_t k = 0;
    switch (len & 3u) {
        case 3: k ^= static_cast<uint32_t>(tail[2]) << 16; [[fallthrough]];
        case 2: k ^= static_cast<uint32_t>(tail[1]) << 8; [[fallthrough]];
        case 1:
            k ^= tail[0];
            k *= c1;
            k = rotl(k, 15);
            k *= c2;
            h ^= k;
    }
    h ^= len;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// The reference's three seeds: first draws of std::mt19937(2) (gpu_hash_table.cuh:303-309).
void orc_hash_seeds(uint32_t out[3]) {
    std::mt19937 rng(2);
    for (int r = 0; r < 3; ++r) out[r] = static_cast<uint32_t>(rng());
}

// Bucket count for at most max_keys pairs: load factor <= 0.7 with 16-slot buckets.
uint32_t orc_hash_buckets(uint32_t max_keys) {
    return static_cast<uint32_t>((static_cast<uint64_t>(max_keys) * 10 + 7 * 16 - 1) / (7 * 16));
}

static uint32_t HashBucketOf(const int dim[3], uint32_t buckets, uint32_t seed, uint32_t key) {
    const int32_t c[3] = {static_cast<int32_t>(key / (static_cast<uint32_t>(dim[1]) * dim[2])),
                          static_cast<int32_t>((key / dim[2]) % dim[1]), static_cast<int32_t>(key % dim[2])};
    uint8_t b[12];
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 4; ++i) b[4 * a + i] = static_cast<uint8_t>(static_cast<uint32_t>(c[a]) >> (8 * i));
    return orc_murmur3_32(b, 12, seed) % buckets;
}

// Deterministic build: pairs in ascending key order; in round r every unplaced pair
// asks for bucket h_r; the pairs asking for one bucket take its free slots in key
// order; leftovers after three rounds go to the stash in key order.  Returns the
// number of stash pairs (also when it exceeds stash_cap; only stash_cap are written).
uint32_t orc_hash_build(const int dim[3], uint32_t buckets, const uint32_t seeds[3], const uint32_t* keys,
                        const uint32_t* vals, uint32_t n, uint64_t* slots, uint64_t* stash, uint32_t stash_cap) {
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (uint64_t i = 0; i < static_cast<uint64_t>(buckets) * 16; ++i) slots[i] = ~0ull;
    std::vector<uint32_t> fill(buckets, 0);
    std::vector<uint32_t> rest = order;
    for (int r = 0; r < 3; ++r) {
        std::vector<uint32_t> asked(buckets, 0), next;
        for (uint32_t p : rest) {  // ascending key order
            const uint32_t b = HashBucketOf(dim, buckets, seeds[r], keys[p]);
            const uint32_t rank = asked[b]++;
            if (rank < 16 - fill[b]) slots[static_cast<uint64_t>(b) * 16 + fill[b] + rank] =
                (static_cast<uint64_t>(keys[p]) << 32) | vals[p];
            else next.push_back(p);
        }
        for (uint32_t b = 0; b < buckets; ++b) fill[b] = std::min<uint32_t>(16, fill[b] + asked[b]);
        rest.swap(next);
    }
    for (uint32_t i = 0; i < rest.size() && i < stash_cap; ++i)
        stash[i] = (static_cast<uint64_t>(keys[rest[i]]) << 32) | vals[rest[i]];
    return static_cast<uint32_t>(rest.size());
}

// Find with the reference's protocol: probe h0, h1, h2; stop at a bucket that is not
// full; then the stash.  Missing keys give 0xFFFFFFFF.
void orc_hash_find(const int dim[3], uint32_t buckets, const uint32_t seeds[3], const uint64_t* slots,
                   const uint64_t* stash, uint32_t nstash, const uint32_t* q, uint32_t nq, uint32_t* out) {
    for (uint32_t i = 0; i < nq; ++i) {
        uint32_t res = 0xFFFFFFFFu;
        bool done = false;
        for (int r = 0; r < 3 && !done; ++r) {
            const uint64_t* b = slots + static_cast<uint64_t>(HashBucketOf(dim, buckets, seeds[r], q[i])) * 16;
            bool full = true;
            for (int j = 0; j < 16; ++j) {
                if (b[j] == ~0ull) full = false;
                else if (static_cast<uint32_t>(b[j] >> 32) == q[i]) {
                    res = static_cast<uint32_t>(b[j]);
                    done = true;
                    break;
                }
            }
            if (!full) done = true;
        }
        if (!done)
            for (uint32_t j = 0; j < nstash; ++j)
                if (static_cast<uint32_t>(stash[j] >> 32) == q[i]) {
                    res = static_cast<uint32_t>(stash[j]);
                    break;
                }
        out[i] = res;
    }
}

}  // extern "C"
