// Shared definitions of the CK-MPM HIP engine (device and host).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ckmpm {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kDirections = 27;      // new block - old block in {-1,0,1}^3
constexpr int kMoveShift = 27;       // ordering word: slot (27 bits) | move << 27
constexpr uint32_t kSlotMask = (1u << kMoveShift) - 1u;
constexpr uint32_t kSetupRankClamp = (1u << 26) - 1u;  // within-cell rank in the setup sort key
constexpr int kNodesPerBlock = 128;  // 2 sub-grids x 4^3 nodes
constexpr int kMChannels = 4;        // m, px, py, pz
constexpr int kVChannels = 3;        // vx, vy, vz
constexpr int kMFloats = kMChannels * kNodesPerBlock;  // 512 floats = 2 KiB per block
constexpr int kVFloats = kVChannels * kNodesPerBlock;  // 384 floats = 1.5 KiB per block

enum Kind : int { kFixedCorotated = 0, kNeoHookean = 1, kDruckerPrager = 2, kFluid = 3, kVonMises = 4, kCamClay = 5 };

enum Flag : uint32_t {
    kFlagOutOfWindow = 1u << 0,
    kFlagOutOfDomain = 1u << 1,
    kFlagNonFinite = 1u << 2,
    kFlagCapacity = 1u << 3,
    kFlagMissingBlock = 1u << 4,
};

// Device-resident counters of one block layout.
struct LayoutCounts {
    uint32_t A;   // particle blocks: [0, A)
    uint32_t Nb;  // + neighbour blocks b + {0,1}^3: [A, Nb)  (the grid)
    uint32_t Ex;  // + exterior blocks b + {-1,0,1}^3: [Nb, Ex)
    uint32_t pad;
};

// Device-resident engine status.
constexpr int kMaxModels = 8;

// The per-substep maxima are double-buffered by substep parity (slot = substep & 1): the
// grid update of substep k reduces into slot k & 1 and clears slot (k + 1) & 1 for the
// next substep, so no separate reset launch is needed.  The G2P2G of substep k reads
// slot k & 1; the host reads the slot of the last grid update.
struct Status {
    uint32_t flags;
    uint32_t p2g_global;     // particles deposited through the FP32 global-atomic path
    uint32_t vmax2_bits[2];  // max |v|^2 of the grid update (non-negative float bits)
    // per model: largest P2G deposit bound of the previous substep (float bits), reduced
    // by the grid update from the per-block maxima; sets the fixed-point scale of P2G
    uint32_t dep_bits[2][kMaxModels];
};

struct DomainDev {
    int dim[3];
    int off[3];
    uint32_t nblocks;
    float dx, inv_dx;
};

struct MaterialDev {
    int kind;
    float lambda, mu;
    float bulk, gamma, viscosity, J0;
    int gamma_int;              // gamma when it is an integer in [1, 16], else 0
    float sin_phi;
    float yield;                // von Mises yield stress
    float beta, xi, msq, alpha0;  // Cam-clay: beta, hardening factor, M^2, initial logJp
    int hardening;              // Cam-clay hardening on
    float mass, volume;
};

// initial value of the material's scalar state attribute: J0 of a fluid, the initial
// logJp (alpha0) of a Cam-clay model (mpm_algorithm.cuh:1043-1049)
__host__ __device__ inline float InitAttr(const MaterialDev& m) { return m.kind == kCamClay ? m.alpha0 : m.J0; }

__host__ __device__ inline uint32_t KeyOf(const DomainDev& d, int bx, int by, int bz) {
    return static_cast<uint32_t>((bx * d.dim[1] + by) * d.dim[2] + bz);
}
__host__ __device__ inline void CoordsOf(const DomainDev& d, uint32_t key, int& bx, int& by, int& bz) {
    bz = static_cast<int>(key % static_cast<uint32_t>(d.dim[2]));
    const uint32_t r = key / static_cast<uint32_t>(d.dim[2]);
    by = static_cast<int>(r % static_cast<uint32_t>(d.dim[1]));
    bx = static_cast<int>(r / static_cast<uint32_t>(d.dim[1]));
}
__host__ __device__ inline bool InDomain(const DomainDev& d, int bx, int by, int bz) {
    return bx >= 0 && by >= 0 && bz >= 0 && bx < d.dim[0] && by < d.dim[1] && bz < d.dim[2];
}
__host__ __device__ inline int FloorDiv4(int v) { return v >> 2; }  // arithmetic shift = floor

// Grid-0 cell of an FP32 coordinate: int(x * invDx + 0.25f) with the multiply and
// the add rounded separately (mpm_algorithm.cuh:59-67).
__device__ inline int Cell0(float x, float inv_dx) {
    const float t = __fmul_rn(x, inv_dx);
    return static_cast<int>(floorf(__fadd_rn(t, 0.25f)));
}

__device__ inline void SetFlag(Status* st, uint32_t f) { atomicOr(&st->flags, f); }

}  // namespace ckmpm
