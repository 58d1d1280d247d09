// Reverse-mode (adjoint) substep for fixed-corotated and Neo-Hookean models, with
// substep tapes and checkpoints, and its C-ABI.
//
// Forward substep k (dt, next dt nd):  grid update  M_k -> V_k;  G2P at x_k (v, B, D,
// C = B D^-1);  x' = x + dt v;  F' = (I + dt C) F;  tau(F'; lambda_p, mu_p);  P2G at x'
// (D', K = m C - nd V (D'/2)^-1 tau) -> M_{k+1}.  The backward of substep k takes the
// adjoints of (x', F', lambda, mu) per particle and of M_{k+1} per node and returns
// those of (x, F, lambda, mu) and of M_k:
//   GridBackward      V_k-bar -> M_k-bar (through v = p/m, in the previous layout)
//   G2P2GBackward     per particle block: recomputes the forward particle math in
//                     registers, gathers M_{k+1}-bar at x' from an LDS halo,
//                     back-propagates P2G, stress (SVD polar derivative for FC), F
//                     update, x update and G2P, and scatters V_k-bar through an FP64
//                     LDS accumulator (the P2G pattern of the forward kernel).
//   RasterizeBackward M_0-bar -> initial velocity adjoint.
// The reference has no adjoint; the arithmetic differentiated is the forward kernel's
// (mpm_algorithm.cuh:1365-1845, mpm_material.cuh:318-357).
#include <algorithm>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "engine_impl.h"
#include "material.h"
#include "stencil.h"
#include "svd.h"

namespace {

constexpr int kAT = 256;         // threads per workgroup
constexpr int kAWin = 512;       // 8^3 nodes of one sub-grid window

uint32_t AGrid(uint64_t n) {
    uint64_t b = (n + kAT - 1) / kAT;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(b, 4096)));
}

// ------------------------------------------------------------------ small kernels
__global__ void FillZeroKernel(float* p, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0.0f;
}
// copy / clear the first Nb blocks (device count) of a per-block array
// Several device-to-device copies in one launch (a snapshot is a dozen buffers; one
// hipMemcpyAsync each costs a launch apiece).  blockIdx.y picks the segment; a segment
// with lc set copies the first min(lc->Nb, cap) blocks of block_bytes each.  Buffers
// are 256-byte aligned and sizes are multiples of 4 bytes.
constexpr int kMaxCopySegs = 6 + 1 + 4 * kMaxModels;
struct CopySeg {
    const void* src;
    void* dst;
    uint64_t bytes;
    const LayoutCounts* lc;
    uint32_t block_bytes;
};
struct CopyList {
    CopySeg seg[kMaxCopySegs];
    int n;
    uint32_t cap;
};
__global__ void MultiCopyKernel(CopyList L) {
    const CopySeg& c = L.seg[blockIdx.y];
    const uint64_t bytes = c.lc ? static_cast<uint64_t>(min(c.lc->Nb, L.cap)) * c.block_bytes : c.bytes;
    const uint64_t n16 = bytes >> 4;
    const uint4* s16 = static_cast<const uint4*>(c.src);
    uint4* d16 = static_cast<uint4*>(c.dst);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n16; i += stride) d16[i] = s16[i];
    // 4-byte tail (at most 3 words)
    const uint64_t w0 = n16 * 4, nw = bytes >> 2;
    const uint64_t i = w0 + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < nw) static_cast<uint32_t*>(c.dst)[i] = static_cast<const uint32_t*>(c.src)[i];
}
__global__ void ZeroBlocksKernel(float* dst, const LayoutCounts* lc, uint32_t cap, int fpb) {
    const size_t n = static_cast<size_t>(min(lc->Nb, cap)) * fpb;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = 0.0f;
}
// adjoint of positions: input order (AoS) -> processing order (SoA)
__global__ void SeedKernel(const uint32_t* ids, const float* xbar_in, uint32_t n, float* xbar) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t id = ids[t];
#pragma unroll
        for (int k = 0; k < 3; ++k) xbar[k * static_cast<size_t>(n) + t] = xbar_in[3 * static_cast<size_t>(id) + k];
    }
}
// SoA [k][n] in processing order -> AoS [id][k] in input order
__global__ void ToInputOrderKernel(const uint32_t* ids, const float* src, uint32_t n, int k_count, float* dst) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t id = ids[t];
        for (int k = 0; k < k_count; ++k)
            dst[static_cast<size_t>(k_count) * id + k] = src[k * static_cast<size_t>(n) + t];
    }
}

// dst[3 ids[t] + k] = src[3 t + k]: AoS rows from processing order to input order
__global__ void AosToInputOrderKernel(const uint32_t* ids, const float* src, uint32_t n, float* dst) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const size_t d = 3 * static_cast<size_t>(ids[t]), r = 3 * static_cast<size_t>(t);
        dst[d] = src[r];
        dst[d + 1] = src[r + 1];
        dst[d + 2] = src[r + 2];
    }
}

// ------------------------------------------------------------------ grid backward
struct GridBwdArgs {
    DomainDev dom;
    const uint32_t* keys;        // layout of substep k
    const LayoutCounts* lc;
    const uint32_t* table_prev;  // layout of M_k
    const LayoutCounts* lc_prev;
    const float* M;              // M_k (previous layout)
    float* Vbar;                 // V_k-bar (layout of substep k), cleared after use
    float* Mbar;                 // out: M_k-bar (previous layout), zeroed beforehand
    float mass_clamp;
    uint32_t cap;
};
__global__ __launch_bounds__(kAT) void GridBackwardKernel(GridBwdArgs a) {
    const uint32_t Nb = min(a.lc->Nb, a.cap), Nbp = a.lc_prev->Nb;
    const uint32_t n = threadIdx.x & 127u;
    for (uint32_t i = blockIdx.x * 2 + (threadIdx.x >> 7); i < Nb; i += gridDim.x * 2) {
        const uint32_t key = a.keys[i];
        // V-bar is read once, here: clear it for the next substep's accumulation
        float* vb = a.Vbar + static_cast<size_t>(i) * kVFloats;
        const float v[3] = {vb[n], vb[128 + n], vb[256 + n]};
        vb[n] = 0.0f;
        vb[128 + n] = 0.0f;
        vb[256 + n] = 0.0f;
        int b[3];
        CoordsOf(a.dom, key, b[0], b[1], b[2]);
        bool oob = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) oob |= (b[k] + a.dom.off[k] < 2) || (b[k] + a.dom.off[k] + 2 >= a.dom.dim[k]);
        const uint32_t io = a.table_prev[key];
        if (io == kEmpty || io >= Nbp) continue;
        const float* src = a.M + static_cast<size_t>(io) * kMFloats;
        const float m = src[n];
        if (oob || !(m > a.mass_clamp)) continue;
        const float p[3] = {src[128 + n], src[256 + n], src[384 + n]};
        const float im = 1.0f / m;
        float* dst = a.Mbar + static_cast<size_t>(io) * kMFloats;
        // v = p / m (+ g dt): p-bar = v-bar / m, m-bar = -(v-bar . p) / m^2
        dst[n] = -(v[0] * p[0] + v[1] * p[1] + v[2] * p[2]) * im * im;
        dst[128 + n] = v[0] * im;
        dst[256 + n] = v[1] * im;
        dst[384 + n] = v[2] * im;
    }
}

// ------------------------------------------------------------------ particle backward
struct AdjArgs {
    DomainDev dom;
    MaterialDev mat;
    const uint32_t* keys;    // layout of substep k
    const uint32_t* table;
    const LayoutCounts* lc;
    const float* V;          // V_k (recorded)
    const float* Mbar;       // M_{k+1}-bar, same layout
    float* Vbar;             // V_k-bar, accumulated (zero beforehand: the grid backward clears it)
    const float* st_in;      // particle state read by substep k (slot order)
    const uint32_t* perm;
    const uint32_t* off;
    uint32_t n;
    int pp_attr;
    const float* xb_in;      // adjoints of the substep's output, processing order
    const float* Eb_in;
    const float* tb_in;      // (lambda-bar, mu-bar)
    float* xb_out;           // adjoints of its input, slot order
    float* Eb_out;
    float* tb_out;
    float dt, next_dt;
    uint32_t split;          // workgroups per particle block (each takes every split-th chunk)
    float* Mzero;            // cleared first (the next grid backward's output), or null
    const LayoutCounts* lc_zero;  // its Nb blocks
    uint32_t cap;
};

__device__ __forceinline__ float W0p(float t) { return -1.0f + cospif(2.0f * t); }

// Gradient (w.r.t. the position, per axis) of the node-(ox,oy,oz) weight.
__device__ __forceinline__ void WeightGrad(const Stencil32& s, const float (&dw)[3], int ox, int oy, int oz,
                                           float inv_dx, float (&g)[3]) {
    const float W[3] = {ox ? 1.0f - s.w0[0] : s.w0[0], oy ? 1.0f - s.w0[1] : s.w0[1], oz ? 1.0f - s.w0[2] : s.w0[2]};
    const float dW[3] = {(ox ? -dw[0] : dw[0]) * inv_dx, (oy ? -dw[1] : dw[1]) * inv_dx, (oz ? -dw[2] : dw[2]) * inv_dx};
    g[0] = dW[0] * W[1] * W[2];
    g[1] = W[0] * dW[1] * W[2];
    g[2] = W[0] * W[1] * dW[2];
}

// Adjoint of the closed-form moment matrix sum_g D_g(t_g) * scale w.r.t. the
// position: D_aa = f(t_a), D_ab = m1(t_a) m1(t_b), f = w0 t^2 + (1-w0)(1-t)^2,
// m1 = 1 - w0 - t, d t / d x = 1 / dx.
__device__ __forceinline__ void MomentAdjoint(const Stencil32& s, const float (&dw)[3], const float (&Db)[9], float scale,
                                              float inv_dx, float (&xb)[3]) {
    float m1[3], m1p[3], fp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float t = s.t[a], w0 = s.w0[a];
        m1[a] = 1.0f - w0 - t;
        m1p[a] = -dw[a] - 1.0f;
        fp[a] = dw[a] * (2.0f * t - 1.0f) + 2.0f * (w0 + t - 1.0f);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float acc = Db[a * 4] * fp[a];
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (b != a) acc += (Db[a * 3 + b] + Db[b * 3 + a]) * m1p[a] * m1[b];
        xb[a] += scale * inv_dx * acc;
    }
}

__device__ __forceinline__ void Cofactor(const float (&F)[9], float (&c)[9]) {
    c[0] = F[4] * F[8] - F[5] * F[7];
    c[1] = F[5] * F[6] - F[3] * F[8];
    c[2] = F[3] * F[7] - F[4] * F[6];
    c[3] = F[2] * F[7] - F[1] * F[8];
    c[4] = F[0] * F[8] - F[2] * F[6];
    c[5] = F[1] * F[6] - F[0] * F[7];
    c[6] = F[1] * F[5] - F[2] * F[4];
    c[7] = F[2] * F[3] - F[0] * F[5];
    c[8] = F[0] * F[4] - F[1] * F[3];
}

// 2 waves per SIMD: 256 VGPRs (a few spills for FC) instead of ~275 and one wave
template <int KIND>
__global__ __launch_bounds__(kAT) __attribute__((amdgpu_waves_per_eu(2, 2))) void G2P2GBackwardKernel(AdjArgs a) {
    __shared__ float sV[kVChannels * 2 * kAWin];
    __shared__ float sMb[kMChannels * 2 * kAWin];
    __shared__ double sVb[kVChannels * 2 * kAWin];
    __shared__ uint32_t sHalo[8];
    const uint32_t A = a.lc->A, Nb = a.lc->Nb;
    const uint32_t tid = threadIdx.x;
    const float dx = a.dom.dx, inv_dx = a.dom.inv_dx;
    const size_t N = a.n;
    const float mass = a.mat.mass;
    if (a.Mzero) {
        const size_t nz = static_cast<size_t>(min(a.lc_zero->Nb, a.cap)) * kMFloats;
        for (size_t i = blockIdx.x * static_cast<size_t>(kAT) + tid; i < nz; i += static_cast<size_t>(gridDim.x) * kAT)
            a.Mzero[i] = 0.0f;
    }
    // workgroup w serves part w % split of blocks w / split, w / split + gridDim / split, ...
    const uint32_t part = blockIdx.x % a.split, per = gridDim.x / a.split;
    for (uint32_t j = blockIdx.x / a.split; j < A; j += per) {
        const uint32_t beg = a.off[j], end = a.off[j + 1];
        if (beg + part * kAT >= end) continue;  // no chunk for this part
        int bb[3];
        CoordsOf(a.dom, a.keys[j], bb[0], bb[1], bb[2]);
        if (tid < 8) {
            const int bx = bb[0] + ((tid >> 2) & 1), by = bb[1] + ((tid >> 1) & 1), bz = bb[2] + (tid & 1);
            uint32_t bi = kEmpty;
            if (InDomain(a.dom, bx, by, bz)) bi = a.table[KeyOf(a.dom, bx, by, bz)];
            if (bi != kEmpty && bi >= Nb) bi = kEmpty;
            sHalo[tid] = bi;
        }
        for (int i = tid; i < kVChannels * 2 * kAWin; i += kAT) sVb[i] = 0.0;
        __syncthreads();
        for (int i = tid; i < 8 * kMFloats; i += kAT) {
            const int o = i / kMFloats, r = i % kMFloats;
            const int ch = r >> 7, node = r & 127, g = node >> 6;
            const int lx = ((node >> 4) & 3) + 4 * ((o >> 2) & 1);
            const int ly = ((node >> 2) & 3) + 4 * ((o >> 1) & 1);
            const int lz = (node & 3) + 4 * (o & 1);
            const uint32_t bi = sHalo[o];
            const int l = (ch * 2 + g) * kAWin + lx * 64 + ly * 8 + lz;
            sMb[l] = bi != kEmpty ? a.Mbar[static_cast<size_t>(bi) * kMFloats + r] : 0.0f;
            if (ch < 3) sV[l] = bi != kEmpty ? a.V[static_cast<size_t>(bi) * kVFloats + r] : 0.0f;
        }
        __syncthreads();
        for (uint32_t c0 = beg + part * kAT; c0 < end; c0 += a.split * kAT) {
            const uint32_t t = c0 + tid;
            if (t < end) {
                const uint32_t pe = a.perm[t];
                uint32_t s = pe & kSlotMask;
                int mvv = static_cast<int>(pe >> kMoveShift);
                if (s >= a.n || mvv >= kDirections) {
                    s = 0;
                    mvv = 13;
                }
                const int mov[3] = {mvv / 9 - 1, (mvv / 3) % 3 - 1, mvv % 3 - 1};
                float x[3], E[9];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    x[k] = __fsub_rn(a.st_in[k * N + s], __fmul_rn(static_cast<float>(4 * mov[k]), dx));
#pragma unroll
                for (int k = 0; k < 9; ++k) E[k] = a.st_in[(3 + k) * N + s];
                float lam = a.mat.lambda, mu = a.mat.mu;
                if (a.pp_attr) {
                    lam = a.st_in[a.pp_attr * N + s];
                    mu = a.st_in[(a.pp_attr + 1) * N + s];
                }
                // ======================= forward recompute (as G2P2GKernel)
                float xs[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) xs[k] = __fmul_rn(x[k], inv_dx);
                Stencil32 sp[2];
                MakeStencilPair(xs, sp[0], sp[1]);
                int lw[2][3];
                float v[3] = {0.f, 0.f, 0.f}, B[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                float D[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    const Stencil32& sg = sp[g];
                    float m1[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        lw[g][k] = min(max(sg.c[k], 0), 6);
                        const float w0 = sg.w0[k], w1 = 1.0f - w0, tt = sg.t[k];
                        m1[k] = w1 - tt;
                        D[k * 4] += w0 * tt * tt + w1 * (1.0f - tt) * (1.0f - tt);
                    }
                    D[1] += m1[0] * m1[1];
                    D[2] += m1[0] * m1[2];
                    D[5] += m1[1] * m1[2];
                    const float* base = sV + g * kAWin + lw[g][0] * 64 + lw[g][1] * 8 + lw[g][2];
                    float vg[3] = {0.f, 0.f, 0.f}, S[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int o = 0; o < 8; ++o) {
                        const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
                        const float w = NodeWeight(sg, ox, oy, oz);
                        const int off = ox * 64 + oy * 8 + oz;
                        const float wv[3] = {w * base[off], w * base[2 * kAWin + off], w * base[4 * kAWin + off]};
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            vg[r] += wv[r];
                            if (ox) S[r * 3 + 0] += wv[r];
                            if (oy) S[r * 3 + 1] += wv[r];
                            if (oz) S[r * 3 + 2] += wv[r];
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        v[r] += vg[r];
#pragma unroll
                        for (int c = 0; c < 3; ++c) B[r * 3 + c] += dx * (S[r * 3 + c] - vg[r] * sg.t[c]);
                    }
                }
                const float dx2 = dx * dx;
                D[0] *= dx2;
                D[4] *= dx2;
                D[8] *= dx2;
                D[1] *= dx2;
                D[2] *= dx2;
                D[5] *= dx2;
                D[3] = D[1];
                D[6] = D[2];
                D[7] = D[5];
                float Di[9], C[9];
                InverseSym3(D, Di);
                MatMul3(B, Di, C);
                float xn[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    v[k] *= 0.5f;
                    xn[k] = x[k] + a.dt * v[k];
                }
                float CE[9], En[9];
                MatMul3(C, E, CE);
#pragma unroll
                for (int k = 0; k < 9; ++k) En[k] = E[k] + a.dt * (C[k] + CE[k]);
                float Fn[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) Fn[k] = En[k] + ((k % 4 == 0) ? 1.0f : 0.0f);
                float tau[9];
                Svd3 sv;
                float dsv[3] = {0.f, 0.f, 0.f}, jm1;
                if (KIND == kFixedCorotated) {
                    SignedSvdStrain(En, sv, dsv);
                    jm1 = dsv[0] + dsv[1] + dsv[2] + dsv[0] * dsv[1] + dsv[0] * dsv[2] + dsv[1] * dsv[2] +
                          dsv[0] * dsv[1] * dsv[2];
                    const float vol = lam * (1.0f + jm1) * jm1;
                    float kk[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) kk[q] = 2.0f * mu * sv.s[q] * dsv[q] + vol;
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int jj = 0; jj < 3; ++jj)
                            tau[i * 3 + jj] = sv.U[i * 3 + 0] * kk[0] * sv.U[jj * 3 + 0] +
                                              sv.U[i * 3 + 1] * kk[1] * sv.U[jj * 3 + 1] +
                                              sv.U[i * 3 + 2] * kk[2] * sv.U[jj * 3 + 2];
                } else {
                    jm1 = DetIPlusEMinus1(En);
                    StressNeoHookean(lam, mu, En, tau);
                }
                // P2G geometry at x'
#pragma unroll
                for (int k = 0; k < 3; ++k) xs[k] = __fmul_rn(xn[k], inv_dx);
                Stencil32 sn[2];
                MakeStencilPair(xs, sn[0], sn[1]);
                bool inwin = true;
                float Dn[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    float m1[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const int lc = sn[g].c[k];
                        inwin &= g == 0 ? (lc >= 1 && lc <= 6) : (lc >= 0 && lc <= 6);
                        const float w0 = sn[g].w0[k], w1 = 1.0f - w0, tt = sn[g].t[k];
                        m1[k] = w1 - tt;
                        Dn[k * 4] += w0 * tt * tt + w1 * (1.0f - tt) * (1.0f - tt);
                    }
                    Dn[1] += m1[0] * m1[1];
                    Dn[2] += m1[0] * m1[2];
                    Dn[5] += m1[1] * m1[2];
                }
                const float h = 0.5f * dx2;
                Dn[0] *= h;
                Dn[4] *= h;
                Dn[8] *= h;
                Dn[1] *= h;
                Dn[2] *= h;
                Dn[5] *= h;
                Dn[3] = Dn[1];
                Dn[6] = Dn[2];
                Dn[7] = Dn[5];
                float Dh[9], DT[9], K[9];
                InverseSym3(Dn, Dh);
                MatMul3(Dh, tau, DT);
                const float sv_ = a.next_dt * a.mat.volume;
#pragma unroll
                for (int k = 0; k < 9; ++k) K[k] = C[k] * mass - sv_ * DT[k];

                // ======================= backward
                float xbn[3], Fb[9];
#pragma unroll
                for (int k = 0; k < 3; ++k) xbn[k] = a.xb_in[k * N + t];
#pragma unroll
                for (int k = 0; k < 9; ++k) Fb[k] = a.Eb_in[k * N + t];
                float lamb = a.tb_in[t], mub = a.tb_in[N + t];
                float vb[3] = {0.f, 0.f, 0.f}, Cb[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (inwin) {
                    // ---- P2G: m_i += w m;  p_i += w (m v + K xx_i), xx_i = (o - t') dx
                    float Kb[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    const float mv3[3] = {mass * v[0], mass * v[1], mass * v[2]};
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        const Stencil32& sg = sn[g];
                        float dw[3];
#pragma unroll
                        for (int k = 0; k < 3; ++k) dw[k] = W0p(sg.t[k]);
                        const float* base = sMb + g * kAWin + sg.c[0] * 64 + sg.c[1] * 8 + sg.c[2];
#pragma unroll
                        for (int o = 0; o < 8; ++o) {
                            const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
                            const int off = ox * 64 + oy * 8 + oz;
                            const float w = NodeWeight(sg, ox, oy, oz);
                            float gw[3];
                            WeightGrad(sg, dw, ox, oy, oz, inv_dx, gw);
                            const float xx[3] = {(ox - sg.t[0]) * dx, (oy - sg.t[1]) * dx, (oz - sg.t[2]) * dx};
                            const float mb = base[off];
                            const float pb[3] = {base[2 * kAWin + off], base[4 * kAWin + off], base[6 * kAWin + off]};
                            float gi = mb * mass;
#pragma unroll
                            for (int r = 0; r < 3; ++r) {
                                const float q = mv3[r] + K[r * 3 + 0] * xx[0] + K[r * 3 + 1] * xx[1] + K[r * 3 + 2] * xx[2];
                                gi += pb[r] * q;
                                vb[r] += w * mass * pb[r];
#pragma unroll
                                for (int c = 0; c < 3; ++c) Kb[r * 3 + c] += w * pb[r] * xx[c];
                            }
#pragma unroll
                            for (int c = 0; c < 3; ++c)
                                xbn[c] += gi * gw[c] -
                                          w * (K[0 * 3 + c] * pb[0] + K[1 * 3 + c] * pb[1] + K[2 * 3 + c] * pb[2]);
                        }
                    }
                    // ---- K = m C - nd V Dh tau
                    float taub[9], Dhb[9], T1[9], Dnb[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) Cb[k] = mass * Kb[k];
                    MatMul3(Dh, Kb, taub);
#pragma unroll
                    for (int k = 0; k < 9; ++k) taub[k] *= -sv_;
                    // Dh-bar = -nd V Kb tau^T (tau symmetric);  Dn-bar = -Dh Dh-bar Dh
                    MatMul3(Kb, tau, Dhb);
                    MatMul3(Dh, Dhb, T1);
                    MatMul3(T1, Dh, Dnb);
#pragma unroll
                    for (int k = 0; k < 9; ++k) Dnb[k] *= sv_;
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        float dw[3];
#pragma unroll
                        for (int k = 0; k < 3; ++k) dw[k] = W0p(sn[g].t[k]);
                        MomentAdjoint(sn[g], dw, Dnb, h, inv_dx, xbn);
                    }
                    // ---- stress: tau(F'; lambda, mu)
                    const float J = 1.0f + jm1;
                    const float trb = taub[0] + taub[4] + taub[8];
                    float cof[9];
                    Cofactor(Fn, cof);
                    if (KIND == kFixedCorotated) {
                        // tau = 2 mu (F - R) F^T + lambda J (J - 1) I, F = U S V^T, R = U V^T
                        float T[9];  // U^T taub U
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj) {
                                float acc = 0.0f;
#pragma unroll
                                for (int p = 0; p < 3; ++p)
#pragma unroll
                                    for (int q = 0; q < 3; ++q) acc += sv.U[p * 3 + i] * taub[p * 3 + q] * sv.U[q * 3 + jj];
                                T[i * 3 + jj] = acc;
                            }
                        mub += 2.0f * (T[0] * dsv[0] * sv.s[0] + T[4] * dsv[1] * sv.s[1] + T[8] * dsv[2] * sv.s[2]);
                        lamb += J * jm1 * trb;
                        // Z_ij = -2 mu (T_ij s_j - T_ji s_i) / (s_i + s_j), i != j  (polar factor adjoint)
                        float Z[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj)
                                if (i != jj) {
                                    float den = sv.s[i] + sv.s[jj];
                                    if (fabsf(den) < 1e-6f) den = den < 0.0f ? -1e-6f : 1e-6f;
                                    Z[i * 3 + jj] = -2.0f * mu * (T[i * 3 + jj] * sv.s[jj] - T[jj * 3 + i] * sv.s[i]) / den;
                                }
                        // F-bar += 2 mu (taub F + taub^T (F - R)) + U Z V^T + lambda (2J - 1) tr(taub) cof(F)
                        float FmR[9], UZ[9];
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj) {
                                FmR[i * 3 + jj] = sv.U[i * 3 + 0] * dsv[0] * sv.V[jj * 3 + 0] +
                                                  sv.U[i * 3 + 1] * dsv[1] * sv.V[jj * 3 + 1] +
                                                  sv.U[i * 3 + 2] * dsv[2] * sv.V[jj * 3 + 2];
                                UZ[i * 3 + jj] = sv.U[i * 3 + 0] * Z[0 * 3 + jj] + sv.U[i * 3 + 1] * Z[1 * 3 + jj] +
                                                 sv.U[i * 3 + 2] * Z[2 * 3 + jj];
                            }
                        const float vol = lam * (2.0f * J - 1.0f) * trb;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj) {
                                float acc = 0.0f;
#pragma unroll
                                for (int p = 0; p < 3; ++p)
                                    acc += taub[i * 3 + p] * Fn[p * 3 + jj] + taub[p * 3 + i] * FmR[p * 3 + jj];
                                const float uzv = UZ[i * 3 + 0] * sv.V[jj * 3 + 0] + UZ[i * 3 + 1] * sv.V[jj * 3 + 1] +
                                                  UZ[i * 3 + 2] * sv.V[jj * 3 + 2];
                                Fb[i * 3 + jj] += 2.0f * mu * acc + uzv + vol * cof[i * 3 + jj];
                            }
                    } else {
                        // tau = mu (F F^T - I) + lambda ln J I
                        const float lnJ = log1pf(jm1);
                        lamb += lnJ * trb;
                        float acc_mu = 0.0f;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj) {
                                const float bb_ = En[i * 3 + jj] + En[jj * 3 + i] + En[i * 3 + 0] * En[jj * 3 + 0] +
                                                  En[i * 3 + 1] * En[jj * 3 + 1] + En[i * 3 + 2] * En[jj * 3 + 2];
                                acc_mu += bb_ * taub[i * 3 + jj];
                            }
                        mub += acc_mu;
                        const float iJ = 1.0f / J;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj) {
                                float acc = 0.0f;
#pragma unroll
                                for (int p = 0; p < 3; ++p) acc += (taub[i * 3 + p] + taub[p * 3 + i]) * Fn[p * 3 + jj];
                                Fb[i * 3 + jj] += mu * acc + lam * trb * cof[i * 3 + jj] * iJ;
                            }
                    }
                }
                // ---- F' = (I + dt C) F:  C-bar += dt F'-bar F^T;  F-bar = (I + dt C)^T F'-bar
                float Fo[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) Fo[k] = E[k] + ((k % 4 == 0) ? 1.0f : 0.0f);
                float Fbo[9];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) {
                        Cb[i * 3 + jj] += a.dt * (Fb[i * 3 + 0] * Fo[jj * 3 + 0] + Fb[i * 3 + 1] * Fo[jj * 3 + 1] +
                                                  Fb[i * 3 + 2] * Fo[jj * 3 + 2]);
                        Fbo[i * 3 + jj] = Fb[i * 3 + jj] + a.dt * (C[0 * 3 + i] * Fb[0 * 3 + jj] + C[1 * 3 + i] * Fb[1 * 3 + jj] +
                                                                   C[2 * 3 + i] * Fb[2 * 3 + jj]);
                    }
                // ---- x' = x + dt v
                float xb[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    xb[k] = xbn[k];
                    vb[k] += a.dt * xbn[k];
                }
                // ---- C = B D^-1:  B-bar = C-bar D^-1;  D-bar = -C^T B-bar
                float Bb[9], Db[9];
                MatMul3(Cb, Di, Bb);
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj)
                        Db[i * 3 + jj] = -(C[0 * 3 + i] * Bb[0 * 3 + jj] + C[1 * 3 + i] * Bb[1 * 3 + jj] +
                                           C[2 * 3 + i] * Bb[2 * 3 + jj]);
                // ---- G2P: v = 1/2 sum w v_i;  B = sum v_i (w xx_i)^T;  D = sum w xx xx^T
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    const Stencil32& sg = sp[g];
                    float dw[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) dw[k] = W0p(sg.t[k]);
                    MomentAdjoint(sg, dw, Db, dx2, inv_dx, xb);
                    const int lbase = g * kAWin + lw[g][0] * 64 + lw[g][1] * 8 + lw[g][2];
#pragma unroll
                    for (int o = 0; o < 8; ++o) {
                        const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
                        const int off = lbase + ox * 64 + oy * 8 + oz;
                        const float w = NodeWeight(sg, ox, oy, oz);
                        float gw[3];
                        WeightGrad(sg, dw, ox, oy, oz, inv_dx, gw);
                        const float xx[3] = {(ox - sg.t[0]) * dx, (oy - sg.t[1]) * dx, (oz - sg.t[2]) * dx};
                        const float vi[3] = {sV[off], sV[2 * kAWin + off], sV[4 * kAWin + off]};
                        float bx[3];
#pragma unroll
                        for (int r = 0; r < 3; ++r) bx[r] = Bb[r * 3 + 0] * xx[0] + Bb[r * 3 + 1] * xx[1] + Bb[r * 3 + 2] * xx[2];
                        const float ai = 0.5f * (vi[0] * vb[0] + vi[1] * vb[1] + vi[2] * vb[2]) +
                                         vi[0] * bx[0] + vi[1] * bx[1] + vi[2] * bx[2];
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            xb[c] += ai * gw[c] - w * (Bb[0 * 3 + c] * vi[0] + Bb[1 * 3 + c] * vi[1] + Bb[2 * 3 + c] * vi[2]);
#pragma unroll
                        for (int r = 0; r < 3; ++r)
                            atomicAdd(&sVb[2 * r * kAWin + off], static_cast<double>(w * (0.5f * vb[r] + bx[r])));
                    }
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) a.xb_out[k * N + s] = xb[k];
#pragma unroll
                for (int k = 0; k < 9; ++k) a.Eb_out[k * N + s] = Fbo[k];
                a.tb_out[s] = lamb;
                a.tb_out[N + s] = mub;
            }
        }
        __syncthreads();
        for (int i = tid; i < kVChannels * 2 * kAWin; i += kAT) {
            const double val = sVb[i];
            if (val == 0.0) continue;
            const int ch = i / (2 * kAWin), g = (i / kAWin) & 1, l = i & (kAWin - 1);
            const int lx = l >> 6, ly = (l >> 3) & 7, lz = l & 7;
            const int o = ((lx >> 2) << 2) | ((ly >> 2) << 1) | (lz >> 2);
            const uint32_t bi = sHalo[o];
            if (bi == kEmpty) continue;
            const int node = g * 64 + (lx & 3) * 16 + (ly & 3) * 4 + (lz & 3);
            unsafeAtomicAdd(a.Vbar + static_cast<size_t>(bi) * kVFloats + ch * 128 + node, static_cast<float>(val));
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ rasterize backward
__global__ void RasterizeBackwardKernel(DomainDev d, const float* xyz, uint32_t n, float mass, const uint32_t* table,
                                        const LayoutCounts* lc, const float* Mbar, float* vbar) {
    const uint32_t Nb = lc->Nb;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        float xs[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) xs[k] = __fmul_rn(xyz[3 * static_cast<size_t>(p) + k], d.inv_dx);
        Stencil32 sp[2];
        MakeStencilPair(xs, sp[0], sp[1]);
        float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
                const int cx = sp[g].c[0] + ox, cy = sp[g].c[1] + oy, cz = sp[g].c[2] + oz;
                const int bx = FloorDiv4(cx), by = FloorDiv4(cy), bz = FloorDiv4(cz);
                if (!InDomain(d, bx, by, bz)) continue;
                const uint32_t bi = table[KeyOf(d, bx, by, bz)];
                if (bi == kEmpty || bi >= Nb) continue;
                const int node = g * 64 + (cx - 4 * bx) * 16 + (cy - 4 * by) * 4 + (cz - 4 * bz);
                const float wm = NodeWeight(sp[g], ox, oy, oz) * mass;
                const float* src = Mbar + static_cast<size_t>(bi) * kMFloats;
                acc[0] += wm * src[128 + node];
                acc[1] += wm * src[256 + node];
                acc[2] += wm * src[384 + node];
            }
#pragma unroll
        for (int k = 0; k < 3; ++k) vbar[3 * static_cast<size_t>(p) + k] = acc[k];
    }
}

}  // namespace

// ================================================================== host side
struct Snapshot {
    // inputs of one substep: current and previous layout, grid mass/momentum, per-model
    // ordering, state and particle identities, and the engine's buffer roles
    uint32_t* keys[2] = {nullptr, nullptr};   // [0] current, [1] previous layout
    uint32_t* table[2] = {nullptr, nullptr};
    LayoutCounts* lc[2] = {nullptr, nullptr};
    float* M = nullptr;
    float* V = nullptr;  // grid velocities of the substep (current layout)
    std::vector<uint32_t*> perm, off, ids;
    std::vector<float*> st;
    double dt = 0, nd = 0;
    int cur = 0, sc = 0, mprev = 0, idc = 0;
    uint64_t substeps = 0;
    uint32_t a_hint = 0;
    double time = 0, cur_dt = 0;
};

struct AdjointState {
    uint32_t seg = 0;
    bool recording = false;
    // Two substep tapes.  ckmpm_adjoint_substep runs on its own stream (ustream) and
    // undoes tape und while the replay of the previous segment records into the other
    // tape on the engine's stream: at a few hundred particle blocks both are launch- and
    // latency-bound, so they overlap.  A tape is allocated when first recorded into.
    std::vector<Snapshot> tape[2];
    uint32_t tlen[2] = {0, 0};
    int rec = 0;                     // tape being recorded into
    int und = 0;                     // tape being undone
    bool undone[2] = {false, false};  // free_ev[t] follows the last undo queued from tape t
    hipStream_t ustream = nullptr;
    hipEvent_t join_ev = nullptr;    // engine stream -> ustream
    hipEvent_t back_ev = nullptr;    // ustream -> engine stream
    hipEvent_t free_ev[2] = {nullptr, nullptr};
    bool need_join = true;           // the engine stream holds work the next undo reads
    std::vector<uint32_t*> ids[2];
    int idc = 0;
    std::vector<uint32_t*> setup_order;
    std::vector<float*> x0;
    std::vector<float*> xb[2], Eb[2], tb[2];
    int ab = 0;
    float* Mbar[2] = {nullptr, nullptr};
    int mb = 0;
    float* Vbar = nullptr;
    float* host_io = nullptr;  // device staging of 3 n floats
    const uint32_t* setup_table = nullptr;    // layout of M_0 (set when the first substep is undone)
    const LayoutCounts* setup_lc = nullptr;
    bool at_setup = false;
    std::vector<void*> allocs;
    // state right after the initial setup (ckmpm_adjoint_rewind)
    Snapshot setup;
    bool has_setup = false;
    // checkpoint memory: blocks of ck_bytes, reused after ckmpm_checkpoint_free
    size_t ck_bytes = 0;
    std::vector<void*> ck_pool;
    std::set<ckmpm_checkpoint*> live;
};

struct ckmpm_checkpoint {
    Snapshot snap;
    void* mem = nullptr;
    ckmpm_engine* eng = nullptr;  // null once the engine is destroyed (memory freed with it)
};

namespace {

template <typename T>
bool DevAlloc(std::vector<void*>& list, T** p, size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return false;
    list.push_back(q);
    *p = static_cast<T*>(q);
    return true;
}

// Carves snapshots out of one allocation (a tape of hundreds of snapshots would
// otherwise cost thousands of hipMalloc / hipFree calls).
struct Carver {
    char* base = nullptr;
    size_t used = 0;
    template <typename T>
    T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
        used += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        return p;
    }
};

// with_v: room for the grid velocities (substep tape; checkpoints do not keep them)
void CarveSnapshot(ckmpm_engine* e, Snapshot& s, Carver& c, bool with_v) {
    for (int l = 0; l < 2; ++l) {
        s.keys[l] = c.take<uint32_t>(e->cap);
        s.table[l] = c.take<uint32_t>(e->dom.nblocks);
        s.lc[l] = c.take<LayoutCounts>(1);
    }
    s.M = c.take<float>(static_cast<size_t>(e->cap) * kMFloats);
    s.V = with_v ? c.take<float>(static_cast<size_t>(e->cap) * kVFloats) : nullptr;
    const size_t nm = e->models.size();
    s.perm.assign(nm, nullptr);
    s.off.assign(nm, nullptr);
    s.ids.assign(nm, nullptr);
    s.st.assign(nm, nullptr);
    for (size_t i = 0; i < nm; ++i) {
        const ModelBuf& m = e->models[i];
        s.perm[i] = c.take<uint32_t>(m.n);
        s.off[i] = c.take<uint32_t>(static_cast<size_t>(e->cap) + 1);
        s.ids[i] = c.take<uint32_t>(m.n);
        s.st[i] = c.take<float>(static_cast<size_t>(m.n) * m.nattr);
    }
}

// count snapshots in one device allocation (recorded in list)
bool AllocSnapshots(ckmpm_engine* e, Snapshot* s, size_t count, std::vector<void*>& list, bool with_v) {
    Carver sizer;
    Snapshot probe;
    CarveSnapshot(e, probe, sizer, with_v);
    const size_t per = sizer.used;
    char* base = nullptr;
    if (!DevAlloc(list, &base, per * count)) return false;
    for (size_t k = 0; k < count; ++k) {
        Carver c{base + k * per, 0};
        CarveSnapshot(e, s[k], c, with_v);
    }
    return true;
}

// The engine buffers of a snapshot, paired with the snapshot's copies: to_snap copies
// engine -> snapshot (Take), otherwise snapshot -> engine (Put).  The engine's buffer
// roles (cur, sc, mprev, idc) must already be those of the snapshot.
void LaunchSnapshotCopy(ckmpm_engine* e, const Snapshot& s, bool to_snap) {
    AdjointState* A = e->adj;
    CopyList cl{};
    cl.cap = e->cap;
    auto add = [&](void* eng, void* snap, uint64_t bytes, const LayoutCounts* lc = nullptr, uint32_t bb = 0) {
        CopySeg& c = cl.seg[cl.n++];
        c.src = to_snap ? eng : snap;
        c.dst = to_snap ? snap : eng;
        c.bytes = bytes;
        c.lc = lc;
        c.block_bytes = bb;
    };
    const LayoutBuf* L[2] = {&e->L[e->cur], &e->L[1 - e->cur]};
    for (int l = 0; l < 2; ++l) {
        add(L[l]->keys, s.keys[l], sizeof(uint32_t) * e->cap);
        add(L[l]->table, s.table[l], sizeof(uint32_t) * e->dom.nblocks);
        add(L[l]->lc, s.lc[l], sizeof(LayoutCounts));
    }
    // M_k is indexed by the previous layout: its Nb blocks are the ones in use (the count
    // is read from the side being copied, before the copy of that count lands)
    add(e->M[e->mprev], s.M, 0, to_snap ? L[1]->lc : s.lc[1], kMFloats * sizeof(float));
    // V of the current layout: recorded for the backward, not restored (the next grid
    // phase recomputes it)
    if (to_snap && s.V) add(e->V, s.V, 0, L[0]->lc, kVFloats * sizeof(float));
    for (size_t i = 0; i < e->models.size(); ++i) {
        const ModelBuf& m = e->models[i];
        add(m.perm[e->sc], s.perm[i], sizeof(uint32_t) * m.n);
        add(m.off[e->sc], s.off[i], sizeof(uint32_t) * (static_cast<uint64_t>(e->cap) + 1));
        add(A->ids[A->idc][i], s.ids[i], sizeof(uint32_t) * m.n);
        add(m.st[e->sc], s.st[i], sizeof(float) * static_cast<uint64_t>(m.n) * m.nattr);
    }
    // enough workgroups for about four 16-byte loads per thread in the largest segment
    uint64_t most = 0;
    for (int k = 0; k < cl.n; ++k)
        most = std::max<uint64_t>(most, cl.seg[k].lc ? static_cast<uint64_t>(std::min<uint32_t>(e->cap, 2 * e->a_hint + 64)) *
                                                           cl.seg[k].block_bytes
                                                     : cl.seg[k].bytes);
    const uint32_t gx = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(most / (kAT * 64), 16), 1024));
    MultiCopyKernel<<<dim3(gx, cl.n), kAT, 0, e->stream>>>(cl);
}

int Take(ckmpm_engine* e, Snapshot& s) {
    AdjointState* A = e->adj;
    LaunchSnapshotCopy(e, s, true);
    CK_HIP(e, hipGetLastError());
    s.cur = e->cur;
    s.sc = e->sc;
    s.mprev = e->mprev;
    s.idc = A->idc;
    s.substeps = e->substeps;
    s.a_hint = e->a_hint;
    s.time = e->time;
    s.cur_dt = e->cur_dt;
    return CKMPM_OK;
}

int Put(ckmpm_engine* e, const Snapshot& s) {
    hipStream_t st = e->stream;
    AdjointState* A = e->adj;
    e->cur = s.cur;
    e->sc = s.sc;
    e->mprev = s.mprev;
    A->idc = s.idc;
    LaunchSnapshotCopy(e, s, false);
    CK_HIP(e, hipGetLastError());
    CK_HIP(e, hipMemsetAsync(e->status, 0, sizeof(Status), st));
    e->substeps = s.substeps;
    e->a_hint = s.a_hint;
    e->time = s.time;
    e->cur_dt = s.cur_dt;
    return CKMPM_OK;
}

// The engine stream waits for every undo queued so far.
void JoinBack(ckmpm_engine* e) {
    AdjointState* A = e->adj;
    (void)hipEventRecord(A->back_ev, A->ustream);
    (void)hipStreamWaitEvent(e->stream, A->back_ev, 0);
}

uint32_t GridSize(const ckmpm_engine* e) {
    uint32_t g = std::max<uint32_t>(e->a_hint + e->a_hint / 4 + 8, 8u);
    g = std::min<uint32_t>(g, 16384u);
    return (g + 7u) & ~7u;
}

}  // namespace

// ------------------------------------------------------------------ engine hooks
namespace ckmpm {

void AdjointOnSetupOrder(ckmpm_engine* e, size_t model, const uint32_t* order, const float* xyz) {
    AdjointState* A = e->adj;
    if (!A) return;
    const ModelBuf& m = e->models[model];
    (void)hipMemcpyAsync(A->setup_order[model], order, sizeof(uint32_t) * m.n, hipMemcpyDeviceToDevice, e->stream);
    (void)hipMemcpyAsync(A->ids[A->idc][model], order, sizeof(uint32_t) * m.n, hipMemcpyDeviceToDevice, e->stream);
    (void)hipMemcpyAsync(A->x0[model], xyz, sizeof(float) * 3 * m.n, hipMemcpyDeviceToDevice, e->stream);
}

void AdjointAfterSetup(ckmpm_engine* e) {
    AdjointState* A = e->adj;
    if (!A) return;
    Carver sizer;
    Snapshot probe;
    CarveSnapshot(e, probe, sizer, false);
    char* base = nullptr;
    if (!DevAlloc(A->allocs, &base, sizer.used)) return;  // rewind then reports "not set up"
    Carver c{base, 0};
    CarveSnapshot(e, A->setup, c, false);
    A->has_setup = Take(e, A->setup) == CKMPM_OK;
}

int AdjointCheckTape(ckmpm_engine* e) {
    AdjointState* A = e->adj;
    if (A && A->recording && A->tlen[A->rec] >= A->seg) return Fail(e, CKMPM_ERR_CAPACITY, "substep tape full");
    return CKMPM_OK;
}

int AdjointBeforeSubstep(ckmpm_engine* e, double dt, double nd) {
    AdjointState* A = e->adj;
    if (!A || !A->recording) return CKMPM_OK;
    uint32_t& len = A->tlen[A->rec];
    if (len >= A->seg) return Fail(e, CKMPM_ERR_CAPACITY, "substep tape full");
    Snapshot& s = A->tape[A->rec][len];
    CK_TRY(Take(e, s));
    s.dt = dt;
    s.nd = nd;
    ++len;
    A->need_join = true;
    return CKMPM_OK;
}

void AdjointIdBuffers(ckmpm_engine* e, size_t model, const uint32_t** in, uint32_t** out) {
    AdjointState* A = e->adj;
    *in = A ? A->ids[A->idc][model] : nullptr;
    *out = A ? A->ids[1 - A->idc][model] : nullptr;
}

void AdjointAfterSubstep(ckmpm_engine* e) {
    AdjointState* A = e->adj;
    if (!A) return;
    // G2P2G wrote ids[1 - idc][t] = ids[idc][slot of t]: the state written by the
    // substep holds, at position t, the particle of that slot
    A->idc = 1 - A->idc;
}

void AdjointDestroy(ckmpm_engine* e) {
    AdjointState* A = e->adj;
    if (!A) return;
    for (ckmpm_checkpoint* c : A->live) {
        (void)hipFree(c->mem);
        c->mem = nullptr;
        c->eng = nullptr;
    }
    if (A->ustream) (void)hipStreamSynchronize(A->ustream);
    for (void* p : A->ck_pool) (void)hipFree(p);
    for (void* p : A->allocs) (void)hipFree(p);
    for (hipEvent_t ev : {A->join_ev, A->back_ev, A->free_ev[0], A->free_ev[1]})
        if (ev) (void)hipEventDestroy(ev);
    if (A->ustream) (void)hipStreamDestroy(A->ustream);
    delete A;
    e->adj = nullptr;
}

}  // namespace ckmpm

// ------------------------------------------------------------------ C-ABI
extern "C" {

int ckmpm_adjoint_enable(ckmpm_engine* e, uint32_t segment) {
    if (!e || segment == 0) return CKMPM_ERR_INVALID;
    if (e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "enable the adjoint before ckmpm_initial_setup");
    if (e->adj) return Fail(e, CKMPM_ERR_INVALID, "adjoint already enabled");
    if (e->models.empty()) return Fail(e, CKMPM_ERR_INVALID, "add the models first");
    if (e->bc_custom) return Fail(e, CKMPM_ERR_INVALID, "the adjoint differentiates the default boundary only");
    for (const ModelBuf& m : e->models) {
        if (m.kind != kFixedCorotated && m.kind != kNeoHookean)
            return Fail(e, CKMPM_ERR_INVALID, "the adjoint supports fixed-corotated and Neo-Hookean models");
        if (!m.pp_attr) return Fail(e, CKMPM_ERR_INVALID, "the adjoint needs per-particle lambda and mu");
        if (m.rigid_attr) return Fail(e, CKMPM_ERR_INVALID, "the adjoint does not differentiate rigid particles");
    }
    AdjointState* A = new AdjointState();
    A->seg = segment;
    bool ok = CreateOwnQueueStream(&A->ustream) == hipSuccess;
    for (hipEvent_t* ev : {&A->join_ev, &A->back_ev, &A->free_ev[0], &A->free_ev[1]})
        ok = ok && hipEventCreateWithFlags(ev, hipEventDisableTiming) == hipSuccess;
    const size_t nm = e->models.size();
    for (int k = 0; k < 2; ++k) {
        A->ids[k].assign(nm, nullptr);
        A->xb[k].assign(nm, nullptr);
        A->Eb[k].assign(nm, nullptr);
        A->tb[k].assign(nm, nullptr);
    }
    A->setup_order.assign(nm, nullptr);
    A->x0.assign(nm, nullptr);
    uint32_t nmax = 0;
    for (size_t i = 0; ok && i < nm; ++i) {
        const uint32_t n = e->models[i].n;
        nmax = std::max(nmax, n);
        for (int k = 0; k < 2; ++k) {
            ok = ok && DevAlloc(A->allocs, &A->ids[k][i], n);
            ok = ok && DevAlloc(A->allocs, &A->xb[k][i], 3 * static_cast<size_t>(n));
            ok = ok && DevAlloc(A->allocs, &A->Eb[k][i], 9 * static_cast<size_t>(n));
            ok = ok && DevAlloc(A->allocs, &A->tb[k][i], 2 * static_cast<size_t>(n));
        }
        ok = ok && DevAlloc(A->allocs, &A->setup_order[i], n);
        ok = ok && DevAlloc(A->allocs, &A->x0[i], 3 * static_cast<size_t>(n));
    }
    for (int k = 0; k < 2; ++k) ok = ok && DevAlloc(A->allocs, &A->Mbar[k], static_cast<size_t>(e->cap) * kMFloats);
    ok = ok && DevAlloc(A->allocs, &A->Vbar, static_cast<size_t>(e->cap) * kVFloats);
    ok = ok && DevAlloc(A->allocs, &A->host_io, 3 * static_cast<size_t>(nmax));
    e->adj = A;
    if (!ok) {
        AdjointDestroy(e);
        return Fail(e, CKMPM_ERR_HIP, "adjoint allocation failed");
    }
    return CKMPM_OK;
}

int ckmpm_checkpoint_save(ckmpm_engine* e, ckmpm_checkpoint** out) {
    if (!e || !out) return CKMPM_ERR_INVALID;
    if (!e->adj || !e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "checkpoints need the adjoint enabled and set up");
    AdjointState* A = e->adj;
    if (A->ck_bytes == 0) {
        Carver sizer;
        Snapshot probe;
        CarveSnapshot(e, probe, sizer, false);
        A->ck_bytes = sizer.used;
    }
    void* mem = nullptr;
    if (!A->ck_pool.empty()) {
        mem = A->ck_pool.back();
        A->ck_pool.pop_back();
    } else if (hipMalloc(&mem, A->ck_bytes) != hipSuccess) {
        return Fail(e, CKMPM_ERR_HIP, "checkpoint allocation failed");
    }
    ckmpm_checkpoint* c = new ckmpm_checkpoint();
    c->mem = mem;
    c->eng = e;
    Carver carve{static_cast<char*>(mem), 0};
    CarveSnapshot(e, c->snap, carve, false);
    A->live.insert(c);
    const int rc = Take(e, c->snap);
    if (rc != CKMPM_OK) {
        ckmpm_checkpoint_free(c);
        return rc;
    }
    *out = c;
    return CKMPM_OK;
}

int ckmpm_checkpoint_restore(ckmpm_engine* e, const ckmpm_checkpoint* c) {
    if (!e || !c || !e->adj) return CKMPM_ERR_INVALID;
    // a checkpoint holds its own engine's sizes and memory: another engine's, or one whose
    // engine is gone, would copy the wrong extents
    if (c->eng != e || !c->mem || !e->adj->live.count(const_cast<ckmpm_checkpoint*>(c)))
        return Fail(e, CKMPM_ERR_INVALID, "checkpoint belongs to another engine or was released");
    return Put(e, c->snap);
}

int ckmpm_checkpoint_free(ckmpm_checkpoint* c) {
    if (!c) return CKMPM_ERR_INVALID;
    if (c->eng && c->eng->adj) {
        // the memory goes back to the engine's pool; a pending copy into it is
        // stream-ordered before any later checkpoint's use
        AdjointState* A = c->eng->adj;
        A->live.erase(c);
        A->ck_pool.push_back(c->mem);
    }
    delete c;
    return CKMPM_OK;
}

int ckmpm_adjoint_rewind(ckmpm_engine* e, const float* vel, uint64_t nvel, const float* lambda_p, const float* mu_p) {
    if (!e || !vel || !lambda_p || !mu_p) return CKMPM_ERR_INVALID;
    AdjointState* A = e->adj;
    if (!A || !A->has_setup) return Fail(e, CKMPM_ERR_INVALID, "rewind needs the adjoint enabled and set up");
    if (e->models.size() != 1) return Fail(e, CKMPM_ERR_INVALID, "rewind supports engines with one model");
    const ModelBuf& m = e->models[0];
    if (nvel != 1 && nvel != m.n) return Fail(e, CKMPM_ERR_INVALID, "vel holds 1 or n velocities");
    hipStream_t st = e->stream;
    CK_TRY(Put(e, A->setup));
    // initial state with the new Lame parameters (setup order, as the setup wrote it)
    CK_HIP(e, hipMemcpyAsync(A->host_io, lambda_p, sizeof(float) * m.n, hipMemcpyHostToDevice, st));
    CK_HIP(e, hipMemcpyAsync(A->host_io + m.n, mu_p, sizeof(float) * m.n, hipMemcpyHostToDevice, st));
    // (the kernel turns its order argument into the identity ordering: give it a copy in
    // the move-code buffer, which the next G2P2G overwrites)
    CK_HIP(e, hipMemcpyAsync(m.code, A->setup_order[0], sizeof(uint32_t) * m.n, hipMemcpyDeviceToDevice, st));
    LaunchInitState(e->dom, m.kind, A->x0[0], A->host_io, A->host_io + m.n, m.pp_attr, m.code, m.n, InitAttr(m.mat),
                    m.st[e->sc], st);
    // grid mass / momentum of the new velocities on the setup layout
    const LayoutBuf& L = e->L[e->cur];
    CK_HIP(e, hipMemsetAsync(e->M[e->mprev], 0, static_cast<size_t>(e->cap) * kMFloats * sizeof(float), st));
    CK_HIP(e, hipMemcpyAsync(A->host_io, vel, sizeof(float) * 3 * nvel, hipMemcpyHostToDevice, st));
    LaunchRasterize(e->dom, A->x0[0], A->host_io, m.n, static_cast<uint32_t>(nvel), m.mat.mass, L.table, L.lc,
                    e->M[e->mprev], e->status, st);
    CK_HIP(e, hipMemsetAsync(m.blkmax, 0, static_cast<size_t>(e->cap) * sizeof(float), st));
    A->tlen[0] = A->tlen[1] = 0;
    A->recording = false;
    A->at_setup = false;
    CK_HIP(e, hipStreamSynchronize(st));
    return CKMPM_OK;
}

int ckmpm_adjoint_record(ckmpm_engine* e, int32_t on) {
    if (!e || !e->adj) return CKMPM_ERR_INVALID;
    AdjointState* A = e->adj;
    if (on && !A->recording) {
        // the tape that is not being undone; its last undos must be done before it is
        // overwritten
        const int t = 1 - A->und;
        if (A->tape[t].empty()) {
            A->tape[t].resize(A->seg);
            if (!AllocSnapshots(e, A->tape[t].data(), A->seg, A->allocs, true)) {
                A->tape[t].clear();
                return Fail(e, CKMPM_ERR_HIP, "substep tape allocation failed");
            }
        }
        if (A->undone[t]) CK_HIP(e, hipStreamWaitEvent(e->stream, A->free_ev[t], 0));
        A->rec = t;
    } else if (!on && A->recording) {
        A->und = A->rec;  // undone next
    }
    if (on) A->tlen[A->rec] = 0;
    A->recording = on != 0;
    return CKMPM_OK;
}

int ckmpm_adjoint_reset(ckmpm_engine* e) {
    if (!e || !e->adj) return CKMPM_ERR_INVALID;
    AdjointState* A = e->adj;
    JoinBack(e);  // undos still queued write the buffers cleared here
    A->need_join = true;
    for (size_t i = 0; i < e->models.size(); ++i) {
        const uint32_t n = e->models[i].n;
        FillZeroKernel<<<AGrid(3ull * n), kAT, 0, e->stream>>>(A->xb[A->ab][i], 3ull * n);
        FillZeroKernel<<<AGrid(9ull * n), kAT, 0, e->stream>>>(A->Eb[A->ab][i], 9ull * n);
        FillZeroKernel<<<AGrid(2ull * n), kAT, 0, e->stream>>>(A->tb[A->ab][i], 2ull * n);
    }
    FillZeroKernel<<<AGrid(static_cast<size_t>(e->cap) * kMFloats), kAT, 0, e->stream>>>(
        A->Mbar[A->mb], static_cast<size_t>(e->cap) * kMFloats);
    // V-bar accumulates from zero; each grid backward clears what it read
    FillZeroKernel<<<AGrid(static_cast<size_t>(e->cap) * kVFloats), kAT, 0, e->stream>>>(
        A->Vbar, static_cast<size_t>(e->cap) * kVFloats);
    A->at_setup = false;
    CK_HIP(e, hipGetLastError());
    return CKMPM_OK;
}

int ckmpm_adjoint_seed(ckmpm_engine* e, int32_t model, const float* xbar) {
    if (!e || !e->adj || !xbar || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    AdjointState* A = e->adj;
    const uint32_t n = e->models[model].n;
    CK_HIP(e, hipMemcpyAsync(A->host_io, xbar, sizeof(float) * 3 * n, hipMemcpyHostToDevice, e->stream));
    JoinBack(e);
    SeedKernel<<<AGrid(n), kAT, 0, e->stream>>>(A->ids[A->idc][model], A->host_io, n, A->xb[A->ab][model]);
    A->need_join = true;
    CK_HIP(e, hipStreamSynchronize(e->stream));
    return CKMPM_OK;
}

int ckmpm_adjoint_substep(ckmpm_engine* e) {
    if (!e || !e->adj) return CKMPM_ERR_INVALID;
    AdjointState* A = e->adj;
    const int tp = A->recording ? A->rec : A->und;
    if (A->tlen[tp] == 0) return Fail(e, CKMPM_ERR_INVALID, "no recorded substep to undo");
    const Snapshot& s = A->tape[tp][A->tlen[tp] - 1];
    hipStream_t st = A->ustream;
    if (A->need_join) {
        // the recording of this tape (or the reset / seed) is queued on the engine stream
        CK_HIP(e, hipEventRecord(A->join_ev, e->stream));
        CK_HIP(e, hipStreamWaitEvent(st, A->join_ev, 0));
        A->need_join = false;
    }
    const uint32_t grid = GridSize(e);
    float* Mout = A->Mbar[1 - A->mb];
    const int nb = 1 - A->ab;
    for (size_t i = 0; i < e->models.size(); ++i) {
        const ModelBuf& m = e->models[i];
        AdjArgs a{};
        a.dom = e->dom;
        a.mat = m.mat;
        a.keys = s.keys[0];
        a.table = s.table[0];
        a.lc = s.lc[0];
        a.V = s.V;
        a.Mbar = A->Mbar[A->mb];
        a.Vbar = A->Vbar;
        a.st_in = s.st[i];
        a.perm = s.perm[i];
        a.off = s.off[i];
        a.n = m.n;
        a.pp_attr = m.pp_attr;
        a.xb_in = A->xb[A->ab][i];
        a.Eb_in = A->Eb[A->ab][i];
        a.tb_in = A->tb[A->ab][i];
        a.xb_out = A->xb[nb][i];
        a.Eb_out = A->Eb[nb][i];
        a.tb_out = A->tb[nb][i];
        a.dt = static_cast<float>(s.dt);
        a.next_dt = static_cast<float>(s.nd);
        // M_k-bar lives in the previous layout; blocks beyond its Nb are never read
        a.Mzero = i == 0 ? Mout : nullptr;
        a.lc_zero = s.lc[1];
        a.cap = e->cap;
        // few particle blocks (small scenes): several workgroups per block so that the
        // launch still covers the 256 CUs, but no more than the block's 256-particle
        // chunks (a workgroup without a chunk would only hold a CU slot)
        const uint32_t blocks = std::max<uint32_t>(e->a_hint, 1);
        const uint32_t chunks = static_cast<uint32_t>((static_cast<uint64_t>(m.n) + static_cast<uint64_t>(blocks) * kAT - 1) /
                                                      (static_cast<uint64_t>(blocks) * kAT));
        a.split = std::min<uint32_t>({8u, std::max<uint32_t>(1, (2048 + blocks) / blocks), std::max<uint32_t>(chunks, 1)});
        const uint32_t bgrid = std::min<uint32_t>(grid * a.split, 65536u / a.split * a.split);
        if (m.kind == kFixedCorotated) G2P2GBackwardKernel<kFixedCorotated><<<bgrid, kAT, 0, st>>>(a);
        else G2P2GBackwardKernel<kNeoHookean><<<bgrid, kAT, 0, st>>>(a);
    }
    GridBwdArgs gb{};
    gb.dom = e->dom;
    gb.keys = s.keys[0];
    gb.lc = s.lc[0];
    gb.table_prev = s.table[1];
    gb.lc_prev = s.lc[1];
    gb.M = s.M;
    gb.Vbar = A->Vbar;
    gb.Mbar = Mout;
    gb.mass_clamp = e->sim.mass_clamp;
    gb.cap = e->cap;
    GridBackwardKernel<<<std::max<uint32_t>(2 * grid, 64u), kAT, 0, st>>>(gb);
    A->ab = nb;
    A->mb = 1 - A->mb;
    A->at_setup = s.substeps == 0;
    A->setup_table = s.table[1];
    A->setup_lc = s.lc[1];
    --A->tlen[tp];
    CK_HIP(e, hipEventRecord(A->free_ev[tp], st));
    A->undone[tp] = true;
    CK_HIP(e, hipGetLastError());
    return CKMPM_OK;
}

int ckmpm_adjoint_gradients(ckmpm_engine* e, int32_t model, float* lambda_bar, float* mu_bar, float* v0_bar) {
    if (!e || !e->adj || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    AdjointState* A = e->adj;
    if (!A->at_setup) return Fail(e, CKMPM_ERR_INVALID, "undo every substep back to the first before reading gradients");
    const ModelBuf& m = e->models[model];
    hipStream_t st = e->stream;
    const uint32_t n = m.n;
    std::vector<float> tmp(3 * static_cast<size_t>(n));
    JoinBack(e);
    if (v0_bar) {
        RasterizeBackwardKernel<<<AGrid(n), kAT, 0, st>>>(e->dom, A->x0[model], n, m.mat.mass, A->setup_table, A->setup_lc,
                                                          A->Mbar[A->mb], A->host_io);
        CK_HIP(e, hipMemcpyAsync(v0_bar, A->host_io, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
        CK_HIP(e, hipStreamSynchronize(st));
    }
    if (lambda_bar || mu_bar) {
        // (lambda-bar, mu-bar) are in setup processing order: map to input order
        ToInputOrderKernel<<<AGrid(n), kAT, 0, st>>>(A->setup_order[model], A->tb[A->ab][model], n, 2, A->host_io);
        CK_HIP(e, hipMemcpyAsync(tmp.data(), A->host_io, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, st));
        CK_HIP(e, hipStreamSynchronize(st));
        for (uint32_t p = 0; p < n; ++p) {
            if (lambda_bar) lambda_bar[p] = tmp[2 * static_cast<size_t>(p)];
            if (mu_bar) mu_bar[p] = tmp[2 * static_cast<size_t>(p) + 1];
        }
    }
    CK_HIP(e, hipGetLastError());
    return CKMPM_OK;
}

int ckmpm_get_positions_input_order(ckmpm_engine* e, int32_t model, float* xyz) {
    if (!e || !e->adj || !xyz || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    if (!e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "call ckmpm_initial_setup first");
    AdjointState* A = e->adj;
    ModelBuf& m = e->models[model];
    const uint32_t n = m.n;
    if (!e->pos_stage) {  // the staging buffer of ckmpm_get_positions, kept
        uint32_t nmax = 0;
        for (const ModelBuf& q : e->models) nmax = std::max(nmax, q.n);
        CK_TRY(Alloc(e, &e->pos_stage, 3 * static_cast<size_t>(nmax)));
    }
    // processing-order AoS positions, then to input order through the identities on the
    // device (host_io holds 3 n floats), then one copy to the caller
    JoinBack(e);  // host_io: after any undo still queued
    const LayoutBuf& P = e->L[1 - e->cur];
    const uint32_t wgs = std::max<uint32_t>(8u, std::min<uint32_t>(e->a_hint + 8u, 16384u));
    LaunchExportPositions(e->dom, P.keys, P.lc, m.off[1 - e->sc], m.st[e->sc], n, e->pos_stage, wgs, e->stream);
    AosToInputOrderKernel<<<AGrid(n), kAT, 0, e->stream>>>(A->ids[A->idc][model], e->pos_stage, n, A->host_io);
    CK_HIP(e, hipMemcpyAsync(xyz, A->host_io, 3 * sizeof(float) * static_cast<size_t>(n), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    return CKMPM_OK;
}

}  // extern "C"
