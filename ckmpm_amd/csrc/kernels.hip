// CK-MPM kernels for gfx950 (wave64).
//
// Data layout in HBM (all FP32 / u32, structure of arrays):
//   particle state   st[attr][n]      u (3: position relative to the origin 4 b dx of the
//                                     block b that processed the particle last), then
//                                     E = F - I (9, row-major) | q = J - 1 (fluid), then
//                                     logJp (DP)
//   ordering         perm[t] = slot | move << 27: slot of the previous state array and
//                    the particle's block move since then; off[j] block ranges
//   grid mass/mom.   M[block][ch 4][node 128]   node = g*64 + lx*16 + ly*4 + lz
//   grid velocity    V[block][ch 3][node 128]
//   layout           keys[list index] -> linear block key, table[key] -> list index
#include "kernels.h"

#include <algorithm>
#include "stencil.h"
#include "material.h"

namespace ckmpm {

namespace {

constexpr int kThreads = 256;
// Bitmap words per thread of the emit kernel (a thread emits the marked blocks of its
// words one after the other: 2 for small tiles, 8 for large) and exterior blocks per
// thread of the rebin scan.  Small tiles spread both over more workgroups: with 8 words
// and 8 blocks per thread the 8M scene's scan ran in 12 workgroups, and the substep took
// 0.706 ms against 0.690 ms.
static_assert(kEmitTileWordsSmall == kThreads * 2 && kEmitTileWordsLarge == kThreads * 8, "emit tiles");
constexpr int kCountThreads = 1024;                    // rebin count workgroup
constexpr int kCountTile = kCountThreads / 32;         // exterior blocks per count tile
constexpr int kRebinScanItems = 1;
constexpr uint32_t kRebinScanTile = kThreads * kRebinScanItems;
static_assert(kRebinScanTile % kCountTile == 0, "scan tiles hold whole count tiles");

__device__ __forceinline__ uint32_t LaneId() { return __lane_id(); }

__device__ __forceinline__ uint32_t WaveInclusiveScan(uint32_t v) {
    const uint32_t lane = LaneId();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= static_cast<uint32_t>(o)) v += u;
    }
    return v;
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t& total, uint32_t* sh) {
    const uint32_t lane = LaneId(), wave = threadIdx.x >> 6;
    const uint32_t inc = WaveInclusiveScan(v);
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < kThreads / 64; ++w) {
        const uint32_t x = sh[w];
        wbase += (w < wave) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return wbase + inc - v;
}

// Number of set bits of a wave mask below this lane (mbcnt).
__device__ __forceinline__ uint32_t LanesBelow(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Lanes of the wave holding the same small integer value (bits significant bits).
template <int bits>
__device__ __forceinline__ uint64_t MatchPeers(uint32_t v) {
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < bits; ++b) {
        const uint64_t m = __ballot((v >> b) & 1u);
        peers &= ((v >> b) & 1u) ? m : ~m;
    }
    return peers;
}

__device__ __forceinline__ uint32_t ReadCount(const uint32_t* n_dev, uint32_t n_imm) { return n_dev ? *n_dev : n_imm; }

// ============================================================== scans
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;

template <bool kPopc>
__device__ __forceinline__ uint32_t ScanValue(const uint32_t* in, uint32_t i, uint32_t n) {
    if (i >= n) return 0u;
    const uint32_t v = in[i];
    return kPopc ? static_cast<uint32_t>(__popc(v)) : v;
}

template <bool kPopc>
__global__ __launch_bounds__(kThreads) void ScanReduceKernel(const uint32_t* in, const uint32_t* n_dev, uint32_t n_imm,
                                                             uint32_t* partial) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint32_t n = ReadCount(n_dev, n_imm);
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += ScanValue<kPopc>(in, base + k * kThreads + threadIdx.x, n);
    uint32_t tot;
    BlockExclusiveScan(s, tot, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void ScanPartialsKernel(uint32_t* partial, const uint32_t* n_dev, uint32_t n_imm,
                                                               uint32_t* total) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint32_t n = ReadCount(n_dev, n_imm);
    const uint32_t ntiles = (n + kScanTile - 1) / kScanTile;
    uint32_t carry = 0;
    for (uint32_t c = 0; c < ntiles; c += kThreads) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < ntiles ? partial[i] : 0u;
        uint32_t tot;
        const uint32_t ex = BlockExclusiveScan(v, tot, sh);
        if (i < ntiles) partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <bool kPopc>
__global__ __launch_bounds__(kThreads) void ScanDownsweepKernel(const uint32_t* in, uint32_t* out, const uint32_t* n_dev,
                                                                uint32_t n_imm, const uint32_t* partial) {
    __shared__ uint32_t sh[kThreads / 64];
    const uint32_t n = ReadCount(n_dev, n_imm);
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;
    // thread-contiguous items: thread i owns [base + i*8, base + i*8 + 8)
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = ScanValue<kPopc>(in, base + threadIdx.x * kScanItems + k, n);
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = partial[blockIdx.x] + BlockExclusiveScan(s, tot, sh);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t i = base + threadIdx.x * kScanItems + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

// The reduce kernel sums strided items and the downsweep contiguous ones; both cover
// the same tile, so the tile sums agree.

// ============================================================== radix sort (setup)
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kThreads * kRadixItems;

__global__ __launch_bounds__(kThreads) void RadixHistKernel(const uint32_t* keys, uint32_t n, int shift, uint32_t* hist,
                                                            uint32_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        const uint32_t i = base + k * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void RadixScatterKernel(const uint32_t* keys, const uint32_t* vals,
                                                               uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                                               int shift, const uint32_t* hist, uint32_t ntiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kThreads / 64][257];
    run[threadIdx.x] = hist[threadIdx.x * ntiles + blockIdx.x];
    const uint32_t lane = LaneId(), wave = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kRadixTile;
    for (int k = 0; k < kRadixItems; ++k) {
        for (int i = threadIdx.x; i < (kThreads / 64) * 257; i += kThreads) (&wcnt[0][0])[i] = 0;
        __syncthreads();
        const uint32_t i = base + k * kThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t val = valid ? vals[i] : 0u;
        const uint32_t digit = valid ? ((key >> shift) & 255u) : 256u;
        const uint64_t peers = MatchPeers<9>(digit);
        const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt));
        if (rank == 0) wcnt[wave][digit] = static_cast<uint32_t>(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t pos = run[digit] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][digit];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t w = 0; w < kThreads / 64; ++w) add += wcnt[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

// ============================================================== setup kernels
__global__ void SetupKeysKernel(DomainDev d, const float* xyz, uint32_t n, uint32_t* keys, uint32_t* vals,
                                uint32_t* bitmap, Status* st) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        int b[3];
        bool out = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            b[a] = FloorDiv4(Cell0(xyz[3 * static_cast<size_t>(p) + a], d.inv_dx) - 2);
            const int c = min(max(b[a], 0), d.dim[a] - 1);
            out |= c != b[a];
            b[a] = c;
        }
        if (out) SetFlag(st, kFlagOutOfDomain);
        const uint32_t key = KeyOf(d, b[0], b[1], b[2]);
        // binning cell inside the block, from the unclamped cell: ((c - 2) mod 4) per axis
        uint32_t cb = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            cb = cb * 4u + static_cast<uint32_t>((Cell0(xyz[3 * static_cast<size_t>(p) + a], d.inv_dx) - 2) & 3);
        keys[p] = (key << 6) | cb;
        vals[p] = p;
        atomicOr(&bitmap[key >> 5], 1u << (key & 31u));
    }
}

// Setup ordering inside a block: (rank of the particle among the earlier input
// particles of its binning cell, cell).  A block's particles then cycle through its
// cells, so the lanes of a wavefront deposit into distinct grid nodes in P2G.
// Computed from the cell-key-sorted list: heads of equal-key runs, their positions,
// then the rank = position - run start.
__global__ void CellHeadsKernel(const uint32_t* ck, uint32_t n, uint32_t* head) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        head[i] = (i == 0 || ck[i] != ck[i - 1]) ? 1u : 0u;
}
__global__ void CellStartsKernel(const uint32_t* head, const uint32_t* ex, uint32_t n, uint32_t* start) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (head[i]) start[ex[i]] = i;
}
// in place: head -> within-block sort key (rank << 6 | cell), ex -> block key
__global__ void CellRankKernel(const uint32_t* ck, uint32_t* head_k2, uint32_t* ex_bk, const uint32_t* start,
                               uint32_t n, uint32_t* idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t seg = ex_bk[i] + head_k2[i] - 1u;
        const uint32_t rank = min(i - start[seg], kSetupRankClamp);
        head_k2[i] = (rank << 6) | (ck[i] & 63u);
        ex_bk[i] = ck[i] >> 6;
        idx[i] = i;
    }
}
__global__ void SetupGatherKernel(const uint32_t* idx, const uint32_t* bk, const uint32_t* orig, uint32_t n,
                                  uint32_t* keys, uint32_t* order) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t i = idx[j];
        keys[j] = bk[i];
        order[j] = orig[i];
    }
}

__global__ void SetupOffsetsKernel(const uint32_t* sorted, uint32_t n, const uint32_t* akeys, const LayoutCounts* lc,
                                   uint32_t* off) {
    const uint32_t A = lc->A;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= A; j += gridDim.x * blockDim.x) {
        if (j == A) {
            off[j] = n;
            continue;
        }
        const uint32_t key = akeys[j];
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sorted[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        off[j] = lo;
    }
}

// Initial state in processing order (CopyParticleArrayToParticleBuffer,
// mpm_algorithm.cuh:973-1081): slot t holds input particle order[t], its position
// relative to its block origin 4 b dx (exact for power-of-two dx: the product is exact
// and the difference of two nearby numbers is exact); F = I (E = 0), logJp = 0,
// J = J0 (q = J0 - 1).  The
// ordering becomes the identity with move 13 (no move), like InitializeAdvectionBucket
// (mpm_algorithm.cuh:949-964).
__global__ void InitStateKernel(DomainDev d, int kind, const float* xyz, const float* lam_p, const float* mu_p,
                                int pp_attr, uint32_t* order, uint32_t n, float init, float* st, int rigid_attr,
                                uint64_t rigid_count) {
    const size_t N = n;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t src = order[p];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float x = xyz[3 * static_cast<size_t>(src) + a];
            const int b = min(max(FloorDiv4(Cell0(x, d.inv_dx) - 2), 0), d.dim[a] - 1);
            st[a * N + p] = __fsub_rn(x, __fmul_rn(static_cast<float>(4 * b), d.dx));
        }
        order[p] = p | (13u << kMoveShift);
        if (pp_attr) {
            st[pp_attr * N + p] = lam_p[src];
            st[(pp_attr + 1) * N + p] = mu_p[src];
        }
        if (rigid_attr) st[rigid_attr * N + p] = src < rigid_count ? 1.0f : 0.0f;
        if (kind == kFluid) {
            st[3 * N + p] = init - 1.0f;  // the fluid state is q = J - 1 (init = J0)
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) st[(3 + k) * N + p] = 0.0f;  // E = F - I = 0
            if (kind == kDruckerPrager) st[12 * N + p] = 0.0f;
            if (kind == kCamClay) st[12 * N + p] = init;  // logJp = alpha0
        }
    }
}


// mpm_algorithm.cuh:871-947 (Rasterize): m w and m w v0 on both grids, global atomics.
__global__ void RasterizeKernel(DomainDev d, const float* xyz, const float* vel, uint32_t n, uint32_t nvel, float mass,
                                const uint32_t* table, const LayoutCounts* lc, float* M, Status* st) {
    const uint32_t Nb = lc->Nb;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        float xs[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) xs[a] = __fmul_rn(xyz[3 * static_cast<size_t>(p) + a], d.inv_dx);
        const float* v = nvel == 1 ? vel : vel + 3 * static_cast<size_t>(p);
        const float v0 = v[0], v1 = v[1], v2 = v[2];
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            Stencil32 s;
            if (g == 0) MakeStencil<0>(xs, s);
            else MakeStencil<1>(xs, s);
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
                const int cx = s.c[0] + ox, cy = s.c[1] + oy, cz = s.c[2] + oz;
                const int bx = FloorDiv4(cx), by = FloorDiv4(cy), bz = FloorDiv4(cz);
                uint32_t bi = kEmpty;
                if (InDomain(d, bx, by, bz)) bi = table[KeyOf(d, bx, by, bz)];
                if (bi == kEmpty || bi >= Nb) {
                    SetFlag(st, kFlagMissingBlock);
                    continue;
                }
                const int node = g * 64 + (cx - 4 * bx) * 16 + (cy - 4 * by) * 4 + (cz - 4 * bz);
                const float wm = NodeWeight(s, ox, oy, oz) * mass;
                float* blk = M + static_cast<size_t>(bi) * kMFloats;
                unsafeAtomicAdd(blk + node, wm);
                unsafeAtomicAdd(blk + 128 + node, wm * v0);
                unsafeAtomicAdd(blk + 256 + node, wm * v1);
                unsafeAtomicAdd(blk + 384 + node, wm * v2);
            }
        }
    }
}

// Global positions of the stored state: slot t was written while processing block j of
// the previous ordering (off_prev), relative to that block's origin.
__global__ __launch_bounds__(kThreads) void ExportPositionsKernel(DomainDev d, const uint32_t* keys,
                                                                  const LayoutCounts* lc, const uint32_t* off,
                                                                  const float* st, uint32_t n, float* xyz) {
    const uint32_t A = lc->A;
    const size_t N = n;
    for (uint32_t j = blockIdx.x; j < A; j += gridDim.x) {
        int b[3];
        CoordsOf(d, keys[j], b[0], b[1], b[2]);
        for (uint32_t t = off[j] + threadIdx.x; t < off[j + 1] && t < n; t += kThreads)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                xyz[3 * static_cast<size_t>(t) + k] = __fadd_rn(__fmul_rn(static_cast<float>(4 * b[k]), d.dx), st[k * N + t]);
    }
}

// ============================================================== layout kernels
__global__ void TableInsertAKernel(const uint32_t* keys, const LayoutCounts* lc, uint32_t cap, uint32_t* table) {
    const uint32_t A = min(lc->A, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A; i += gridDim.x * blockDim.x) table[keys[i]] = i;
}

// offset of move d (d = (ox + 1) * 9 + (oy + 1) * 3 + oz + 1)
__device__ __forceinline__ void DirOffset(int d, int& ox, int& oy, int& oz) {
    ox = d / 9 - 1;
    oy = (d / 3) % 3 - 1;
    oz = d % 3 - 1;
}

// Neighbour (b + {0,1}^3) and exterior (b + {-1,0,1}^3) candidates of the particle blocks
// that are not in the table, in one pass over the 27 offsets (mpm_algorithm.cuh:657-695
// marks them in two passes, the exterior one after the neighbours are registered).
// Block key k owns bits 2 (k % 16) (neighbour mark) and 2 (k % 16) + 1 (exterior mark)
// of map word k / 16: a block with the neighbour mark is a neighbour block, one with
// only the exterior mark an exterior block.  Each tile of map words
// counts both kinds (tile_cnt[2t], tile_cnt[2t+1]) as marks are set: the atomic OR on
// the key's word orders the marks of one block, so the first neighbour mark that finds
// the exterior mark set takes that block's exterior count back, and an exterior mark
// counts only where neither mark was set.  The emit kernel then knows every tile's
// offsets without a separate scan pass.  The counts are aggregated per wavefront (its
// candidates span one to three tiles): one atomic per distinct tile and kind instead of
// one per mark, which would serialize on the few counter addresses.
// wg / ng: workgroup index and workgroup count of the launch (the scene's part of a fused
// launch); the same holds for every kernel body below.
__device__ __forceinline__ void MarkExtensionBody(const MarkArgs& a, uint32_t wg, uint32_t ng) {
    const DomainDev& d = a.d;
    const uint32_t* keys = a.keys;
    const uint32_t* table = a.table;
    uint32_t* map = a.map;
    uint32_t* tile_cnt = a.tile_cnt;
    const uint32_t A = min(a.lc->A, a.cap);
    const uint32_t total = A * static_cast<uint32_t>(kDirections);
    for (uint32_t q = wg * kThreads + threadIdx.x; q < total; q += ng * kThreads) {
        const uint32_t i = q / static_cast<uint32_t>(kDirections);
        const int o = static_cast<int>(q - i * static_cast<uint32_t>(kDirections));
        int bx, by, bz, ox, oy, oz;
        CoordsOf(d, keys[i], bx, by, bz);
        DirOffset(o, ox, oy, oz);
        const bool nbr = ox >= 0 && oy >= 0 && oz >= 0;
        bx += ox;
        by += oy;
        bz += oz;
        uint32_t tile = kEmpty;  // tile of a count changed by this lane
        int dn = 0, de = 0;      // neighbour and exterior count changes
        if (InDomain(d, bx, by, bz)) {
            const uint32_t key = KeyOf(d, bx, by, bz);
            if (table[key] == kEmpty) {
                const uint32_t sh = 2u * (key % kExtKeysPerWord);
                const uint32_t old = atomicOr(&map[key / kExtKeysPerWord], (nbr ? 1u : 2u) << sh) >> sh;
                if (nbr) {
                    dn = (old & 1u) ? 0 : 1;
                    de = (dn && (old & 2u)) ? -1 : 0;
                } else {
                    de = (old & 3u) ? 0 : 1;
                }
                if (dn || de) tile = key >> a.tile_shift;
            }
        }
        uint64_t pending = __ballot(tile != kEmpty);
        while (pending) {
            const int leader = __builtin_ctzll(pending);
            const uint32_t t = __shfl(tile, leader, 64);
            const bool mine = tile == t;
            const uint64_t same = __ballot(mine);
            const uint32_t nn = static_cast<uint32_t>(__popcll(__ballot(mine && dn > 0)));
            const uint32_t ep = static_cast<uint32_t>(__popcll(__ballot(mine && de > 0)));
            const uint32_t em = static_cast<uint32_t>(__popcll(__ballot(mine && de < 0)));
            if (LaneId() == static_cast<uint32_t>(leader)) {
                if (nn) atomicAdd(&tile_cnt[2u * t], nn);
                if (ep != em) atomicAdd(&tile_cnt[2u * t + 1u], ep - em);  // modulo 2^32
            }
            pending &= ~same;
        }
    }
}
__global__ __launch_bounds__(kThreads) void MarkExtensionKernel(MarkArgs a) { MarkExtensionBody(a, blockIdx.x, gridDim.x); }

// Reduce per-thread values over a workgroup into red[0, nq) (all threads see the result
// after the call).  red must hold nq entries.
template <int NQ>
__device__ __forceinline__ void BlockSums(const uint32_t (&v)[NQ], int nq, uint32_t* red) {
    __syncthreads();  // earlier reads of red are done
    if (threadIdx.x < static_cast<uint32_t>(nq)) red[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        if (j < nq) {
            uint32_t x = v[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            if (LaneId() == 0 && x) atomicAdd(&red[j], x);
        }
    }
    __syncthreads();
}

// Layout extension in one kernel: workgroup w owns map words [w, w+1) * kThreads * ITEMS,
// takes its first positions from the mark kernel's tile counts (neighbour blocks follow
// A, exterior blocks follow all neighbour blocks), scans its words' mark counts and
// writes the blocks of each kind in ascending key order to keys[] and the table,
// clearing the map.  Tiles without a mark return at once.  Workgroup 0 sets Nb and Ex.
template <int ITEMS>
__device__ __forceinline__ void BitmapScanEmitBody(const EmitArgs& a, uint32_t wg, uint32_t) {
    static_assert(ITEMS == 2 || ITEMS % 4 == 0, "emit words per thread: 2 or a multiple of 4");
    __shared__ uint32_t sh[kThreads / 64];
    __shared__ uint32_t red[4];
    uint32_t* map = a.map;
    const uint32_t* tile_cnt = a.tile_cnt;
    const uint32_t ntiles = a.ntiles, cap = a.cap;
    uint32_t* keys = a.keys;
    uint32_t* table = a.table;
    LayoutCounts* lc = a.lc;
    Status* st = a.st;
    const uint32_t w = wg, tid = threadIdx.x;
    const bool marked = tile_cnt[2u * w] != 0u || tile_cnt[2u * w + 1u] != 0u;
    if (w != 0 && !marked) return;
    // neighbour total; neighbour and exterior blocks of the tiles before w; exterior total
    uint32_t v[4] = {0u, 0u, 0u, 0u};
    for (uint32_t t = tid; t < ntiles; t += kThreads) {
        const uint32_t cn = tile_cnt[2u * t], ce = tile_cnt[2u * t + 1u];
        v[0] += cn;
        v[3] += ce;
        if (t < w) {
            v[1] += cn;
            v[2] += ce;
        }
    }
    BlockSums<4>(v, 4, red);
    const uint32_t A = lc->A;
    uint32_t nb = A + red[0];
    if (nb > cap) nb = cap;
    if (w == 0 && tid == 0) {
        uint32_t ex = nb + red[3];
        if (A + red[0] > cap || ex > cap) {
            SetFlag(st, kFlagCapacity);
            ex = min(ex, cap);
        }
        lc->Nb = nb;
        lc->Ex = ex;
    }
    if (!marked) return;
    const uint32_t w0 = (w * kThreads + tid) * ITEMS;
    uint32_t words[ITEMS];
    if constexpr (ITEMS % 4 == 0) {
#pragma unroll
        for (int q = 0; q < ITEMS / 4; ++q) {
            const uint4 v = reinterpret_cast<const uint4*>(map + w0)[q];
            words[4 * q] = v.x;
            words[4 * q + 1] = v.y;
            words[4 * q + 2] = v.z;
            words[4 * q + 3] = v.w;
        }
    } else {
        const uint2 v = *reinterpret_cast<const uint2*>(map + w0);
        words[0] = v.x;
        words[1] = v.y;
    }
    constexpr uint32_t kLow = 0x55555555u;  // the neighbour mark of each block of a word
    uint32_t pn = 0, pe = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        pn += static_cast<uint32_t>(__popc(words[k] & kLow));
        pe += static_cast<uint32_t>(__popc((words[k] >> 1) & ~words[k] & kLow));
    }
    uint32_t tot;
    uint32_t posn = A + red[1] + BlockExclusiveScan(pn, tot, sh);
    uint32_t pose = nb + red[2] + BlockExclusiveScan(pe, tot, sh);
    if (pn + pe == 0) return;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (!words[k]) continue;
        map[w0 + k] = 0u;
        uint32_t bn = words[k] & kLow, be = (words[k] >> 1) & ~words[k] & kLow;
        const uint32_t key0 = (w0 + k) * kExtKeysPerWord;
        while (bn) {
            const uint32_t key = key0 + static_cast<uint32_t>(__ffs(bn) - 1) / 2u;
            bn &= bn - 1u;
            if (posn < cap) {
                keys[posn] = key;
                table[key] = posn;
            }
            ++posn;
        }
        while (be) {
            const uint32_t key = key0 + static_cast<uint32_t>(__ffs(be) - 1) / 2u;
            be &= be - 1u;
            if (pose < cap) {
                keys[pose] = key;
                table[key] = pose;
            }
            ++pose;
        }
    }
}
template <int ITEMS>
__global__ __launch_bounds__(kThreads) void BitmapScanEmitKernel(EmitArgs a) {
    BitmapScanEmitBody<ITEMS>(a, blockIdx.x, gridDim.x);
}

// Emit set bits in ascending key order at position base + word_off[w] + rank.
// pass 0 / 1: append after A / after Nb and register in the table; pass 2: particle
// blocks of the setup (positions from 0, no table).
__global__ void BitmapEmitKernel(uint32_t* bitmap, uint32_t nwords, const uint32_t* word_off, uint32_t* keys,
                                 uint32_t* table, const LayoutCounts* lc, uint32_t cap, int pass, Status* st) {
    const uint32_t base = pass == 0 ? lc->A : (pass == 1 ? lc->Nb : 0u);
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x) {
        uint32_t bits = bitmap[w];
        if (!bits) continue;
        uint32_t pos = base + word_off[w];
        bitmap[w] = 0;
        while (bits) {
            const uint32_t b = static_cast<uint32_t>(__ffs(bits) - 1);
            bits &= bits - 1u;
            const uint32_t key = w * 32u + b;
            if (pos < cap) {
                keys[pos] = key;
                if (table) table[key] = pos;
            } else {
                SetFlag(st, kFlagCapacity);
            }
            ++pos;
        }
    }
}

// counts.field = base + total (field 0: A = total, 1: Nb = A + total, 2: Ex = Nb + total)
__global__ void SetCountKernel(LayoutCounts* lc, int field, const uint32_t* total, uint32_t cap, Status* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t v = *total;
    if (field == 1) v += lc->A;
    if (field == 2) v += lc->Nb;
    if (v > cap) {
        SetFlag(st, kFlagCapacity);
        v = cap;
    }
    if (field == 0) lc->A = v;
    if (field == 1) lc->Nb = v;
    if (field == 2) lc->Ex = v;
}

// ============================================================== grid update

// atomicMax of a non-negative float (as bits) that skips the atomic when the stored
// value is already at least v: the stored value only grows during a launch, so a
// stale read can only cause an unneeded atomic, never a missed one.
__device__ __forceinline__ void MaxBits(uint32_t* p, float v) {
    if (v > 0.0f && __float_as_uint(v) > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(p, __float_as_uint(v));
}

// mpm_algorithm.cuh:738-869 fused with CopySelectedGridBlocks (:1920-1945): each grid
// block of the current layout reads its mass / momentum from the previous layout
// through the previous key table, writes velocities (p/m, gravity on y, sticky
// boundary within 2 blocks of the domain edge), clears the next P2G target and
// reduces max |v|^2 (NaN -> +inf) and the P2G deposit bounds of the previous substep.
// 32 threads per block (4 consecutive nodes each, 16-byte loads and stores), 8 blocks
// per workgroup and step.  Each thread serves kGridIlp blocks at a time: the key ->
// table -> mass / momentum load chains of those blocks are issued together, so a wave
// has kGridIlp independent latency chains in flight instead of one.
constexpr int kGridIlp = 4;
// A scene's boundary description (the reference's ProcessGridCellVelocity hook,
// mpm_algorithm.cuh:795-799, as data): regions in order on the node's integer cell, then
// the walls within 8 cells of the domain faces.
__device__ __forceinline__ void ApplyBoundary(const ckmpm_boundary_desc& bc, const DomainDev& dom, const int (&cell)[3],
                                           float (&v)[3]) {
    for (uint32_t r = 0; r < bc.nregions && r < CKMPM_MAX_REGIONS; ++r) {
        const ckmpm_region& g = bc.regions[r];
        bool in = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) in &= cell[k] >= g.lo[k] && cell[k] < g.hi[k];
        if (!in) continue;
        if (g.kind == CKMPM_REGION_FIXED) {
            v[0] = v[1] = v[2] = 0.0f;
        } else if (g.kind == CKMPM_REGION_VELOCITY) {
            v[0] = g.value[0];
            v[1] = g.value[1];
            v[2] = g.value[2];
        } else {
            float q[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = static_cast<float>(cell[k]) * dom.dx - g.center[k];
            v[0] = g.value[1] * q[2] - g.value[2] * q[1];
            v[1] = g.value[2] * q[0] - g.value[0] * q[2];
            v[2] = g.value[0] * q[1] - g.value[1] * q[0];
        }
    }
    if (bc.walls == CKMPM_WALLS_NONE) return;
    float nrm[3];
    bool on = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        nrm[k] = cell[k] < 8 ? 1.0f : (cell[k] >= 4 * dom.dim[k] - 8 ? -1.0f : 0.0f);
        on |= nrm[k] != 0.0f;
    }
    if (!on) return;
    if (bc.walls == CKMPM_WALLS_STICKY) {
        v[0] = v[1] = v[2] = 0.0f;
        return;
    }
    const float inv = rsqrtf(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) nrm[k] *= inv;
    const float vn = nrm[0] * v[0] + nrm[1] * v[1] + nrm[2] * v[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] -= vn * nrm[k];
}

// kCustom: the scene's boundary description replaces the default sticky margin (a
// separate instantiation, so that the default path keeps its registers and code size)
template <bool kCustom>
__device__ __forceinline__ void GridUpdateBody(const GridUpdateArgs& a, uint32_t wg, uint32_t ng) {
    const uint32_t Nb = min(a.lc->Nb, a.cap), Nbp = a.lc_prev->Nb;
    const uint32_t n0 = (threadIdx.x & 31u) * 4u;  // nodes n0 .. n0 + 3 of each of its blocks
    constexpr uint32_t kPer = kThreads / 32;       // blocks per workgroup and step
    const uint32_t stride = ng * kPer;
    // clear the other parity slot for the next substep (nothing else touches it now)
    if (wg == 0) {
        if (threadIdx.x == 0) a.st->vmax2_bits[a.slot ^ 1] = 0u;
        if (threadIdx.x < static_cast<uint32_t>(kMaxModels)) a.st->dep_bits[a.slot ^ 1][threadIdx.x] = 0u;
    }
    float vmax = 0.0f;
    for (uint32_t i0 = wg * kPer + (threadIdx.x >> 5); i0 < Nb; i0 += stride * kGridIlp) {
        uint32_t key[kGridIlp], io[kGridIlp];
#pragma unroll
        for (int u = 0; u < kGridIlp; ++u) {
            const uint32_t i = i0 + u * stride;
            key[u] = i < Nb ? a.keys[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kGridIlp; ++u) {
            const uint32_t i = i0 + u * stride;
            io[u] = i < Nb ? a.table_prev[key[u]] : kEmpty;
        }
        float4 m[kGridIlp], p[kGridIlp][3];
#pragma unroll
        for (int u = 0; u < kGridIlp; ++u) {
            m[u] = p[u][0] = p[u][1] = p[u][2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (io[u] != kEmpty && io[u] < Nbp) {
                const float4* src = reinterpret_cast<const float4*>(a.M_prev + static_cast<size_t>(io[u]) * kMFloats + n0);
                m[u] = src[0];
                p[u][0] = src[32];
                p[u][1] = src[64];
                p[u][2] = src[96];
            }
        }
#pragma unroll
        for (int u = 0; u < kGridIlp; ++u) {
            const uint32_t i = i0 + u * stride;
            if (i >= Nb) continue;
            int b[3];
            CoordsOf(a.dom, key[u], b[0], b[1], b[2]);
            bool oob = false;
#pragma unroll
            for (int k = 0; k < 3; ++k) oob |= (b[k] + a.dom.off[k] < 2) || (b[k] + a.dom.off[k] + 2 >= a.dom.dim[k]);
            const float mk[4] = {m[u].x, m[u].y, m[u].z, m[u].w};
            const float pk[3][4] = {{p[u][0].x, p[u][0].y, p[u][0].z, p[u][0].w},
                                    {p[u][1].x, p[u][1].y, p[u][1].z, p[u][1].w},
                                    {p[u][2].x, p[u][2].y, p[u][2].z, p[u][2].w}};
            float vo[3][4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t n = n0 + q;
                float v[3] = {0.0f, 0.0f, 0.0f};
                float v2 = 0.0f;
                if (mk[q] > a.mass_clamp) {
                    const float im = __builtin_amdgcn_rcpf(mk[q]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] = pk[k][q] * im;
                    v[1] += a.gravity * a.dt;
                    if (kCustom) {
                        // node (x, y, z) of either sub-grid: cell 4 b + (x, y, z), as the hook sees it
                        const int cell[3] = {4 * b[0] + static_cast<int>((n >> 4) & 3u),
                                             4 * b[1] + static_cast<int>((n >> 2) & 3u),
                                             4 * b[2] + static_cast<int>(n & 3u)};
                        ApplyBoundary(a.bc->d, a.dom, cell, v);
                    } else if (oob) {
                        v[0] = v[1] = v[2] = 0.0f;
                    }
                    v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) vo[k][q] = v[k];
                if (isnan(v2)) {
                    v2 = INFINITY;
                    SetFlag(a.st, kFlagNonFinite);
                }
                if (mk[q] >= a.vmax_mass_floor || v2 == INFINITY) vmax = fmaxf(vmax, v2);
            }
            float4* dst = reinterpret_cast<float4*>(a.V + static_cast<size_t>(i) * kVFloats + n0);
#pragma unroll
            for (int k = 0; k < 3; ++k) dst[32 * k] = make_float4(vo[k][0], vo[k][1], vo[k][2], vo[k][3]);
            float4* mz = reinterpret_cast<float4*>(a.M_zero + static_cast<size_t>(i) * kMFloats + n0);
#pragma unroll
            for (int k = 0; k < 4; ++k) mz[32 * k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    // largest deposit bound per model over the previous layout's particle blocks
    const uint32_t Ap = a.lc_prev->A;
    float bmax[kMaxModels];
#pragma unroll
    for (int m = 0; m < kMaxModels; ++m) {
        bmax[m] = 0.0f;
        if (m < a.nmodels)
            for (uint32_t j = wg * kThreads + threadIdx.x; j < Ap; j += ng * kThreads)
                bmax[m] = fmaxf(bmax[m], a.blkmax[m][j]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
#pragma unroll
        for (int m = 0; m < kMaxModels; ++m)
            if (m < a.nmodels) bmax[m] = fmaxf(bmax[m], __shfl_xor(bmax[m], o, 64));
    }
    // one atomic per workgroup and value at most: the per-wave atomics on one word serialise
    __shared__ float sMax[1 + kMaxModels][kThreads / 64];
    if (LaneId() == 0) {
        sMax[0][threadIdx.x >> 6] = vmax;
#pragma unroll
        for (int m = 0; m < kMaxModels; ++m) sMax[1 + m][threadIdx.x >> 6] = bmax[m];
    }
    __syncthreads();
    if (threadIdx.x <= static_cast<uint32_t>(a.nmodels)) {
        float mx = 0.0f;
        for (int w = 0; w < kThreads / 64; ++w) mx = fmaxf(mx, sMax[threadIdx.x][w]);
        MaxBits(threadIdx.x == 0 ? &a.st->vmax2_bits[a.slot] : &a.st->dep_bits[a.slot][threadIdx.x - 1], mx);
    }
}
template <bool kCustom>
__global__ __launch_bounds__(kThreads) void GridUpdateKernel(GridUpdateArgs a) {
    GridUpdateBody<kCustom>(a, blockIdx.x, gridDim.x);
}

// ============================================================== G2P2G
// Fused G2P / constitutive update / P2G per particle block (mpm_algorithm.cuh:1365-1845):
// one workgroup per particle block; the 2x2x2-block velocity halo and the P2G
// accumulator live in LDS as 8^3 node windows of both sub-grids; each non-zero LDS
// node is flushed once with a global float atomic.  The kernel also emits each
// particle's block move (direction) and its deterministic rank among the block's
// particles with the same move, for the counting-sort rebin.
constexpr int kWin = 512;  // 8^3 nodes of one sub-grid window

// LDS windows of a particle block (node coordinates relative to the block's first
// node, 0..7 over the block and its +1 neighbours).  Stencil bases lie in [1,6] for
// grid 0 and [0,6] for grid 1 (window check), so grid 0 uses nodes [1,7]^3 and grid 1
// nodes [0,7]^3: 343 + 512 nodes per channel instead of 2 x 512.
constexpr int kM0Win = 343;              // 7^3
constexpr int kMWin = kM0Win + 512;      // both sub-grids, one channel
// Halo entry (block o of the 2x2x2 neighbourhood b + {0,1}^3, bit 2 = x, bit 1 = y,
// bit 0 = z; node nd = g*64 + x*16 + y*4 + z of its 128) -> window entry, or -1 for a
// grid-0 node outside [1,7]^3.  The halo load and the flush walk the neighbourhood block
// by block, so a wavefront reads or updates 64 consecutive nodes of one grid block.
__device__ __forceinline__ int WindowOfHalo(int o, int nd) {
    const int lx = ((o >> 2) & 1) * 4 + ((nd >> 4) & 3);
    const int ly = ((o >> 1) & 1) * 4 + ((nd >> 2) & 3);
    const int lz = (o & 1) * 4 + (nd & 3);
    if (nd >= 64) return kM0Win + lx * 64 + ly * 8 + lz;
    if (lx == 0 || ly == 0 || lz == 0) return -1;
    return (lx - 1) * 49 + (ly - 1) * 7 + (lz - 1);
}

// P2G accumulator in LDS: 64-bit fixed point.  On gfx950 a wavefront's ds_add_f32
// runs lane by lane (about 190 cycles whatever the addresses), ds_add_f64 takes about
// 16 cycles and ds_add_u64 about 8, independent of the number of active lanes, so the
// deposit is an integer add: value v goes in as round(v 2^s).  Integer sums are exact,
// so a node's LDS total does not depend on the order of the atomics; the flush rounds
// it once to FP32.
//   mass      deposits are weights w in [0, 1] (the mass is a per-model constant), at
//             s = kq = min(50, 62 - ceil(log2 n)) for a block of n particles: a deposit
//             is below 2^51 and a node total at most n 2^kq <= 2^62.
//   momentum  s = kq - e with one e per launch, read before the first deposit (a
//             workgroup-wide agreement on a scale inside the particle loop would need a
//             barrier between the constitutive update and the deposit, which serialises
//             the workgroup's VALU and LDS phases and costs more than the FP64 atomics):
//             2^e = 2 max(32 b_prev, 8 m vmax), b_prev the largest deposit bound of the
//             previous substep and vmax the grid's current max speed.  A particle whose
//             bound b = max_r (|m v_r| + sum_c |dx K_rc|) (|w (m v + K (x_i - x_p))_r| <= b)
//             reaches 2^(e-1) deposits its momentum with FP32 global atomics instead, so a
//             sudden growth of the deposits is handled exactly, only more slowly.  The
//             momentum resolution 2^(e - kq) is about 1e-13 of the largest deposit.
// The mass resolution must stay at least as fine as the momentum's relative to m vmax:
// at a node that only a stencil tail reaches (w ~ 1e-12) a coarser mass makes p / m, and
// so the grid's max speed that sets the CFL dt, wrong by a large factor.
// Conversion: for an integer-valued |x| < 2^51 the FP64 sum x + 1.5 2^52 has a fixed
// exponent and holds x in its mantissa, so x is that sum's bit pattern minus the constant
// B = 0x4338 << 48 (whose low word is zero: one 32-bit subtract).  The momentum terms are
// scaled by 2^s in FP32 once per particle (a power of two: exact) and converted to FP64
// once per particle and sub-grid; each node's momenta are formed by FP64 adds, the weight
// is converted once per node, and a deposit is one FP64 fma (w p exact in FP64, rounded
// once) and one subtract: 2 VALU instructions per deposit.
// e >= -64 keeps every scale 2^(kq - e) and its inverse a normal FP32 number.
using LdsAcc = long long;
constexpr int kMinExp = -64;
constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr unsigned long long kMagicBits = 0x4338000000000000ull;
__device__ __forceinline__ void LdsAddFixed(LdsAcc* p, double d) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p),
              static_cast<unsigned long long>(__double_as_longlong(d)) - kMagicBits);
}

// Deposit of one particle into both sub-grid windows: node momentum pm - sum_c kc_c t_c
// + sum_c o_c kc_c at stencil offset o (pm = m v, kc_c = dx K e_c), with pm and kc
// already multiplied by the momentum scale 2^(kq - e); weights at the mass scale wscale.
// SLOW: the momentum goes straight to the grid M of the halo blocks, scaled back by
// inv_qs.
template <bool SLOW>
__device__ __forceinline__ void DepositP2G(LdsAcc* sM, const Stencil32 (&sn)[2], const float (&pm)[3],
                                           const float (&kc)[3][3], double wscale, float inv_qs,
                                           const uint32_t* sHalo, float* M) {
    double kd[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) kd[c][r] = static_cast<double>(kc[c][r]);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        const Stencil32& sg = sn[g];
        // grid 0: nodes [1,7]^3 (7 wide), grid 1: nodes [0,7]^3 (8 wide)
        const int s1 = g ? 8 : 7, s2 = g ? 64 : 49, lo = g ? 0 : 1;
        LdsAcc* base = sM + (g ? kM0Win : 0) + (sg.c[0] - lo) * s2 + (sg.c[1] - lo) * s1 + (sg.c[2] - lo);
        float pa[3];
        double pad[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            pa[r] = pm[r] - (kc[0][r] * sg.t[0] + kc[1][r] * sg.t[1] + kc[2][r] * sg.t[2]);
            pad[r] = static_cast<double>(pa[r]);
        }
        float wxy[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            wxy[q] = ((q >> 1) ? 1.0f - sg.w0[0] : sg.w0[0]) * ((q & 1) ? 1.0f - sg.w0[1] : sg.w0[1]);
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
            const float w = wxy[o >> 1] * (oz ? 1.0f - sg.w0[2] : sg.w0[2]);
            const double wd = static_cast<double>(w);
            const int off = ox * s2 + oy * s1 + oz;
            LdsAddFixed(base + off, fma(wd, wscale, kMagic));
            uint32_t bi = kEmpty;
            int node = 0;
            if (SLOW) {
                // node coordinates relative to the block's first node: 0..7
                const int lx = sg.c[0] + ox, ly = sg.c[1] + oy, lz = sg.c[2] + oz;
                bi = sHalo[((lx >> 2) << 2) | ((ly >> 2) << 1) | (lz >> 2)];
                node = g * 64 + (lx & 3) * 16 + (ly & 3) * 4 + (lz & 3);
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (!SLOW) {
                    double pr = pad[r];
                    if (ox) pr += kd[0][r];
                    if (oy) pr += kd[1][r];
                    if (oz) pr += kd[2][r];
                    LdsAddFixed(base + (r + 1) * kMWin + off, fma(wd, pr, kMagic));
                } else if (bi != kEmpty) {
                    float pr = pa[r];
                    if (ox) pr += kc[0][r];
                    if (oy) pr += kc[1][r];
                    if (oz) pr += kc[2][r];
                    unsafeAtomicAdd(M + static_cast<size_t>(bi) * kMFloats + (r + 1) * 128 + node, w * pr * inv_qs);
                }
            }
        }
    }
}

// Particles are taken in 64-particle chunks, one wavefront per chunk (chunk c of a block
// goes to wave c mod waves): with a few hundred particles per block, 256-particle
// chunks would leave about a quarter of the lanes idle.  The rank of a particle among
// the block's particles with the same move (its counting-sort position) is fixed in
// rounds of kRoundChunks chunks: each chunk stores its lanes' move and rank within the
// chunk and its per-move counts in LDS; after the round, one thread per move turns the
// counts into exclusive prefixes over the chunks, and the codes are written.
constexpr int kRoundChunks = 12;
constexpr uint32_t kRoundParticles = kRoundChunks * 64;
constexpr int kHistStride = 28;  // per chunk: 27 move counts (then prefixes), padded

// 4 waves per SIMD: LDS (40.1 KiB per workgroup) allows 4 workgroups per CU; this caps
// the VGPRs at 128
template <int KIND, int THREADS>
__device__ __forceinline__ void G2P2GBody(const G2P2GArgs& a, uint32_t wg, uint32_t ng) {
    constexpr uint32_t kWaves = THREADS / 64;
    __shared__ float sV[kVChannels * kMWin];
    __shared__ __align__(16) LdsAcc sM[kMChannels * kMWin];
    __shared__ uint16_t sCode[kRoundParticles];               // move | rank in the chunk << 5
    __shared__ uint16_t sHist[kRoundChunks * kHistStride];    // per chunk and move
    __shared__ uint32_t sHalo[8];
    __shared__ uint32_t sRun[32];   // particles of the block per move before this round
    __shared__ uint32_t sTot[32];   // particles per move in this round
    __shared__ uint32_t sBlkMax;    // largest deposit bound of the block (FP32 bits)
    __shared__ uint32_t sGlobal;    // particles of the block deposited with global atomics

    const uint32_t A = a.lc->A, Nb = a.lc->Nb;
    const uint32_t tid = threadIdx.x, lane = LaneId(), wave = tid >> 6;
    const float dx = a.dom.dx, inv_dx = a.dom.inv_dx;
    const size_t N = a.n;
    // momentum fixed-point exponent of this launch (LdsAcc): deposits with a bound below
    // 2^(mexp-1) go to LDS
    int mexp = kMinExp;
    {
        const float hint = __uint_as_float(a.st->dep_bits[a.slot][a.model]);
        const float vmax = sqrtf(__uint_as_float(a.st->vmax2_bits[a.slot]));
        float ref = fmaxf(32.0f * hint, 8.0f * a.mat.mass * vmax);
        if (!(ref <= 1.0e37f)) ref = 1.0e37f;  // a non-finite grid is already flagged
        if (ref > 0.0f) mexp = max(kMinExp, static_cast<int>((__float_as_uint(ref) >> 23) & 0xFFu) - 127 + 2);
    }
    const float fast_limit = __int_as_float((mexp - 1 + 127) << 23);  // 2^(mexp-1)

    // The accumulator and the move counts are cleared once here; the flush and the
    // prefix pass leave them zero for the next block.  16-byte clears (the size is even).
    for (int i = 2 * tid; i < kMChannels * kMWin; i += 2 * THREADS)
        *reinterpret_cast<uint4*>(sM + i) = make_uint4(0u, 0u, 0u, 0u);
    for (int i = tid; i < kRoundChunks * kHistStride; i += THREADS) sHist[i] = 0;

    // Round-robin over the block list: particle counts per block vary by orders of
    // magnitude along the list (a model occupies only part of it), so contiguous
    // per-XCD ranges leave most XCDs idle; interleaving balances them.
    for (uint32_t j = wg; j < A; j += ng) {
        const uint32_t beg = a.off[j], end = a.off[j + 1];
        if (beg == end) {
            // no particle of this model here: only the (zero) move counts are written
            if (tid < 27) a.cnt[static_cast<size_t>(j) * kDirections + tid] = 0u;
            if (tid == 0) a.blkmax[j] = 0.0f;
            continue;
        }
        int bb[3];
        CoordsOf(a.dom, a.keys[j], bb[0], bb[1], bb[2]);
        if (tid < 8) {
            const int bx = bb[0] + ((tid >> 2) & 1), by = bb[1] + ((tid >> 1) & 1), bz = bb[2] + (tid & 1);
            uint32_t bi = kEmpty;
            if (InDomain(a.dom, bx, by, bz)) bi = a.table[KeyOf(a.dom, bx, by, bz)];
            if (bi != kEmpty && bi >= Nb) bi = kEmpty;
            sHalo[tid] = bi;
        }
        if (tid < 32) sRun[tid] = 0;
        if (tid == 0) {
            sBlkMax = 0u;
            sGlobal = 0u;
        }
        // fixed-point scales of this block (LdsAcc): weights 2^kq, momentum 2^(kq - mexp)
        const uint32_t nblk = end - beg;
        const int kq = min(50, 62 - (nblk > 1u ? 32 - __clz(nblk - 1u) : 0));
        // powers of two built from their bit patterns: integer (scalar) arithmetic, so the
        // block-uniform scales stay in scalar registers across the particle loop
        const double wscale = __longlong_as_double(static_cast<long long>(kq + 1023) << 52);  // 2^kq
        const float qs = __int_as_float((kq - mexp + 127) << 23);                             // 2^(kq - mexp)
        const float inv_qs = __int_as_float((mexp - kq + 127) << 23);
        float tmax = 0.0f;  // largest deposit bound of this thread's particles
        __syncthreads();
        // velocity halo in the same windows as the P2G accumulator (3 channels): a thread
        // takes a z column of 4 nodes of a grid block (16-byte loads), which fills 4
        // consecutive window entries; in grid 0 the first of them lies outside the window
        // when the block is at z offset 0 (lz = 0), and a column at lx = 0 or ly = 0
        // lies outside altogether
        for (int e = tid; e < 8 * kNodesPerBlock / 4; e += THREADS) {
            const int o = e >> 5, nd0 = (e & 31) * 4;
            const int r3 = WindowOfHalo(o, nd0 + 3);
            if (r3 < 0) continue;
            const uint32_t bi = sHalo[o];
            const float4* src = reinterpret_cast<const float4*>(a.V + static_cast<size_t>(bi) * kVFloats + nd0);
            const bool first = nd0 >= 64 || (o & 1);
#pragma unroll
            for (int ch = 0; ch < kVChannels; ++ch) {
                const float4 v = bi != kEmpty ? src[32 * ch] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                const int r = ch * kMWin + r3;
                if (first) sV[r - 3] = v.x;
                sV[r - 2] = v.y;
                sV[r - 1] = v.z;
                sV[r] = v.w;
            }
        }
        if (tid == 0) {
            bool miss = false;
            for (int o = 0; o < 8; ++o) miss |= sHalo[o] == kEmpty;
            if (miss) SetFlag(a.st, kFlagMissingBlock);
        }
        __syncthreads();

        for (uint32_t r0 = beg; r0 < end; r0 += kRoundParticles) {
            const uint32_t rn = min(end - r0, kRoundParticles);
            const uint32_t nch = (rn + 63u) >> 6;
            const bool last = r0 + rn == end;
            // the ordering word of the wave's next chunk is loaded a chunk ahead, so the
            // state loads of a chunk wait for one memory latency instead of two
            uint32_t pe_next = 0u;
            if (wave < nch && wave * 64u + lane < rn) pe_next = a.perm[r0 + wave * 64u + lane];
            for (uint32_t c = wave; c < nch; c += kWaves) {
                const uint32_t t = r0 + c * 64u + lane;
                const bool active = t < r0 + rn;
                const uint32_t pe = pe_next;
                if (c + kWaves < nch && (c + kWaves) * 64u + lane < rn) pe_next = a.perm[t + kWaves * 64u];
                uint32_t dir = 31u;
                if (active) {
                    uint32_t s = pe & kSlotMask;
                    int mv = static_cast<int>(pe >> kMoveShift);
                    if (s >= a.n || mv >= kDirections) {
                        SetFlag(a.st, kFlagMissingBlock);
                        s = 0;
                        mv = 13;
                    }
                    if (a.ids_out) a.ids_out[t] = a.ids_in[s];
                    // position relative to this block's origin: the stored value is relative
                    // to the block that wrote it, which sits at this block - move
                    const int mov[3] = {mv / 9 - 1, (mv / 3) % 3 - 1, mv % 3 - 1};
                    float x[3], F[9], q = 0.0f, logJp = 0.0f;  // F holds E = F - I; q = J - 1 (fluid)
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        x[k] = __fsub_rn(a.st_in[k * N + s], __fmul_rn(static_cast<float>(4 * mov[k]), dx));
                    if (KIND == kFluid) {
                        q = a.st_in[3 * N + s];
                    } else {
#pragma unroll
                        for (int k = 0; k < 9; ++k) F[k] = a.st_in[(3 + k) * N + s];
                        if (KIND == kDruckerPrager || KIND == kCamClay) logJp = a.st_in[12 * N + s];
                    }
                    float lam = a.mat.lambda, mu = a.mat.mu;
                    if (a.pp_attr) {
                        lam = a.st_in[a.pp_attr * N + s];
                        mu = a.st_in[(a.pp_attr + 1) * N + s];
                    }
                    const float rigid = a.rigid_attr ? a.st_in[a.rigid_attr * N + s] : 0.0f;
                    // ---------------- G2P at x (both sub-grids, 16 nodes)
                    float xs[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) xs[k] = __fmul_rn(x[k], inv_dx);
                    float v[3] = {0.f, 0.f, 0.f}, B[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    float D[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    bool bad = false;
                    Stencil32 sp[2];
                    MakeStencilPair(xs, sp[0], sp[1]);
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        const Stencil32& sg = sp[g];
                        int l[3];
                        float m1[3];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const int lc = sg.c[k];
                            // the binned cell (2..5 in the frame of the block that binned the
                            // particle) can move by one in this block's frame through FP32
                            // rounding of + 0.25, so the bases lie in [1,6] (grid 0), [0,6] (grid 1)
                            const int lo = g ? 0 : 1;
                            bad |= lc < lo || lc > 6;
                            l[k] = min(max(lc, lo), 6) - lo;
                            // D closed form per axis: second moment on the diagonal, products of
                            // first moments off the diagonal (units of dx^2)
                            const float w0 = sg.w0[k], w1 = 1.0f - w0, tt = sg.t[k];
                            m1[k] = w1 - tt;
                            D[k * 4] += w0 * tt * tt + w1 * (1.0f - tt) * (1.0f - tt);
                        }
                        D[1] += m1[0] * m1[1];
                        D[2] += m1[0] * m1[2];
                        D[5] += m1[1] * m1[2];
                        // B = sum_i v_i (x_i - x_p)^T w_i with x_i - x_p = (o - t) dx per axis:
                        // B_rc = dx (S_rc - vg_r t_c), S_rc = sum over the nodes with o_c = 1 of w v_r.
                        // The weights are a tensor product, so vg and S are contracted axis by
                        // axis (z, then y, then x): 22 instructions per channel.
                        const int s1 = g ? 8 : 7, s2 = g ? 64 : 49;
                        const float* base = sV + (g ? kM0Win : 0) + l[0] * s2 + l[1] * s1 + l[2];
                        const float wx0 = sg.w0[0], wy0 = sg.w0[1], wz0 = sg.w0[2];
                        const float wx1 = 1.0f - wx0, wy1 = 1.0f - wy0, wz1 = 1.0f - wz0;
                        float vg[3], S[9];
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            const float* vr = base + r * kMWin;
                            float T[4], U[4];  // per (x, y) column: sum over z, z-moment
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const int off = (q >> 1) * s2 + (q & 1) * s1;
                                U[q] = wz1 * vr[off + 1];
                                T[q] = fmaf(wz0, vr[off], U[q]);
                            }
                            float P[2], Q[2], Z[2];  // per x row: sum over y and z, y-moment, z-moment
#pragma unroll
                            for (int ox = 0; ox < 2; ++ox) {
                                Q[ox] = wy1 * T[ox * 2 + 1];
                                P[ox] = fmaf(wy0, T[ox * 2], Q[ox]);
                                Z[ox] = fmaf(wy0, U[ox * 2], wy1 * U[ox * 2 + 1]);
                            }
                            S[r * 3 + 0] = wx1 * P[1];
                            vg[r] = fmaf(wx0, P[0], S[r * 3 + 0]);
                            S[r * 3 + 1] = fmaf(wx0, Q[0], wx1 * Q[1]);
                            S[r * 3 + 2] = fmaf(wx0, Z[0], wx1 * Z[1]);
                        }
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            v[r] += vg[r];
#pragma unroll
                            for (int c = 0; c < 3; ++c) B[r * 3 + c] += dx * (S[r * 3 + c] - vg[r] * sg.t[c]);
                        }
                    }
                    if (bad) SetFlag(a.st, kFlagOutOfWindow);
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
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] *= 0.5f;
                    if (rigid != 0.0f) {
                        // a rigid particle takes nothing from the grid and moves at the
                        // prescribed velocity (mpm_algorithm.cuh:1609-1617)
#pragma unroll
                        for (int k = 0; k < 9; ++k) C[k] = 0.0f;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            x[k] += a.dt * a.rigid_v[k];
                            v[k] = 0.0f;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < 3; ++k) x[k] += a.dt * v[k];
                    }
                    // ---------------- deformation update and stress
                    float tau[9];
                    if (KIND == kFluid) {
                        StressFluid(a.mat, C, a.dt, q, tau);
                    } else {
                        // F' = (I + dt C) F  <=>  E' = E + dt C (I + E), E = F - I
                        float CE[9];
                        MatMul3(C, F, CE);
#pragma unroll
                        for (int k = 0; k < 9; ++k) F[k] += a.dt * (C[k] + CE[k]);
                        if (KIND == kFixedCorotated) StressFixedCorotated(lam, mu, F, tau);
                        else if (KIND == kNeoHookean) StressNeoHookean(lam, mu, F, tau);
                        else if (KIND == kDruckerPrager) StressDruckerPrager(lam, mu, a.mat.sin_phi, F, logJp, tau);
                        else if (KIND == kVonMises) StressVonMises(lam, mu, a.mat.yield, F, tau);
                        else StressCamClay(a.mat, lam, mu, F, logJp, tau);
                    }
#pragma unroll
                    for (int k = 0; k < 3; ++k) a.st_out[k * N + t] = x[k];
                    if (KIND == kFluid) {
                        a.st_out[3 * N + t] = q;
                    } else {
#pragma unroll
                        for (int k = 0; k < 9; ++k) a.st_out[(3 + k) * N + t] = F[k];
                        if (KIND == kDruckerPrager || KIND == kCamClay) a.st_out[12 * N + t] = logJp;
                    }
                    if (a.pp_attr) {
                        a.st_out[a.pp_attr * N + t] = lam;
                        a.st_out[(a.pp_attr + 1) * N + t] = mu;
                    }
                    if (a.rigid_attr) a.st_out[a.rigid_attr * N + t] = rigid;
                    // ---------------- block move for the rebin (FP32 grid-0 binning)
                    int dlt[3];
                    bool clamped = false;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const int dd = FloorDiv4(Cell0(x[k], inv_dx) - 2);
                        clamped |= dd < -1 || dd > 1;
                        dlt[k] = min(max(dd, -1), 1);
                    }
                    if (clamped) SetFlag(a.st, kFlagOutOfWindow);
                    if (!InDomain(a.dom, bb[0] + dlt[0], bb[1] + dlt[1], bb[2] + dlt[2])) {
                        SetFlag(a.st, kFlagOutOfDomain);
                        dlt[0] = dlt[1] = dlt[2] = 0;
                    }
                    dir = static_cast<uint32_t>((dlt[0] + 1) * 9 + (dlt[1] + 1) * 3 + (dlt[2] + 1));
                    // ---------------- P2G at the new x
#pragma unroll
                    for (int k = 0; k < 3; ++k) xs[k] = __fmul_rn(x[k], inv_dx);
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
                    if (!inwin) {
                        SetFlag(a.st, kFlagOutOfWindow);
                    } else {
                        // D^ = (0.5 D')^-1 = 2 D'^-1; K = m C - nextDt V D^ tau
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
                        const float sv = a.next_dt * a.mat.volume;
#pragma unroll
                        for (int k = 0; k < 9; ++k) K[k] = C[k] * a.mat.mass - sv * DT[k];
                        float pm[3], kc[3][3];  // m v and dx K e_c
#pragma unroll
                        for (int r = 0; r < 3; ++r) pm[r] = a.mat.mass * v[r];
                        // node momentum m v + K (x_i - x_p), x_i - x_p = (o - t) dx:
                        // = a + sum_c o_c kc_c with a = m v - dx K t and kc_c = dx K e_c
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int r = 0; r < 3; ++r) kc[c][r] = dx * K[r * 3 + c];
                        float bound = 0.0f, bsum = fabsf(x[0]) + fabsf(x[1]) + fabsf(x[2]);
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            const float b = fabsf(pm[r]) + fabsf(kc[0][r]) + fabsf(kc[1][r]) + fabsf(kc[2][r]);
                            bound = fmaxf(bound, b);
                            bsum += b;
                        }
                        // a non-finite deposit has no fixed-point value: report it (the sum is
                        // NaN or inf whenever an input is; fmaxf would drop a NaN)
                        if (!(bsum <= 3.0e38f)) {
                            SetFlag(a.st, kFlagNonFinite);
                        } else {
                            tmax = fmaxf(tmax, bound);
#pragma unroll
                            for (int r = 0; r < 3; ++r) {
                                pm[r] *= qs;
#pragma unroll
                                for (int c = 0; c < 3; ++c) kc[c][r] *= qs;
                            }
                            if (bound < fast_limit) {
                                DepositP2G<false>(sM, sn, pm, kc, wscale, inv_qs, sHalo, a.M);
                            } else {
                                DepositP2G<true>(sM, sn, pm, kc, wscale, inv_qs, sHalo, a.M);
                                const uint64_t on = __ballot(1);
                                if (lane == static_cast<uint32_t>(__builtin_ctzll(on)))
                                    atomicAdd(&sGlobal, static_cast<uint32_t>(__popcll(on)));
                            }
                        }
                    }
                }
                // ---------------- rank among the chunk's particles with the same move
                const uint64_t peers = MatchPeers<5>(dir);
                const uint32_t rank = LanesBelow(peers);
                if (active) {
                    sCode[t - r0] = static_cast<uint16_t>(dir | (rank << 5));
                    if (rank == 0) sHist[c * kHistStride + dir] = static_cast<uint16_t>(__popcll(peers));
                }
            }
            // largest deposit bound of the block (for the next substep's scale): one LDS
            // max per wave instruction (non-negative floats order as their bits)
            if (last && tmax > 0.0f) atomicMax(&sBlkMax, __float_as_uint(tmax));
            __syncthreads();
            // exclusive prefix of each move's counts over the round's chunks
            if (tid < kDirections) {
                uint32_t run = 0;
                for (uint32_t c = 0; c < nch; ++c) {
                    const uint32_t h = sHist[c * kHistStride + tid];
                    sHist[c * kHistStride + tid] = static_cast<uint16_t>(run);
                    run += h;
                }
                sTot[tid] = run;
            }
            __syncthreads();
            for (uint32_t i = tid; i < rn; i += THREADS) {
                const uint32_t v = sCode[i], dir = v & 31u;
                const uint32_t r = sRun[dir] + sHist[(i >> 6) * kHistStride + dir] + (v >> 5);
                a.code[r0 + i] = dir | (r << 5);
            }
            if (last) {
                // ---------------- flush: one global atomic per non-zero node and channel,
                // block by block of the neighbourhood (fixed point -> FP32 through an exact
                // double, 2^-kq m for the weights, 2^(mexp-kq) for the momentum); every entry
                // read non-zero is cleared for the next block
                const double unit[2] = {static_cast<double>(__int_as_float((127 - kq) << 23)) * a.mat.mass,
                                        static_cast<double>(__int_as_float((mexp - kq + 127) << 23))};
                for (int e = tid; e < 8 * kNodesPerBlock; e += THREADS) {
                    const int o = e >> 7, nd = e & 127;
                    const int r = WindowOfHalo(o, nd);
                    if (r < 0) continue;
                    LdsAcc qv[kMChannels];
                    bool any = false;
#pragma unroll
                    for (int ch = 0; ch < kMChannels; ++ch) {
                        qv[ch] = sM[ch * kMWin + r];
                        any |= qv[ch] != 0;
                    }
                    if (!any) continue;
#pragma unroll
                    for (int ch = 0; ch < kMChannels; ++ch) sM[ch * kMWin + r] = 0;
                    const uint32_t bi = sHalo[o];
                    if (bi == kEmpty) continue;
                    float* dst = a.M + static_cast<size_t>(bi) * kMFloats + nd;
#pragma unroll
                    for (int ch = 0; ch < kMChannels; ++ch) {
                        if (qv[ch] == 0) continue;
                        // |q| < 2^63: hi * 2^32 + lo is exact to double rounding (2^-53)
                        const double d = fma(static_cast<double>(static_cast<int>(qv[ch] >> 32)), 4294967296.0,
                                             static_cast<double>(static_cast<uint32_t>(qv[ch])));
                        const float val = static_cast<float>(d * unit[ch == 0 ? 0 : 1]);
                        if (val != 0.0f) unsafeAtomicAdd(dst + ch * 128, val);
                    }
                }
            }
            __syncthreads();
            // next round (or block): move totals carried, counts cleared
            for (uint32_t i = tid; i < nch * kHistStride; i += THREADS) sHist[i] = 0;
            if (tid < kDirections) sRun[tid] += sTot[tid];
            if (!last) __syncthreads();
        }
        if (tid < 27) a.cnt[static_cast<size_t>(j) * kDirections + tid] = sRun[tid];
        if (tid == 0) {
            a.blkmax[j] = __uint_as_float(sBlkMax);
            if (sGlobal) atomicAdd(&a.st->p2g_global, sGlobal);
        }
    }
}
// The body reads the arguments through the kernel-argument segment pointer (the explicit
// argument a sits at offset 0), hidden from the optimizer by an empty asm: the fields are
// then reloaded with scalar loads where they are needed instead of being held in SGPRs.
// Read from a directly, the plastic kinds spilled 13 SGPRs and 2-4 VGPRs; this way none.
template <int KIND, int THREADS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void G2P2GKernel(G2P2GArgs a) {
    using CT = const __attribute__((address_space(4))) G2P2GArgs;
    CT* p = (CT*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    G2P2GBody<KIND, THREADS>(*(const G2P2GArgs*)p, blockIdx.x, gridDim.x);
}

// ============================================================== rebin
// Particles entering each exterior block e: sum over the 27 possible source blocks
// (e - delta, particle blocks of the current layout) of their count for move delta.
// A 32-lane group per exterior block, lane = move (27 independent lookups instead of a
// dependent chain of 27 per thread); a workgroup owns a tile of kCountTile consecutive
// exterior blocks and writes the tile's sums (occupied blocks, entering particles per
// model) for the scan kernel.  A block's particles are ordered by the move that brought
// them in, so the group also writes base[m][e][d], the particles of model m entering e
// with a move before d (an exclusive scan over the lanes): the first destination of the
// particles of source e - d with move d is e's offset in the next layout plus that count.
// The same launch clears the next layout's table entries (its keys are still those of
// two substeps ago) and, in workgroup 0, the tile counts of the extension marking.
__device__ __forceinline__ void RebinCountBody(const RebinCountArgs& a, uint32_t wg, uint32_t ng) {
    __shared__ uint32_t sG[kCountTile][kTileSums];
    const DomainDev& d = a.d;
    const uint32_t* keys = a.keys;
    const uint32_t* table = a.table;
    const RebinModels& rm = a.rm;
    uint32_t* tsum = a.tsum;
    const uint32_t cap = a.cap;
    const LayoutCounts* lc = a.lc;
    const uint32_t exn = min(a.lc_next->Ex, cap);
    for (uint32_t i = wg * kCountThreads + threadIdx.x; i < exn; i += ng * kCountThreads)
        a.table_next[a.keys_next[i]] = kEmpty;
    if (wg == 0)
        for (uint32_t i = threadIdx.x; i < a.ext_cnt_words; i += kCountThreads) a.ext_cnt[i] = 0u;
    const uint32_t A = lc->A, Ex = min(lc->Ex, cap);
    const uint32_t lane = threadIdx.x & 31u, grp = threadIdx.x >> 5;
    const int nq = 1 + rm.nmodels;
    const uint32_t ntiles = (Ex + kCountTile - 1) / kCountTile;
    for (uint32_t tile = wg; tile < ntiles; tile += ng) {
        const uint32_t e = tile * kCountTile + grp;
        if (e < Ex) {
            int bx, by, bz;
            CoordsOf(d, keys[e], bx, by, bz);
            uint32_t js = kEmpty;
            if (lane < static_cast<uint32_t>(kDirections)) {
                int ox, oy, oz;
                DirOffset(static_cast<int>(lane), ox, oy, oz);
                const int sx = bx - ox, sy = by - oy, sz = bz - oz;
                if (InDomain(d, sx, sy, sz)) js = table[KeyOf(d, sx, sy, sz)];
            }
            uint32_t any = 0;
            for (int m = 0; m < rm.nmodels; ++m) {
                const uint32_t v = js < A ? rm.cnt[m][static_cast<size_t>(js) * kDirections + lane] : 0u;  // also kEmpty
                // inclusive scan over the group's lanes (lane = move)
                uint32_t inc = v;
#pragma unroll
                for (int o = 1; o < 32; o <<= 1) {
                    const uint32_t u = __shfl_up(inc, o, 32);
                    if (lane >= static_cast<uint32_t>(o)) inc += u;
                }
                if (lane < static_cast<uint32_t>(kDirections))
                    rm.base[m][static_cast<size_t>(e) * kDirections + lane] = inc - v;
                const uint32_t total = __shfl(inc, 31, 32);
                if (lane == 0) {
                    rm.cnt_e[m][e] = total;
                    sG[grp][1 + m] = total;
                }
                any |= total;
            }
            if (lane == 0) sG[grp][0] = any ? 1u : 0u;
        } else if (lane < static_cast<uint32_t>(nq)) {
            sG[grp][lane] = 0u;
        }
        __syncthreads();
        if (threadIdx.x < static_cast<uint32_t>(nq)) {
            uint32_t sum = 0;
            for (int g = 0; g < kCountTile; ++g) sum += sG[g][threadIdx.x];
            tsum[static_cast<size_t>(tile) * kTileSums + threadIdx.x] = sum;
        }
        __syncthreads();
    }
}
__global__ __launch_bounds__(kCountThreads) void RebinCountKernel(RebinCountArgs a) {
    RebinCountBody(a, blockIdx.x, gridDim.x);
}

// Next layout's particle blocks and particle offsets in one pass: workgroup w owns
// exterior blocks [w, w+1) * kRebinScanTile of the current layout, takes the sums of
// the count tiles before it, and scans the occupied flags (new list index of each
// occupied block; its key goes to keys_next and the table) and each model's entering
// counts (offsets of the next layout).  Unoccupied blocks enter no particle of any
// model, so the scan over all exterior blocks equals the scan over the compacted
// list.  Workgroup 0 also sets A of the next layout and the end offsets (= n).
// The launch covers the capacity: ng * kRebinScanTile >= cap.  Thread i owns
// exterior blocks [base + i, base + i + 1) * kRebinScanItems.
__device__ __forceinline__ void RebinScanBody(const RebinScanArgs& a, uint32_t wg, uint32_t) {
    __shared__ uint32_t sh[kThreads / 64];
    __shared__ uint32_t red[kTileSums];
    __shared__ uint32_t sPre[kTileSums];
    const uint32_t* keys = a.keys;
    const RebinModels& rm = a.rm;
    const uint32_t* tsum = a.tsum;
    uint32_t* keys_next = a.keys_next;
    uint32_t* table_next = a.table_next;
    const uint32_t cap = a.cap;
    const uint32_t Ex = min(a.lc->Ex, cap), tid = threadIdx.x;
    const int nq = 1 + rm.nmodels;
    const uint32_t base = wg * kRebinScanTile;
    if (base >= Ex && wg != 0) return;
    // sums of the count tiles before this tile (workgroup 0: of all tiles, the totals)
    const uint32_t ntc = (Ex + kCountTile - 1) / kCountTile;
    const uint32_t upto = wg == 0 ? ntc : min(ntc, base / kCountTile);
    // every global load of the workgroup is issued before the first barrier: the count
    // tiles before this tile, and the keys and entering counts of the thread's blocks
    // (a block is occupied iff a model's particles enter it)
    uint32_t v[kTileSums];
#pragma unroll
    for (int j = 0; j < kTileSums; ++j) v[j] = 0u;
    for (uint32_t t = tid; t < upto; t += kThreads) {
#pragma unroll
        for (int j = 0; j < kTileSums; ++j)
            if (j < nq) v[j] += tsum[static_cast<size_t>(t) * kTileSums + j];
    }
    const uint32_t e0 = base + tid * kRebinScanItems;
    uint32_t c[kMaxModels][kRebinScanItems], key[kRebinScanItems];
#pragma unroll
    for (int k = 0; k < kRebinScanItems; ++k) key[k] = e0 + k < Ex ? keys[e0 + k] : 0u;
#pragma unroll
    for (int m = 0; m < kMaxModels; ++m)
#pragma unroll
        for (int k = 0; k < kRebinScanItems; ++k) c[m][k] = m < rm.nmodels && e0 + k < Ex ? rm.cnt_e[m][e0 + k] : 0u;
    BlockSums<kTileSums>(v, nq, red);
    if (tid < static_cast<uint32_t>(nq)) sPre[tid] = wg == 0 ? 0u : red[tid];
    if (wg == 0 && tid == 0) {
        uint32_t an = red[0];
        if (an > cap) {
            SetFlag(a.st, kFlagCapacity);
            an = cap;
        }
        a.lc_next->A = an;
        for (int m = 0; m < rm.nmodels; ++m) rm.off_next[m][an] = rm.n[m];
    }
    __syncthreads();
    uint32_t omask = 0;
#pragma unroll
    for (int k = 0; k < kRebinScanItems; ++k) {
        uint32_t any = 0;
#pragma unroll
        for (int m = 0; m < kMaxModels; ++m) any |= c[m][k];
        if (any) omask |= 1u << k;
    }
    uint32_t tot;
    const uint32_t i0 = sPre[0] + BlockExclusiveScan(static_cast<uint32_t>(__popc(omask)), tot, sh);
    uint32_t i = i0;
#pragma unroll
    for (int k = 0; k < kRebinScanItems; ++k) {
        if (!((omask >> k) & 1u)) continue;
        if (i < cap) {
            keys_next[i] = key[k];
            table_next[key[k]] = i;
        }
        ++i;
    }
#pragma unroll
    for (int m = 0; m < kMaxModels; ++m) {
        if (m < rm.nmodels) {
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < kRebinScanItems; ++k) sum += c[m][k];
            uint32_t run = sPre[1 + m] + BlockExclusiveScan(sum, tot, sh);
            i = i0;
#pragma unroll
            for (int k = 0; k < kRebinScanItems; ++k) {
                if ((omask >> k) & 1u) {
                    if (i < cap) rm.off_next[m][i] = run;
                    rm.eoff[m][e0 + k] = run;
                    ++i;
                }
                run += c[m][k];
            }
        }
    }
}
__global__ __launch_bounds__(kThreads) void RebinScanKernel(RebinScanArgs a) { RebinScanBody(a, blockIdx.x, gridDim.x); }

// Particle permutation of the next layout: a wavefront per block j of the current
// layout, for every model.  Lane d < 27 finds the destination e = j + d (an exterior
// block of the current layout); per model, the first destination of move d is e's offset
// in the next layout (eoff) plus base[m][e][d] (the particles entering e with an earlier
// move), and the particle loop takes it by lane shuffle.  The destination lookups start
// after the first pass's move codes are requested, so the two chains of loads overlap.
// kScatterUnroll particles per lane are in flight (a block holds a few hundred
// particles per model, so most blocks take one pass of dependent loads: 512 per pass).
constexpr int kScatterUnroll = 8;
__device__ __forceinline__ void RebinScatterBody(const RebinScatterArgs& a, uint32_t wg, uint32_t ng) {
    const RebinModels& rm = a.rm;
    const uint32_t A = a.lc->A, cap = a.cap;
    const uint32_t lane = LaneId(), waves = kThreads / 64;
    for (uint32_t j = wg * waves + (threadIdx.x >> 6); j < A; j += ng * waves) {
        uint32_t e = kEmpty;
        bool have_e = false;  // wave-uniform
        for (int m = 0; m < rm.nmodels; ++m) {
            const uint32_t* off = rm.off_cur[m];
            const uint32_t beg = off[j], end = off[j + 1];
            if (beg >= end) continue;
            const uint32_t* code = rm.code[m];
            uint32_t* perm_next = rm.perm_next[m];
            const uint32_t n = rm.n[m];
            // a move without a destination (no particle takes it) keeps n: its positions drop
            uint32_t first = n;
            for (uint32_t t0 = beg; t0 < end; t0 += kScatterUnroll * 64) {
                uint32_t c[kScatterUnroll];
#pragma unroll
                for (int u = 0; u < kScatterUnroll; ++u) {
                    const uint32_t t = t0 + u * 64 + lane;
                    c[u] = t < end ? code[t] : 31u;
                }
                if (!have_e) {
                    if (lane < static_cast<uint32_t>(kDirections)) {
                        int bx, by, bz, ox, oy, oz;
                        CoordsOf(a.d, a.keys[j], bx, by, bz);
                        DirOffset(static_cast<int>(lane), ox, oy, oz);
                        bx += ox;
                        by += oy;
                        bz += oz;
                        if (InDomain(a.d, bx, by, bz)) e = a.table[KeyOf(a.d, bx, by, bz)];
                    }
                    have_e = true;
                }
                if (t0 == beg && e < cap)
                    first = rm.eoff[m][e] + rm.base[m][static_cast<size_t>(e) * kDirections + lane];
#pragma unroll
                for (int u = 0; u < kScatterUnroll; ++u) {
                    const uint32_t dir = c[u] & 31u;
                    const uint32_t f = __shfl(first, static_cast<int>(dir), 64);  // every lane active here
                    const uint32_t pos = f + (c[u] >> 5);
                    if (dir < static_cast<uint32_t>(kDirections) && pos < n)
                        perm_next[pos] = (t0 + u * 64 + lane) | (dir << kMoveShift);
                }
            }
        }
    }
}
__global__ __launch_bounds__(kThreads) void RebinScatterKernel(RebinScatterArgs a) {
    RebinScatterBody(a, blockIdx.x, gridDim.x);
}

// The scatter and the extension's mark pass side by side in one launch: a fork to a
// second stream and its join cost more than either kernel's tail.
__device__ __forceinline__ void ScatterMarkBody(const ScatterMarkArgs& a, uint32_t wg, uint32_t ng) {
    if (wg < a.mark_wgs) MarkExtensionBody(a.m, wg, a.mark_wgs);
    else RebinScatterBody(a.x, wg - a.mark_wgs, ng - a.mark_wgs);
}
__global__ __launch_bounds__(kThreads) void ScatterMarkKernel(ScatterMarkArgs a) {
    ScatterMarkBody(a, blockIdx.x, gridDim.x);
}

// rank of each element inside its run of equal keys (from run heads and their positions)
__global__ void RunRankKernel(const uint32_t* head, const uint32_t* ex, const uint32_t* start, uint32_t n,
                              uint32_t* rank) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rank[i] = i - start[ex[i] + head[i] - 1u];
}

uint32_t Blocks(uint64_t n, uint32_t threads, uint32_t maxb = 4096) {
    uint64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > maxb) b = maxb;
    return static_cast<uint32_t>(b);
}

}  // namespace

// ============================================================== launchers
uint32_t LaunchBlocks(uint64_t n) { return Blocks(n, kThreads); }
size_t ScanTempWords(size_t n_cap) { return (n_cap + kScanTile - 1) / kScanTile + 1; }
size_t RebinTileSumWords(uint32_t cap) { return (static_cast<size_t>(cap) / kCountTile + 1) * kTileSums; }

void ExclusiveScan(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_imm, uint32_t n_cap,
                   uint32_t* total, bool popcount, uint32_t* tmp, hipStream_t s) {
    const uint32_t tiles = static_cast<uint32_t>((static_cast<uint64_t>(n_cap) + kScanTile - 1) / kScanTile);
    const uint32_t grid = tiles ? tiles : 1;
    if (popcount) ScanReduceKernel<true><<<grid, kThreads, 0, s>>>(in, n_dev, n_imm, tmp);
    else ScanReduceKernel<false><<<grid, kThreads, 0, s>>>(in, n_dev, n_imm, tmp);
    ScanPartialsKernel<<<1, kThreads, 0, s>>>(tmp, n_dev, n_imm, total);
    if (popcount) ScanDownsweepKernel<true><<<grid, kThreads, 0, s>>>(in, out, n_dev, n_imm, tmp);
    else ScanDownsweepKernel<false><<<grid, kThreads, 0, s>>>(in, out, n_dev, n_imm, tmp);
}

size_t RadixTempWords(size_t n) {
    const size_t tiles = (n + kRadixTile - 1) / kRadixTile;
    const size_t hist = 256 * (tiles ? tiles : 1);
    return hist + ScanTempWords(hist);
}

void RadixSortPairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n, int key_bits,
                    uint32_t* tmp, hipStream_t s) {
    if (n == 0) return;
    const uint32_t tiles = (n + kRadixTile - 1) / kRadixTile;
    uint32_t* hist = tmp;
    uint32_t* scan_tmp = tmp + 256 * static_cast<size_t>(tiles);
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8, ++passes) {
        RadixHistKernel<<<tiles, kThreads, 0, s>>>(ki, n, shift, hist, tiles);
        ExclusiveScan(hist, hist, nullptr, 256 * tiles, 256 * tiles, nullptr, false, scan_tmp, s);
        RadixScatterKernel<<<tiles, kThreads, 0, s>>>(ki, vi, ko, vo, n, shift, hist, tiles);
        uint32_t* t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    if (passes & 1) {
        (void)hipMemcpyAsync(keys, ki, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(vals, vi, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
    }
}

void SegmentRanks(const uint32_t* sorted, uint32_t n, uint32_t* rank, uint32_t* t0, uint32_t* t1, uint32_t* t2,
                  uint32_t* scan_tmp, hipStream_t s) {
    if (n == 0) return;
    const uint32_t g = Blocks(n, kThreads);
    CellHeadsKernel<<<g, kThreads, 0, s>>>(sorted, n, t0);
    ExclusiveScan(t0, t1, nullptr, n, n, nullptr, false, scan_tmp, s);
    CellStartsKernel<<<g, kThreads, 0, s>>>(t0, t1, n, t2);
    RunRankKernel<<<g, kThreads, 0, s>>>(t0, t1, t2, n, rank);
}

void SetupOrder(const DomainDev& d, const float* xyz, uint32_t n, int key_bits, uint32_t* keys, uint32_t* order,
                const SetupTemps& t, uint32_t* bitmap, Status* st, hipStream_t s) {
    if (n == 0) return;
    const uint32_t g = Blocks(n, kThreads);
    // cell keys (block << 6 | cell) in input order, stably sorted: runs of one cell
    SetupKeysKernel<<<g, kThreads, 0, s>>>(d, xyz, n, keys, t.t[0], bitmap, st);
    RadixSortPairs(keys, t.t[0], t.keys_alt, t.vals_alt, n, key_bits + 6, t.radix, s);
    CellHeadsKernel<<<g, kThreads, 0, s>>>(keys, n, t.t[1]);
    ExclusiveScan(t.t[1], t.t[2], nullptr, n, n, nullptr, false, t.radix, s);
    CellStartsKernel<<<g, kThreads, 0, s>>>(t.t[1], t.t[2], n, t.t[3]);
    CellRankKernel<<<g, kThreads, 0, s>>>(keys, t.t[1], t.t[2], t.t[3], n, order);
    // LSD: by (rank, cell), then stably by block
    RadixSortPairs(t.t[1], order, t.keys_alt, t.vals_alt, n, 32, t.radix, s);
    SetupGatherKernel<<<g, kThreads, 0, s>>>(order, t.t[2], t.t[0], n, keys, t.t[3]);
    RadixSortPairs(keys, t.t[3], t.keys_alt, t.vals_alt, n, key_bits, t.radix, s);
    (void)hipMemcpyAsync(order, t.t[3], sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
}
void LaunchSetupOffsets(const uint32_t* sorted_keys, uint32_t n, const uint32_t* akeys, const LayoutCounts* lc,
                        uint32_t a_cap, uint32_t* off, hipStream_t s) {
    SetupOffsetsKernel<<<Blocks(a_cap + 1, kThreads), kThreads, 0, s>>>(sorted_keys, n, akeys, lc, off);
}
void LaunchInitState(const DomainDev& d, int kind, const float* xyz, const float* lam_p, const float* mu_p, int pp_attr,
                     uint32_t* order, uint32_t n, float init, float* st, hipStream_t s, int rigid_attr,
                     uint64_t rigid_count) {
    InitStateKernel<<<Blocks(n, kThreads), kThreads, 0, s>>>(d, kind, xyz, lam_p, mu_p, pp_attr, order, n, init, st,
                                                             rigid_attr, rigid_count);
}
void LaunchExportPositions(const DomainDev& d, const uint32_t* keys, const LayoutCounts* lc, const uint32_t* off,
                           const float* st, uint32_t n, float* xyz, uint32_t grid, hipStream_t s) {
    ExportPositionsKernel<<<grid, kThreads, 0, s>>>(d, keys, lc, off, st, n, xyz);
}
void LaunchRasterize(const DomainDev& d, const float* xyz, const float* vel, uint32_t n, uint32_t nvel, float mass,
                     const uint32_t* table, const LayoutCounts* lc, float* M, Status* st, hipStream_t s) {
    RasterizeKernel<<<Blocks(n, kThreads), kThreads, 0, s>>>(d, xyz, vel, n, nvel, mass, table, lc, M, st);
}
void LaunchTableInsertA(const uint32_t* keys, LayoutCounts* lc, uint32_t cap, uint32_t* table, hipStream_t s) {
    TableInsertAKernel<<<Blocks(cap, kThreads, 1024), kThreads, 0, s>>>(keys, lc, cap, table);
}
void LaunchBitmapToKeysA(uint32_t* bitmap, uint32_t nwords, const uint32_t* word_off, const uint32_t* total,
                         uint32_t* keys, LayoutCounts* lc, uint32_t cap, Status* st, hipStream_t s) {
    BitmapEmitKernel<<<Blocks(nwords, kThreads), kThreads, 0, s>>>(bitmap, nwords, word_off, keys, nullptr, lc, cap, 2,
                                                                    st);
    SetCountKernel<<<1, 64, 0, s>>>(lc, 0, total, cap, st);
}
template <int T>
void LaunchG2P2GT(const G2P2GArgs& a, uint32_t grid, hipStream_t s) {
    switch (a.mat.kind) {
        case kFixedCorotated: G2P2GKernel<kFixedCorotated, T><<<grid, T, 0, s>>>(a); break;
        case kNeoHookean: G2P2GKernel<kNeoHookean, T><<<grid, T, 0, s>>>(a); break;
        case kDruckerPrager: G2P2GKernel<kDruckerPrager, T><<<grid, T, 0, s>>>(a); break;
        case kVonMises: G2P2GKernel<kVonMises, T><<<grid, T, 0, s>>>(a); break;
        case kCamClay: G2P2GKernel<kCamClay, T><<<grid, T, 0, s>>>(a); break;
        default: G2P2GKernel<kFluid, T><<<grid, T, 0, s>>>(a); break;
    }
}
uint32_t G2P2GGrid(uint32_t grid) { return ((grid + 7) / 8) * 8; }
// one tile per workgroup for up to ex_hint exterior blocks; the loops cover any count
uint32_t RebinCountGrid(uint32_t cap, uint32_t ex_hint) { return Blocks(std::min(ex_hint, cap), kCountTile, 2048); }
uint32_t RebinScanGrid(uint32_t cap) { return std::max<uint32_t>(1u, (cap + kRebinScanTile - 1) / kRebinScanTile); }
uint32_t MarkGrid(uint32_t cap, uint32_t a_hint) {
    return Blocks(static_cast<uint64_t>(std::min(cap, a_hint)) * kDirections, kThreads);
}
uint32_t ExtTiles(uint64_t nblocks) {
    const uint64_t words = (nblocks + kExtKeysPerWord - 1) / kExtKeysPerWord, tw = ExtTileWords(nblocks);
    return static_cast<uint32_t>((words + tw - 1) / tw);
}
uint32_t ExtTileShift(uint64_t nblocks) {
    static_assert(kExtKeysPerWord == 16, "16 keys per map word");
    return 4u + static_cast<uint32_t>(__builtin_ctz(ExtTileWords(nblocks)));
}

void LaunchOne(int kernel, int variant, const void* args, uint32_t grid, hipStream_t s) {
    switch (kernel) {
        case kFkGrid: {
            const auto& a = *static_cast<const GridUpdateArgs*>(args);
            if (variant) GridUpdateKernel<true><<<grid, kThreads, 0, s>>>(a);
            else GridUpdateKernel<false><<<grid, kThreads, 0, s>>>(a);
            break;
        }
        case kFkG2P2G: {
            const auto& a = *static_cast<const G2P2GArgs*>(args);
            if (variant & 1) LaunchG2P2GT<512>(a, grid, s);
            else LaunchG2P2GT<256>(a, grid, s);
            break;
        }
        case kFkRebinCount:
            RebinCountKernel<<<grid, kCountThreads, 0, s>>>(*static_cast<const RebinCountArgs*>(args));
            break;
        case kFkRebinScan: RebinScanKernel<<<grid, kThreads, 0, s>>>(*static_cast<const RebinScanArgs*>(args)); break;
        case kFkRebinScatter:
            RebinScatterKernel<<<grid, kThreads, 0, s>>>(*static_cast<const RebinScatterArgs*>(args));
            break;
        case kFkMark: MarkExtensionKernel<<<grid, kThreads, 0, s>>>(*static_cast<const MarkArgs*>(args)); break;
        case kFkScatterMark:
            ScatterMarkKernel<<<grid, kThreads, 0, s>>>(*static_cast<const ScatterMarkArgs*>(args));
            break;
        default:
            if (variant) BitmapScanEmitKernel<8><<<grid, kThreads, 0, s>>>(*static_cast<const EmitArgs*>(args));
            else BitmapScanEmitKernel<2><<<grid, kThreads, 0, s>>>(*static_cast<const EmitArgs*>(args));
            break;
    }
}

// ---------------------------------------------------------------- fused batches
// Scene of this workgroup in a fused launch: the last i with begin[i] <= blockIdx.x
// (scenes of zero workgroups share their begin with the next scene and are skipped).
// Every wave counts the entries at or below blockIdx.x with a ballot over 64 entries at
// a time, so the search costs one round of loads for up to 64 scenes.
__device__ __forceinline__ uint32_t FusedScene(const uint32_t* __restrict__ begin, uint32_t n, uint32_t& wg,
                                               uint32_t& ng) {
    const uint32_t b = blockIdx.x, lane = LaneId();
    uint32_t cnt = 0;
    for (uint32_t c = 0; c < n; c += 64) {
        const uint32_t t = c + lane;
        cnt += static_cast<uint32_t>(__popcll(__ballot(t < n && begin[t] <= b)));
    }
    const uint32_t i = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(cnt - 1u)));
    const uint32_t b0 = begin[i];
    wg = b - b0;
    ng = begin[i + 1] - b0;
    return i;
}

// The scene's argument struct, read through the constant address space: the table is
// never written while a launch runs, so every field load can be a scalar load (through
// a generic pointer, the loads after the kernel's first stores become vector loads; the
// fused G2P2G spilled 4-10 VGPRs that way, none this way).
template <class T>
__device__ __forceinline__ const T& FusedArgs(const T* tab, uint32_t i) {
    using CT = const __attribute__((address_space(4))) T;
    return *(const T*)((CT*)tab + i);
}

template <bool kCustom>
__global__ __launch_bounds__(kThreads) void GridUpdateFused(const GridUpdateArgs* __restrict__ tab,
                                                            const uint32_t* __restrict__ begin, uint32_t n) {
    uint32_t wg, ng;
    const uint32_t i = FusedScene(begin, n, wg, ng);
    GridUpdateBody<kCustom>(FusedArgs(tab, i), wg, ng);
}
template <int KIND, int THREADS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void G2P2GFused(
    const G2P2GArgs* __restrict__ tab, const uint32_t* __restrict__ begin, uint32_t n) {
    uint32_t wg, ng;
    const uint32_t i = FusedScene(begin, n, wg, ng);
    G2P2GBody<KIND, THREADS>(FusedArgs(tab, i), wg, ng);
}
#define CKMPM_FUSED(NAME, BODY, ARGS, THREADS)                                                                  \
    __global__ __launch_bounds__(THREADS) void NAME(const ARGS* __restrict__ tab, const uint32_t* __restrict__ begin, \
                                                    uint32_t n) {                                                 \
        uint32_t wg, ng;                                                                                        \
        const uint32_t i = FusedScene(begin, n, wg, ng);                                                        \
        BODY(FusedArgs(tab, i), wg, ng);                                                                        \
    }
CKMPM_FUSED(RebinCountFused, RebinCountBody, RebinCountArgs, kCountThreads)
CKMPM_FUSED(RebinScanFused, RebinScanBody, RebinScanArgs, kThreads)
CKMPM_FUSED(RebinScatterFused, RebinScatterBody, RebinScatterArgs, kThreads)
CKMPM_FUSED(MarkExtensionFused, MarkExtensionBody, MarkArgs, kThreads)
CKMPM_FUSED(BitmapScanEmitFused2, BitmapScanEmitBody<2>, EmitArgs, kThreads)
CKMPM_FUSED(BitmapScanEmitFused8, BitmapScanEmitBody<8>, EmitArgs, kThreads)
CKMPM_FUSED(ScatterMarkFused, ScatterMarkBody, ScatterMarkArgs, kThreads)
#undef CKMPM_FUSED

size_t FusedArgsSize(int kernel) {
    switch (kernel) {
        case kFkGrid: return sizeof(GridUpdateArgs);
        case kFkG2P2G: return sizeof(G2P2GArgs);
        case kFkRebinCount: return sizeof(RebinCountArgs);
        case kFkRebinScan: return sizeof(RebinScanArgs);
        case kFkRebinScatter: return sizeof(RebinScatterArgs);
        case kFkMark: return sizeof(MarkArgs);
        case kFkScatterMark: return sizeof(ScatterMarkArgs);
        default: return sizeof(EmitArgs);
    }
}

template <int T>
void LaunchG2P2GFusedT(int kind, const G2P2GArgs* t, const uint32_t* begin, uint32_t n, uint32_t total, hipStream_t s) {
    switch (kind) {
        case kFixedCorotated: G2P2GFused<kFixedCorotated, T><<<total, T, 0, s>>>(t, begin, n); break;
        case kNeoHookean: G2P2GFused<kNeoHookean, T><<<total, T, 0, s>>>(t, begin, n); break;
        case kDruckerPrager: G2P2GFused<kDruckerPrager, T><<<total, T, 0, s>>>(t, begin, n); break;
        case kVonMises: G2P2GFused<kVonMises, T><<<total, T, 0, s>>>(t, begin, n); break;
        case kCamClay: G2P2GFused<kCamClay, T><<<total, T, 0, s>>>(t, begin, n); break;
        default: G2P2GFused<kFluid, T><<<total, T, 0, s>>>(t, begin, n); break;
    }
}

void LaunchFused(int kernel, int variant, const void* table, const uint32_t* begin, uint32_t n, uint32_t total,
                 hipStream_t s) {
    if (n == 0 || total == 0) return;
    switch (kernel) {
        case kFkGrid: {
            const auto* t = static_cast<const GridUpdateArgs*>(table);
            if (variant) GridUpdateFused<true><<<total, kThreads, 0, s>>>(t, begin, n);
            else GridUpdateFused<false><<<total, kThreads, 0, s>>>(t, begin, n);
            break;
        }
        case kFkG2P2G: {
            const auto* t = static_cast<const G2P2GArgs*>(table);
            if (variant & 1) LaunchG2P2GFusedT<512>(variant >> 1, t, begin, n, total, s);
            else LaunchG2P2GFusedT<256>(variant >> 1, t, begin, n, total, s);
            break;
        }
        case kFkRebinCount:
            RebinCountFused<<<total, kCountThreads, 0, s>>>(static_cast<const RebinCountArgs*>(table), begin, n);
            break;
        case kFkRebinScan:
            RebinScanFused<<<total, kThreads, 0, s>>>(static_cast<const RebinScanArgs*>(table), begin, n);
            break;
        case kFkRebinScatter:
            RebinScatterFused<<<total, kThreads, 0, s>>>(static_cast<const RebinScatterArgs*>(table), begin, n);
            break;
        case kFkMark:
            MarkExtensionFused<<<total, kThreads, 0, s>>>(static_cast<const MarkArgs*>(table), begin, n);
            break;
        case kFkScatterMark:
            ScatterMarkFused<<<total, kThreads, 0, s>>>(static_cast<const ScatterMarkArgs*>(table), begin, n);
            break;
        default:
            if (variant) BitmapScanEmitFused8<<<total, kThreads, 0, s>>>(static_cast<const EmitArgs*>(table), begin, n);
            else BitmapScanEmitFused2<<<total, kThreads, 0, s>>>(static_cast<const EmitArgs*>(table), begin, n);
            break;
    }
}

}  // namespace ckmpm
