// Setup ordering for domains of more than 2^26 blocks.
//
// The setup sorts each model's particles by (block key, binning cell): the cell key
// block << 6 | cell fits 32 bits only up to 2^26 blocks (SetupOrder).  Beyond that the
// same order comes from two stable radix passes, by cell (6 bits) and then by block key,
// with the cell carried beside the key; equal (block, cell) runs and the within-cell
// ranks follow as in SetupOrder, and the rest of the setup is shared.  Domains are at
// most 1024^3 blocks (2^30), so a key stays below 2^30 and the dense tables of both
// layouts take at most 8 GiB of the 288 GB.
#include "setup_wide.h"

namespace ckmpm {

namespace {

constexpr uint32_t kWideThreads = 256;

uint32_t WideBlocks(uint64_t n) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + kWideThreads - 1) / kWideThreads, 4096)));
}

// block key and binning cell of every particle (the rule of SetupKeysKernel), the
// input index, and the occupied-block bitmap
__global__ void WideKeysKernel(DomainDev d, const float* xyz, uint32_t n, uint32_t* bkey, uint32_t* cell, uint32_t* idx,
                               uint32_t* bitmap, Status* st) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        int b[3];
        bool out = false;
        uint32_t cb = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int c0 = Cell0(xyz[3 * static_cast<size_t>(p) + a], d.inv_dx);
            b[a] = FloorDiv4(c0 - 2);
            const int c = min(max(b[a], 0), d.dim[a] - 1);
            out |= c != b[a];
            b[a] = c;
            cb = cb * 4u + static_cast<uint32_t>((c0 - 2) & 3);  // from the unclamped cell
        }
        if (out) SetFlag(st, kFlagOutOfDomain);
        const uint32_t key = KeyOf(d, b[0], b[1], b[2]);
        bkey[p] = key;
        cell[p] = cb;
        idx[p] = p;
        atomicOr(&bitmap[key >> 5], 1u << (key & 31u));
    }
}

// dst[i] = src[idx[i]]
__global__ void WideGatherKernel(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}

__global__ void WideIotaKernel(uint32_t n, uint32_t* v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

// heads of equal (block, cell) runs of the sorted arrays
__global__ void WideHeadsKernel(const uint32_t* bkey, const uint32_t* cell, uint32_t n, uint32_t* head) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        head[i] = (i == 0 || bkey[i] != bkey[i - 1] || cell[i] != cell[i - 1]) ? 1u : 0u;
}

__global__ void WideStartsKernel(const uint32_t* head, const uint32_t* ex, uint32_t n, uint32_t* start) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (head[i]) start[ex[i]] = i;
}

// within-block sort key (rank << 6 | cell), rank = position - start of the (block, cell)
// run, clamped as in SetupOrder; idx = position
__global__ void WideRankKernel(const uint32_t* cell, const uint32_t* head, const uint32_t* ex, const uint32_t* start,
                               uint32_t n, uint32_t* k2, uint32_t* idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t rank = min(i - start[ex[i] + head[i] - 1u], kSetupRankClamp);
        k2[i] = (rank << 6) | cell[i];
        idx[i] = i;
    }
}

}  // namespace

void SetupOrderWide(const DomainDev& d, const float* xyz, uint32_t n, int key_bits, uint32_t* keys, uint32_t* order,
                    const SetupTemps& t, uint32_t* bitmap, Status* st, hipStream_t s) {
    if (n == 0) return;
    const uint32_t g = WideBlocks(n);
    uint32_t* bk = t.t[0];    // block keys
    uint32_t* cell = t.t[1];  // binning cells
    uint32_t* p = t.t[2];     // input indices
    uint32_t* j = t.t[3];
    WideKeysKernel<<<g, kWideThreads, 0, s>>>(d, xyz, n, bk, cell, p, bitmap, st);
    // stable LSD over (block, cell): by cell, then by block key
    RadixSortPairs(cell, p, t.keys_alt, t.vals_alt, n, 6, t.radix, s);      // cell, p in cell order
    WideGatherKernel<<<g, kWideThreads, 0, s>>>(bk, p, n, keys);            // block keys in cell order
    WideIotaKernel<<<g, kWideThreads, 0, s>>>(n, j);
    RadixSortPairs(keys, j, t.keys_alt, t.vals_alt, n, key_bits, t.radix, s);  // keys sorted, j: cell-order position
    WideGatherKernel<<<g, kWideThreads, 0, s>>>(cell, j, n, bk);            // bk <- cells in (block, cell) order
    WideGatherKernel<<<g, kWideThreads, 0, s>>>(p, j, n, cell);             // cell <- input indices in that order
    // keys: sorted block keys; bk: their cells; cell: their input indices
    uint32_t* head = p;
    uint32_t* ex = j;
    WideHeadsKernel<<<g, kWideThreads, 0, s>>>(keys, bk, n, head);
    ExclusiveScan(head, ex, nullptr, n, n, nullptr, false, t.radix, s);
    WideStartsKernel<<<g, kWideThreads, 0, s>>>(head, ex, n, t.vals_alt);   // run starts
    WideRankKernel<<<g, kWideThreads, 0, s>>>(bk, head, ex, t.vals_alt, n, t.keys_alt, order);
    // (rank, cell) keys in t.keys_alt, positions in order: move the keys out of the
    // sort's scratch, then sort by (rank, cell) and stably by block
    (void)hipMemcpyAsync(head, t.keys_alt, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
    RadixSortPairs(head, order, t.keys_alt, t.vals_alt, n, 32, t.radix, s);
    WideGatherKernel<<<g, kWideThreads, 0, s>>>(keys, order, n, ex);        // block keys in (rank, cell) order
    WideGatherKernel<<<g, kWideThreads, 0, s>>>(cell, order, n, bk);        // input indices in that order
    (void)hipMemcpyAsync(keys, ex, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
    RadixSortPairs(keys, bk, t.keys_alt, t.vals_alt, n, key_bits, t.radix, s);
    (void)hipMemcpyAsync(order, bk, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
}

}  // namespace ckmpm
