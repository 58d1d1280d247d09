// Sparse block-key -> index map: bucketed three-choice hash with a deterministic build.
//
// Replaces BucketedCuckooHashTable (include/gpu_hash_table.cuh:274-552) for sparse
// domains.  Same key hash (MurmurHash3_x86_32 over the 12-byte int3 block coordinate,
// :39-118), same three seeds (the first three draws of std::mt19937(2), :303-309),
// same bucket choice (hash % bucket count) and the same Find protocol (probe h0, h1,
// h2; a bucket that is not full ends the search, :463-513).
//
// The build differs on purpose: the reference inserts by atomics with random cuckoo
// eviction, so its layout depends on thread timing.  Here the layout is a pure
// function of the (key, value) set: pairs are taken in ascending key order; in round
// r = 0, 1, 2 every still-unplaced pair asks for bucket h_r(key) % buckets, the pairs
// asking for one bucket are ranked by key and fill its free slots in that order, the
// rest go on to the next round; pairs left after round 2 go to a small stash (sorted
// by key) that Find scans last.  A pair overflows bucket h_r only when that bucket
// was filled completely in round r, so the early exit of Find stays valid.
//
// Slot word: key << 32 | value; empty = all ones.  A bucket is 16 slots = 128 B, one
// cache line; a 16-lane quarter of a wavefront reads it with one 8-byte load per lane.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace ckmpm {

constexpr int kHashBucket = 16;
constexpr int kHashRounds = 3;
constexpr uint64_t kHashEmpty = ~0ull;

__host__ __device__ __forceinline__ uint32_t Rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// MurmurHash3_x86_32 (Austin Appleby, public domain algorithm) of three little-endian
// int32 words; no tail, length 12.
__host__ __device__ __forceinline__ uint32_t Murmur3Int3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t h = seed;
    const uint32_t w[3] = {static_cast<uint32_t>(x), static_cast<uint32_t>(y), static_cast<uint32_t>(z)};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        uint32_t k = w[i] * c1;
        k = Rotl32(k, 15) * c2;
        h ^= k;
        h = Rotl32(h, 13) * 5u + 0xe6546b64u;
    }
    h ^= 12u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// Device view of a built table (all pointers device memory).
struct HashView {
    const uint64_t* slots;   // buckets * 16
    const uint64_t* stash;   // stash_cap, first *nstash used, ascending key
    const uint32_t* nstash;  // device count
    uint32_t buckets;
    uint32_t seed[kHashRounds];
    int dim[3];              // block coordinates of a linear key: (x * dim1 + y) * dim2 + z

    __device__ __forceinline__ uint32_t Bucket(uint32_t key, int r) const {
        const int z = static_cast<int>(key % static_cast<uint32_t>(dim[2]));
        const uint32_t xy = key / static_cast<uint32_t>(dim[2]);
        const int y = static_cast<int>(xy % static_cast<uint32_t>(dim[1]));
        const int x = static_cast<int>(xy / static_cast<uint32_t>(dim[1]));
        return Murmur3Int3(x, y, z, seed[r]) % buckets;
    }

    // One thread, one key: the bucket's 16 words as eight 16-byte loads.
    __device__ uint32_t Find(uint32_t key) const {
        for (int r = 0; r < kHashRounds; ++r) {
            const uint4* b = reinterpret_cast<const uint4*>(slots + static_cast<size_t>(Bucket(key, r)) * kHashBucket);
            bool full = true;
#pragma unroll
            for (int i = 0; i < kHashBucket / 2; ++i) {
                const uint4 q = b[i];
                if (q.y == key) return q.x;
                if (q.w == key) return q.z;
                if (q.y == 0xFFFFFFFFu || q.w == 0xFFFFFFFFu) full = false;
            }
            if (!full) return kEmpty;
        }
        const uint32_t ns = *nstash;
        for (uint32_t i = 0; i < ns; ++i) {
            const uint64_t p = stash[i];
            if (static_cast<uint32_t>(p >> 32) == key) return static_cast<uint32_t>(p);
        }
        return kEmpty;
    }
};

}  // namespace ckmpm
