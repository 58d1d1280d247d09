// Deterministic three-choice bucketed hash of block keys, and its C-ABI.
//
// Build (all on the table's stream, no host sync):
//   1. sort the (key, value) pairs by key;
//   2. per round r: bucket[i] = h_r(key_i) % B for unplaced pairs, B (sorts last) for
//      placed ones; stable sort of pair indices by bucket; rank inside each bucket's
//      run; pair with rank < free slots of its bucket -> slot fill + rank; then each
//      run's last element raises the bucket's fill;
//   3. unplaced pairs, in key order, -> stash.
// Find: wave-cooperative, a 16-lane quarter wavefront per query and probe (one 8-byte
// slot per lane, one 128-byte line per bucket), as CooperativeFind
// (gpu_hash_table.cuh:463-513) does with its 16-thread tiles.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/ckmpm.h"
#include "hash.h"
#include "kernels.h"

using namespace ckmpm;

namespace {

constexpr uint32_t kHashThreads = 256;

__global__ void HashPackKernel(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t* sk, uint32_t* sv,
                               uint32_t* placed, uint32_t nblocks, uint32_t* bad) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        if (k >= nblocks) atomicOr(bad, 1u);
        sk[i] = k;
        sv[i] = vals[i];
        placed[i] = 0u;
    }
}

// Empties the slots, zeroes the bucket fills, the error word and the stash count (the
// stash pass sets the count of a non-empty build).
__global__ void HashClearKernel(uint64_t* slots, uint32_t nslots, uint32_t* fill, uint32_t buckets, uint32_t* bad,
                                uint32_t* nstash) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += gridDim.x * blockDim.x) {
        slots[i] = kHashEmpty;
        if (i < buckets) fill[i] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *bad = 0u;
        *nstash = 0u;
    }
}

// duplicate keys are an error (they would make Find ambiguous)
__global__ void HashDupKernel(const uint32_t* sk, uint32_t n, uint32_t* bad) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x + 1; i < n; i += gridDim.x * blockDim.x)
        if (sk[i] == sk[i - 1]) atomicOr(bad, 2u);
}

__global__ void HashBucketKernel(HashView h, int round, const uint32_t* sk, const uint32_t* placed, uint32_t n,
                                 uint32_t* bucket, uint32_t* idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        bucket[i] = placed[i] ? h.buckets : h.Bucket(sk[i], round);
        idx[i] = i;
    }
}

__global__ void HashPlaceKernel(const uint32_t* bucket, const uint32_t* idx, const uint32_t* rank, uint32_t n,
                                uint32_t buckets, const uint32_t* sk, const uint32_t* sv, const uint32_t* fill,
                                uint64_t* slots, uint32_t* placed) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t b = bucket[i];
        if (b >= buckets) continue;
        const uint32_t f = fill[b];
        if (rank[i] < kHashBucket - f) {
            const uint32_t p = idx[i];
            slots[static_cast<size_t>(b) * kHashBucket + f + rank[i]] =
                (static_cast<uint64_t>(sk[p]) << 32) | static_cast<uint64_t>(sv[p]);
            placed[p] = 1u;
        }
    }
}

__global__ void HashFillKernel(const uint32_t* bucket, const uint32_t* rank, uint32_t n, uint32_t buckets,
                               uint32_t* fill) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t b = bucket[i];
        if (b >= buckets) continue;
        if (i + 1 < n && bucket[i + 1] == b) continue;  // not the run's last element
        fill[b] = min(static_cast<uint32_t>(kHashBucket), fill[b] + rank[i] + 1u);
    }
}

__global__ void HashUnplacedKernel(const uint32_t* placed, uint32_t n, uint32_t* flag) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        flag[i] = placed[i] ? 0u : 1u;
}

__global__ void HashStashKernel(const uint32_t* flag, const uint32_t* pos, uint32_t n, const uint32_t* sk,
                                const uint32_t* sv, uint64_t* stash, uint32_t stash_cap, const uint32_t* total,
                                uint32_t* nstash, uint32_t* bad) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (flag[i] && pos[i] < stash_cap)
            stash[pos[i]] = (static_cast<uint64_t>(sk[i]) << 32) | static_cast<uint64_t>(sv[i]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t t = *total;
        *nstash = min(t, stash_cap);
        if (t > stash_cap) atomicOr(bad, 4u);
    }
}

// A 16-lane group serves the queries of its own 16 lanes in turn; lane l reads slot l.
__global__ __launch_bounds__(kHashThreads) void HashFindKernel(HashView h, const uint32_t* keys, uint32_t n,
                                                               uint32_t* out) {
    const uint32_t lane = threadIdx.x & 63u, sub = lane & 15u, gbase = lane & ~15u;
    const uint64_t gmask = 0xFFFFull << gbase;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t iters = (n + stride - 1) / stride;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t i = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const bool have = i < n;
        const uint32_t mykey = have ? keys[i] : 0u;
        uint32_t result = kEmpty;
        for (uint32_t q = 0; q < 16u; ++q) {
            const uint32_t owner = gbase + q;
            const bool owner_has = (__ballot(have) >> owner) & 1ull;
            if (!owner_has) continue;
            const uint32_t key = __shfl(mykey, static_cast<int>(owner), 64);
            uint32_t found = kEmpty;
            bool done = false;
            for (int r = 0; r < kHashRounds && !done; ++r) {
                const uint64_t w = h.slots[static_cast<size_t>(h.Bucket(key, r)) * kHashBucket + sub];
                const uint64_t hit = __ballot(static_cast<uint32_t>(w >> 32) == key) & gmask;
                const uint64_t empty = __ballot(w == kHashEmpty) & gmask;
                if (hit) {
                    const int src = __ffsll(static_cast<unsigned long long>(hit)) - 1;
                    found = __shfl(static_cast<uint32_t>(w), src, 64);
                    done = true;
                } else if (empty) {
                    done = true;
                }
            }
            if (!done) {
                // every probed bucket was full: scan the stash, 16 words at a time
                const uint32_t ns = *h.nstash;
                for (uint32_t b0 = 0; b0 < ns && !done; b0 += 16u) {
                    const uint32_t j = b0 + sub;
                    const uint64_t w = j < ns ? h.stash[j] : kHashEmpty;
                    const uint64_t hit = __ballot(static_cast<uint32_t>(w >> 32) == key) & gmask;
                    if (hit) {
                        const int src = __ffsll(static_cast<unsigned long long>(hit)) - 1;
                        found = __shfl(static_cast<uint32_t>(w), src, 64);
                        done = true;
                    }
                }
            }
            if (sub == q) result = found;
        }
        if (have) out[i] = result;
    }
}

// LDS-probed Find for tables of up to kStagedBuckets buckets.  Each workgroup (1024
// threads, one per CU) first copies the table's keys - the high words of its slots, 4
// bytes per slot, at most the CU's 160 KiB of LDS - into LDS with 16-byte loads, then every thread
// probes its own keys there: a bucket is four 16-byte LDS reads and the same bucket rule
// as HashView::Find (a bucket with an empty slot ends the search).  A hit reads its value
// from the slot in global memory; a key that finds all its buckets full scans the stash
// there.  The copy costs the table's size per workgroup, so this pays off when the keys
// outnumber the slots by more than the workgroup count.
constexpr uint32_t kStagedThreads = 1024;
constexpr uint32_t kStagedBuckets = 2560;  // 40960 slot keys: 160 KiB
__global__ __launch_bounds__(kStagedThreads) void HashFindStagedKernel(HashView h, const uint32_t* keys, uint32_t n,
                                                                       uint32_t* out) {
    __shared__ __align__(16) uint32_t sK[kStagedBuckets * kHashBucket];
    const uint32_t pairs = h.buckets * kHashBucket / 2;  // two slots per 16-byte load
    const uint4* src = reinterpret_cast<const uint4*>(h.slots);
    for (uint32_t i = threadIdx.x; i < pairs; i += kStagedThreads) {
        const uint4 q = src[i];
        *reinterpret_cast<uint2*>(sK + 2 * i) = make_uint2(q.y, q.w);
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * kStagedThreads + threadIdx.x; i < n; i += gridDim.x * kStagedThreads) {
        const uint32_t key = keys[i];
        uint32_t found = kEmpty;
        bool done = false;
        for (int r = 0; r < kHashRounds && !done; ++r) {
            const uint32_t b = h.Bucket(key, r);
            const uint4* row = reinterpret_cast<const uint4*>(sK + static_cast<size_t>(b) * kHashBucket);
            int hit = -1;
            bool full = true;
#pragma unroll
            for (int k = 0; k < kHashBucket / 4; ++k) {
                const uint4 q = row[k];
                const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (hit < 0 && w[c] == key) hit = 4 * k + c;
                    full &= w[c] != 0xFFFFFFFFu;
                }
            }
            if (hit >= 0) {
                found = static_cast<uint32_t>(h.slots[static_cast<size_t>(b) * kHashBucket + hit]);
                done = true;
            } else if (!full) {
                done = true;
            }
        }
        if (!done) {
            const uint32_t ns = *h.nstash;
            for (uint32_t j = 0; j < ns; ++j) {
                const uint64_t p = h.stash[j];
                if (static_cast<uint32_t>(p >> 32) == key) {
                    found = static_cast<uint32_t>(p);
                    break;
                }
            }
        }
        out[i] = found;
    }
}

// One key per thread (HashView::Find), for comparison with the cooperative form.
__global__ __launch_bounds__(kHashThreads) void HashFindScalarKernel(HashView h, const uint32_t* keys, uint32_t n,
                                                                     uint32_t* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = h.Find(keys[i]);
}

__global__ void HashHistKernel(const uint64_t* slots, uint32_t buckets, uint32_t* hist) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < buckets; b += gridDim.x * blockDim.x) {
        uint32_t f = 0;
        for (int i = 0; i < kHashBucket; ++i) f += slots[static_cast<size_t>(b) * kHashBucket + i] != kHashEmpty;
        atomicAdd(&hist[f], 1u);
    }
}

int KeyBits(uint32_t nblocks) {
    int b = 1;
    while ((1ull << b) < nblocks) ++b;
    return b;
}

}  // namespace

struct ckmpm_hash {
    hipStream_t stream = nullptr;
    int dim[3] = {0, 0, 0};
    uint32_t nblocks = 0;
    uint32_t max_keys = 0;
    uint32_t buckets = 0;
    uint32_t stash_cap = 0;
    uint32_t seed[kHashRounds] = {0, 0, 0};
    uint64_t* slots = nullptr;
    uint64_t* stash = nullptr;
    uint32_t* nstash = nullptr;
    uint32_t* bad = nullptr;
    uint32_t* fill = nullptr;
    uint32_t* tmp = nullptr;  // 10 arrays of max_keys words + scan/radix temporaries
    uint32_t* dev_io = nullptr;  // 2 * max_keys words for the host-pointer entry points
    uint32_t n = 0;
    std::string err;
};

namespace {

int HashFail(ckmpm_hash* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HASH_HIP(h, call)                                                                     \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) return HashFail((h), CKMPM_ERR_HIP, std::string(#call) + ": " + \
                                                           hipGetErrorString(_e));            \
    } while (0)

HashView View(const ckmpm_hash* h) {
    HashView v{};
    v.slots = h->slots;
    v.stash = h->stash;
    v.nstash = h->nstash;
    v.buckets = h->buckets;
    for (int r = 0; r < kHashRounds; ++r) v.seed[r] = h->seed[r];
    for (int a = 0; a < 3; ++a) v.dim[a] = h->dim[a];
    return v;
}

size_t TmpWords(uint32_t m) {
    return 10ull * m + std::max(RadixTempWords(m), ScanTempWords(m)) + 64;
}

// Table geometry for at most max_keys pairs: load factor at most 0.7 (three rounds then
// leave almost nothing for the stash).
uint32_t BucketsFor(uint32_t max_keys) {
    return static_cast<uint32_t>((static_cast<uint64_t>(max_keys) * 10 + 7 * kHashBucket - 1) / (7 * kHashBucket));
}
uint32_t StashFor(uint32_t max_keys) { return 256 + max_keys / 64; }

// Device memory of one build: the table (slots, stash, count) and its temporaries.
struct BuildMem {
    HashView v;        // the table as Find reads it: buckets, seeds, domain
    uint64_t* slots;   // [buckets * 16]
    uint64_t* stash;   // [stash_cap]
    uint32_t* nstash;
    uint32_t nblocks;
    uint32_t max_keys;
    uint32_t stash_cap;
    uint32_t* bad;     // error bits: 1 key outside the domain, 2 duplicate, 4 stash overflow
    uint32_t* fill;    // [buckets]
    uint32_t* tmp;     // TmpWords(max_keys)
};

BuildMem MemOf(const ckmpm_hash* h) {
    BuildMem b{};
    b.v = View(h);
    b.slots = h->slots;
    b.stash = h->stash;
    b.nstash = h->nstash;
    b.nblocks = h->nblocks;
    b.max_keys = h->max_keys;
    b.stash_cap = h->stash_cap;
    b.bad = h->bad;
    b.fill = h->fill;
    b.tmp = h->tmp;
    return b;
}

// Enqueue the build from device arrays; no host sync.
void EnqueueBuild(const BuildMem& b, const uint32_t* keys, const uint32_t* vals, uint32_t n, hipStream_t s) {
    const uint32_t m = b.max_keys;
    uint32_t* t = b.tmp;
    uint32_t *sk = t, *sv = t + m, *placed = t + 2 * m, *bucket = t + 3 * m, *idx = t + 4 * m;
    uint32_t *rank = t + 5 * m, *t0 = t + 6 * m, *t1 = t + 7 * m, *t2 = t + 8 * m, *alt = t + 9 * m;
    uint32_t* radix = t + 10 * static_cast<size_t>(m);
    uint32_t* total = radix + std::max(RadixTempWords(m), ScanTempWords(m));
    const uint32_t buckets = b.v.buckets, nslots = buckets * kHashBucket;
    HashClearKernel<<<LaunchBlocks(nslots), kHashThreads, 0, s>>>(b.slots, nslots, b.fill, buckets, b.bad, b.nstash);
    if (n == 0) return;
    const uint32_t g = LaunchBlocks(n);
    HashPackKernel<<<g, kHashThreads, 0, s>>>(keys, vals, n, sk, sv, placed, b.nblocks, b.bad);
    RadixSortPairs(sk, sv, t0, t1, n, KeyBits(b.nblocks), radix, s);
    HashDupKernel<<<g, kHashThreads, 0, s>>>(sk, n, b.bad);
    const int bbits = KeyBits(buckets + 1);
    for (int r = 0; r < kHashRounds; ++r) {
        HashBucketKernel<<<g, kHashThreads, 0, s>>>(b.v, r, sk, placed, n, bucket, idx);
        RadixSortPairs(bucket, idx, t0, alt, n, bbits, radix, s);
        SegmentRanks(bucket, n, rank, t0, t1, t2, radix, s);
        HashPlaceKernel<<<g, kHashThreads, 0, s>>>(bucket, idx, rank, n, buckets, sk, sv, b.fill, b.slots, placed);
        HashFillKernel<<<g, kHashThreads, 0, s>>>(bucket, rank, n, buckets, b.fill);
    }
    HashUnplacedKernel<<<g, kHashThreads, 0, s>>>(placed, n, t0);
    ExclusiveScan(t0, t1, nullptr, n, n, total, false, radix, s);
    HashStashKernel<<<g, kHashThreads, 0, s>>>(t0, t1, n, sk, sv, b.stash, b.stash_cap, total, b.nstash, b.bad);
}

// Waits for the build and turns its error bits into a status code (and message).
int BuildStatus(const uint32_t* bad_dev, hipStream_t s, std::string* err) {
    uint32_t bad = 0;
    hipError_t e = hipMemcpyAsync(&bad, bad_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        if (err) *err = std::string("hash build: ") + hipGetErrorString(e);
        return CKMPM_ERR_HIP;
    }
    const char* msg = nullptr;
    int rc = CKMPM_OK;
    if (bad & 1u) msg = "key outside the domain", rc = CKMPM_ERR_INVALID;
    else if (bad & 2u) msg = "duplicate key", rc = CKMPM_ERR_INVALID;
    else if (bad & 4u) msg = "hash stash overflow (too many keys for the table)", rc = CKMPM_ERR_CAPACITY;
    if (msg && err) *err = msg;
    return rc;
}

int CheckBuild(ckmpm_hash* h) { return BuildStatus(h->bad, h->stream, &h->err); }

// Scratch of the build that allocates nothing (ckmpm_hash_build_noalloc): the
// temporaries, bucket fills and error word of a table of up to
// CKMPM_HASH_NOALLOC_MAX_KEYS pairs.  Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress gives for
// the device that is current, kept per device below.
constexpr size_t kNoAllocScratchWords = 11u << 18;  // 11.5 MB
__device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {};

// Address of this device's scratch (looked up once per device), or null on error.
uint32_t* DeviceScratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" {

int ckmpm_hash_create(const ckmpm_domain_desc* domain, uint32_t max_keys, ckmpm_hash** out) {
    if (!domain || !out || max_keys == 0) return CKMPM_ERR_INVALID;
    *out = nullptr;
    const uint64_t nb = static_cast<uint64_t>(domain->dim[0]) * domain->dim[1] * domain->dim[2];
    if (domain->dim[0] <= 0 || domain->dim[1] <= 0 || domain->dim[2] <= 0 || nb >= 0xFFFFFFFFull)
        return CKMPM_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CKMPM_ERR_NO_DEVICE;
    ckmpm_hash* h = new ckmpm_hash();
    for (int a = 0; a < 3; ++a) h->dim[a] = domain->dim[a];
    h->nblocks = static_cast<uint32_t>(nb);
    h->max_keys = max_keys;
    h->buckets = BucketsFor(max_keys);
    h->stash_cap = StashFor(max_keys);
    std::mt19937 rng(2);  // the reference's seeds (gpu_hash_table.cuh:303-309)
    for (int r = 0; r < kHashRounds; ++r) h->seed[r] = static_cast<uint32_t>(rng());
    bool ok = hipStreamCreate(&h->stream) == hipSuccess;
    ok = ok && hipMalloc(&h->slots, sizeof(uint64_t) * h->buckets * kHashBucket) == hipSuccess;
    ok = ok && hipMalloc(&h->stash, sizeof(uint64_t) * h->stash_cap) == hipSuccess;
    ok = ok && hipMalloc(&h->nstash, sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMalloc(&h->bad, sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMalloc(&h->fill, sizeof(uint32_t) * h->buckets) == hipSuccess;
    ok = ok && hipMalloc(&h->tmp, sizeof(uint32_t) * TmpWords(max_keys)) == hipSuccess;
    ok = ok && hipMalloc(&h->dev_io, sizeof(uint32_t) * 2 * static_cast<size_t>(max_keys)) == hipSuccess;
    if (!ok) {
        ckmpm_hash_destroy(h);
        return CKMPM_ERR_HIP;
    }
    EnqueueBuild(MemOf(h), nullptr, nullptr, 0, h->stream);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        ckmpm_hash_destroy(h);
        return CKMPM_ERR_HIP;
    }
    *out = h;
    return CKMPM_OK;
}

int ckmpm_hash_build_device(ckmpm_hash* h, const uint32_t* keys, const uint32_t* values, uint32_t n) {
    if (!h || (n && (!keys || !values))) return CKMPM_ERR_INVALID;
    if (n > h->max_keys) return HashFail(h, CKMPM_ERR_CAPACITY, "more keys than the table was created for");
    EnqueueBuild(MemOf(h), keys, values, n, h->stream);
    return CheckBuild(h);
}

int ckmpm_hash_build_noalloc(const ckmpm_domain_desc* domain, uint32_t max_keys, const uint32_t* keys,
                             const uint32_t* values, uint32_t n, uint64_t* slots, uint64_t* stash, uint32_t* nstash,
                             void* stream) {
    if (!domain || max_keys == 0 || !slots || !stash || !nstash || (n && (!keys || !values))) return CKMPM_ERR_INVALID;
    const uint64_t nb = static_cast<uint64_t>(domain->dim[0]) * domain->dim[1] * domain->dim[2];
    if (domain->dim[0] <= 0 || domain->dim[1] <= 0 || domain->dim[2] <= 0 || nb >= 0xFFFFFFFFull)
        return CKMPM_ERR_INVALID;
    if (n > max_keys) return CKMPM_ERR_CAPACITY;
    const uint32_t buckets = BucketsFor(max_keys);
    // the scratch holds the temporaries, the bucket fills and the error word
    if (max_keys > CKMPM_HASH_NOALLOC_MAX_KEYS ||
        TmpWords(max_keys) + buckets + 1 > kNoAllocScratchWords)
        return CKMPM_ERR_CAPACITY;
    uint32_t* base = DeviceScratch();
    if (!base) return CKMPM_ERR_HIP;
    BuildMem b{};
    b.slots = slots;
    b.stash = stash;
    b.nstash = nstash;
    b.v.slots = slots;
    b.v.stash = stash;
    b.v.nstash = nstash;
    b.v.buckets = buckets;
    std::mt19937 rng(2);  // the seeds of ckmpm_hash_create (gpu_hash_table.cuh:303-309)
    for (int r = 0; r < kHashRounds; ++r) b.v.seed[r] = static_cast<uint32_t>(rng());
    for (int a = 0; a < 3; ++a) b.v.dim[a] = domain->dim[a];
    b.nblocks = static_cast<uint32_t>(nb);
    b.max_keys = max_keys;
    b.stash_cap = StashFor(max_keys);
    b.tmp = base;
    b.fill = base + TmpWords(max_keys);
    b.bad = b.fill + buckets;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    EnqueueBuild(b, keys, values, n, s);
    if (hipGetLastError() != hipSuccess) return CKMPM_ERR_HIP;
    return BuildStatus(b.bad, s, nullptr);
}

int ckmpm_hash_build(ckmpm_hash* h, const uint32_t* keys, const uint32_t* values, uint32_t n) {
    if (!h || (n && (!keys || !values))) return CKMPM_ERR_INVALID;
    if (n > h->max_keys) return HashFail(h, CKMPM_ERR_CAPACITY, "more keys than the table was created for");
    uint32_t* dk = h->dev_io;
    uint32_t* dv = h->dev_io + h->max_keys;
    if (n) {
        HASH_HIP(h, hipMemcpyAsync(dk, keys, sizeof(uint32_t) * n, hipMemcpyHostToDevice, h->stream));
        HASH_HIP(h, hipMemcpyAsync(dv, values, sizeof(uint32_t) * n, hipMemcpyHostToDevice, h->stream));
    }
    EnqueueBuild(MemOf(h), dk, dv, n, h->stream);
    return CheckBuild(h);
}

int ckmpm_hash_find_device(ckmpm_hash* h, const uint32_t* keys, uint32_t n, uint32_t* values, int32_t cooperative) {
    if (!h || (n && (!keys || !values))) return CKMPM_ERR_INVALID;
    if (cooperative < 0 || cooperative > 2) return HashFail(h, CKMPM_ERR_INVALID, "find mode is 0, 1 or 2");
    if (cooperative == 2 && h->buckets > kStagedBuckets)
        return HashFail(h, CKMPM_ERR_INVALID, "table too large for the LDS-probed find (more than 2560 buckets)");
    if (n == 0) return CKMPM_OK;
    const HashView v = View(h);
    if (cooperative == 2) {
        int dev = 0, cus = 0;
        HASH_HIP(h, hipGetDevice(&dev));
        HASH_HIP(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const uint32_t grid = std::max<uint32_t>(
            1, std::min<uint32_t>(static_cast<uint32_t>(std::max(cus, 1)), (n + kStagedThreads - 1) / kStagedThreads));
        HashFindStagedKernel<<<grid, kStagedThreads, 0, h->stream>>>(v, keys, n, values);
    } else if (cooperative) {
        HashFindKernel<<<LaunchBlocks(n), kHashThreads, 0, h->stream>>>(v, keys, n, values);
    } else {
        HashFindScalarKernel<<<LaunchBlocks(n), kHashThreads, 0, h->stream>>>(v, keys, n, values);
    }
    HASH_HIP(h, hipGetLastError());
    return CKMPM_OK;
}

int ckmpm_hash_find(ckmpm_hash* h, const uint32_t* keys, uint32_t n, uint32_t* values) {
    if (!h || (n && (!keys || !values))) return CKMPM_ERR_INVALID;
    for (uint32_t done = 0; done < n;) {
        const uint32_t m = std::min(n - done, h->max_keys);
        uint32_t* dk = h->dev_io;
        uint32_t* dv = h->dev_io + h->max_keys;
        HASH_HIP(h, hipMemcpyAsync(dk, keys + done, sizeof(uint32_t) * m, hipMemcpyHostToDevice, h->stream));
        const int rc = ckmpm_hash_find_device(h, dk, m, dv, 1);
        if (rc != CKMPM_OK) return rc;
        HASH_HIP(h, hipMemcpyAsync(values + done, dv, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, h->stream));
        HASH_HIP(h, hipStreamSynchronize(h->stream));
        done += m;
    }
    return CKMPM_OK;
}

int ckmpm_hash_synchronize(ckmpm_hash* h) {
    if (!h) return CKMPM_ERR_INVALID;
    HASH_HIP(h, hipStreamSynchronize(h->stream));
    return CKMPM_OK;
}

int ckmpm_hash_info(ckmpm_hash* h, ckmpm_hash_info_t* info) {
    if (!h || !info) return CKMPM_ERR_INVALID;
    std::memset(info, 0, sizeof(*info));
    info->buckets = h->buckets;
    info->slots_per_bucket = kHashBucket;
    info->stash_capacity = h->stash_cap;
    info->keys = h->n;
    for (int r = 0; r < kHashRounds; ++r) info->seeds[r] = h->seed[r];
    uint32_t* hist = h->tmp;  // 17 words of scratch
    HASH_HIP(h, hipMemsetAsync(hist, 0, sizeof(uint32_t) * (kHashBucket + 1), h->stream));
    HashHistKernel<<<LaunchBlocks(h->buckets), kHashThreads, 0, h->stream>>>(h->slots, h->buckets, hist);
    HASH_HIP(h, hipMemcpyAsync(info->fill_histogram, hist, sizeof(uint32_t) * (kHashBucket + 1),
                               hipMemcpyDeviceToHost, h->stream));
    HASH_HIP(h, hipMemcpyAsync(&info->stash, h->nstash, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HASH_HIP(h, hipStreamSynchronize(h->stream));
    for (int f = 0; f <= kHashBucket; ++f) info->occupied_slots += f * info->fill_histogram[f];
    return CKMPM_OK;
}

int ckmpm_hash_export(ckmpm_hash* h, uint64_t* slots, uint64_t slots_capacity, uint64_t* stash,
                      uint32_t stash_capacity, uint32_t* nstash) {
    if (!h || !slots || !nstash) return CKMPM_ERR_INVALID;
    const uint64_t ns = static_cast<uint64_t>(h->buckets) * kHashBucket;
    if (slots_capacity < ns) return HashFail(h, CKMPM_ERR_CAPACITY, "slot buffer too small");
    HASH_HIP(h, hipMemcpyAsync(slots, h->slots, sizeof(uint64_t) * ns, hipMemcpyDeviceToHost, h->stream));
    uint32_t cnt = 0;
    HASH_HIP(h, hipMemcpyAsync(&cnt, h->nstash, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HASH_HIP(h, hipStreamSynchronize(h->stream));
    *nstash = cnt;
    if (cnt) {
        if (!stash || stash_capacity < cnt) return HashFail(h, CKMPM_ERR_CAPACITY, "stash buffer too small");
        HASH_HIP(h, hipMemcpy(stash, h->stash, sizeof(uint64_t) * cnt, hipMemcpyDeviceToHost));
    }
    return CKMPM_OK;
}

int ckmpm_hash_destroy(ckmpm_hash* h) {
    if (!h) return CKMPM_ERR_INVALID;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : {static_cast<void*>(h->slots), static_cast<void*>(h->stash), static_cast<void*>(h->nstash),
                    static_cast<void*>(h->bad), static_cast<void*>(h->fill), static_cast<void*>(h->tmp),
                    static_cast<void*>(h->dev_io)})
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CKMPM_OK;
}

const char* ckmpm_hash_last_error(const ckmpm_hash* h) { return h ? h->err.c_str() : "null table"; }

}  // extern "C"
