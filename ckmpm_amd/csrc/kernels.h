// Host-side launchers of the CK-MPM kernels (implemented in kernels.hip).
#pragma once

#include "common.h"
#include "../../include/ckmpm.h"

namespace ckmpm {


// ---------------------------------------------------------------- scans
// Exclusive scan of n u32 values (n read from n_dev when non-null, else n_imm; the
// launch is sized for n_cap).  When popcount is set the scanned values are the
// population counts of the input words.  total (optional) receives the sum.
// tmp must hold ScanTempWords(n_cap) words.
size_t ScanTempWords(size_t n_cap);
void ExclusiveScan(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_imm, uint32_t n_cap,
                   uint32_t* total, bool popcount, uint32_t* tmp, hipStream_t s);

// rank[i] = i - (first index of the run of equal values containing i) for a sorted
// array; t0..t2 are n-word temporaries, scan_tmp holds ScanTempWords(n) words.
void SegmentRanks(const uint32_t* sorted, uint32_t n, uint32_t* rank, uint32_t* t0, uint32_t* t1, uint32_t* t2,
                  uint32_t* scan_tmp, hipStream_t s);
uint32_t LaunchBlocks(uint64_t n);  // grid size for a grid-stride launch of n items

// ---------------------------------------------------------------- setup
// Stable LSD radix sort of (key, value) pairs on the low key_bits bits.
size_t RadixTempWords(size_t n);
void RadixSortPairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n, int key_bits,
                    uint32_t* tmp, hipStream_t s);

// Setup ordering of one model: particles sorted by (block key, rank among the earlier
// input particles of the same binning cell, cell within the block).  Writes the sorted
// block keys and the input index of each processing position; marks occupied blocks
// in the bitmap.  key_bits + 6 <= 32.
struct SetupTemps {
    uint32_t* t[4];       // n words each
    uint32_t* keys_alt;   // n
    uint32_t* vals_alt;   // n
    uint32_t* radix;      // max(RadixTempWords(n), ScanTempWords(n))
};
void SetupOrder(const DomainDev& d, const float* xyz, uint32_t n, int key_bits, uint32_t* keys, uint32_t* order,
                const SetupTemps& t, uint32_t* bitmap, Status* st, hipStream_t s);
void LaunchSetupOffsets(const uint32_t* sorted_keys, uint32_t n, const uint32_t* akeys, const LayoutCounts* lc,
                        uint32_t a_cap, uint32_t* off, hipStream_t s);
// order: in = setup sort permutation (processing position -> input index),
// out = identity ordering with move 13.
// pp_attr: attribute index of per-particle (lambda, mu), 0 when the material's are used.
// rigid_attr: attribute index of the rigid flag (1 for the input particles below
// rigid_count), 0: none.
void LaunchInitState(const DomainDev& d, int kind, const float* xyz, const float* lam_p, const float* mu_p, int pp_attr,
                     uint32_t* order, uint32_t n, float init, float* st, hipStream_t s, int rigid_attr = 0,
                     uint64_t rigid_count = 0);
void LaunchExportPositions(const DomainDev& d, const uint32_t* keys, const LayoutCounts* lc, const uint32_t* off,
                           const float* st, uint32_t n, float* xyz, uint32_t grid, hipStream_t s);
void LaunchRasterize(const DomainDev& d, const float* xyz, const float* vel, uint32_t n, uint32_t nvel, float mass,
                     const uint32_t* table, const LayoutCounts* lc, float* M, Status* st, hipStream_t s);

// ---------------------------------------------------------------- layout
// Map words per emit tile: the extension map is allocated in whole tiles (zero padded),
// with two counters per tile (neighbour and exterior blocks marked).  Small tiles spread
// the emit over more workgroups; every marked tile's workgroup reads the counters of all
// tiles, so domains of more than kSmallTileBlocks blocks take large tiles (a domain of
// 2^30 blocks then has 32768 tiles, not 131072).
constexpr uint32_t kEmitTileWordsSmall = 512, kEmitTileWordsLarge = 2048;
constexpr uint64_t kSmallTileBlocks = 1ull << 25;
inline uint32_t ExtTileWords(uint64_t nblocks) {
    return nblocks > kSmallTileBlocks ? kEmitTileWordsLarge : kEmitTileWordsSmall;
}
void LaunchTableInsertA(const uint32_t* keys, LayoutCounts* lc, uint32_t cap, uint32_t* table, hipStream_t s);
void LaunchBitmapToKeysA(uint32_t* bitmap, uint32_t nwords, const uint32_t* word_off, const uint32_t* total,
                         uint32_t* keys, LayoutCounts* lc, uint32_t cap, Status* st, hipStream_t s);

// ---------------------------------------------------------------- substep
// Grid boundary description in device memory (ckmpm_set_boundary); custom = 0 stands for
// the reference's default rule (sticky 2-block margin).
struct BoundaryDev {
    int custom;
    ckmpm_boundary_desc d;
};

struct GridUpdateArgs {
    DomainDev dom;
    const uint32_t* keys;        // current layout
    const LayoutCounts* lc;
    const uint32_t* table_prev;  // layout of M_prev
    const LayoutCounts* lc_prev;
    const float* M_prev;
    float* V;
    float* M_zero;               // cleared for the current layout (next P2G target)
    float dt, gravity, mass_clamp;
    // nodes with less mass are left out of the max speed (not of the velocity field): the
    // fixed-point P2G gives such a node's momentum as a few units, so its p / m can be far
    // off, while particles see the node only with a weight of that size
    float vmax_mass_floor;
    uint32_t cap;
    Status* st;
    // per-block P2G deposit bounds of the previous layout's G2P2G, reduced into st->dep_bits
    int nmodels;
    const float* blkmax[kMaxModels];
    const BoundaryDev* bc;       // boundary description (read when custom)
    int custom;                  // host copy of bc->custom: selects the kernel instantiation
    int slot;                    // substep parity: the Status slot reduced into (the other is cleared)
};

struct G2P2GArgs {
    DomainDev dom;
    MaterialDev mat;
    const uint32_t* keys;
    const uint32_t* table;
    const LayoutCounts* lc;
    const float* V;
    float* M;
    const float* st_in;
    float* st_out;
    uint32_t n;
    const uint32_t* perm;
    const uint32_t* off;
    uint32_t* code;
    uint32_t* cnt;
    int pp_attr;  // attribute index of per-particle (lambda, mu) in the state, 0: none
    int rigid_attr;     // attribute index of the rigid flag, 0: none
    float rigid_v[3];   // velocity of the rigid particles
    float dt, next_dt;
    Status* st;
    int model;      // index of st->dep_bits
    int slot;       // substep parity: the Status slot of this substep
    float* blkmax;  // [A] largest deposit bound per particle block (for the next substep)
    // particle identities (differentiation): ids_out[t] = ids_in[slot of t]; null: none
    const uint32_t* ids_in;
    uint32_t* ids_out;
};

struct RebinModels {
    int nmodels;
    const uint32_t* cnt[kMaxModels];   // [A_cur][27] direction counts per old block
    uint32_t* cnt_e[kMaxModels];       // [Ex_cur] particles entering each exterior block
    uint32_t* eoff[kMaxModels];        // [Ex_cur] next-layout offset of each occupied exterior block
    // [Ex_cur][27] per exterior block e and move d: particles of the model entering e
    // with a move before d
    uint32_t* base[kMaxModels];
    uint32_t* off_next[kMaxModels];    // [A_new + 1] block offsets of the next layout
    const uint32_t* off_cur[kMaxModels];
    const uint32_t* code[kMaxModels];  // [n] move | rank << 5 per processing position
    uint32_t* perm_next[kMaxModels];
    uint32_t n[kMaxModels];
};
// count tile sums: kTileSums words per tile of 32 exterior blocks
constexpr int kTileSums = 1 + kMaxModels;
size_t RebinTileSumWords(uint32_t cap);

// Rebin (RebinCount: entering counts per exterior block of the current layout, cnt_e,
// tile sums, the earlier-move counts of every (exterior block, move), the next
// layout's table cleared and the extension tile counts zeroed; RebinScan: next
// layout's particle blocks (keys, table, A) and every model's block offsets;
// RebinScatter: the next particle permutation) and layout extension (Mark: neighbour
// candidates b + {0,1}^3 and exterior candidates b + {-1,0,1}^3 not in the table, in a
// map of two bits per block, counted per tile; Emit: the neighbour blocks in ascending
// key order after A (sets Nb), then the exterior blocks after Nb (sets Ex), registered in
// the table, the map cleared).  ext_cnt holds 2 counters per map tile, zero before Mark.
struct RebinCountArgs {
    DomainDev d;
    const uint32_t* keys;
    const uint32_t* table;
    const LayoutCounts* lc;
    RebinModels rm;
    uint32_t* tsum;
    const uint32_t* keys_next;
    const LayoutCounts* lc_next;
    uint32_t* table_next;
    uint32_t cap;
    uint32_t* ext_cnt;       // cleared by workgroup 0 for the extension that follows
    uint32_t ext_cnt_words;
};
struct RebinScanArgs {
    const uint32_t* keys;
    const LayoutCounts* lc;
    RebinModels rm;
    const uint32_t* tsum;
    uint32_t* keys_next;
    uint32_t* table_next;
    LayoutCounts* lc_next;
    uint32_t cap;
    Status* st;
};
struct RebinScatterArgs {
    DomainDev d;
    const uint32_t* keys;    // current layout
    const uint32_t* table;
    const LayoutCounts* lc;
    RebinModels rm;
    uint32_t cap;
};
// Extension map: two bits per domain block, kExtKeysPerWord blocks per word.
constexpr uint32_t kExtKeysPerWord = 16;
struct MarkArgs {
    DomainDev d;
    const uint32_t* keys;
    const LayoutCounts* lc;
    uint32_t cap;
    const uint32_t* table;
    uint32_t* map;
    uint32_t* tile_cnt;  // [tiles][2]: neighbour, exterior blocks newly marked per tile
    uint32_t tile_shift; // tile of block key k: k >> tile_shift (16 keys per map word)
};
struct EmitArgs {
    uint32_t* map;
    const uint32_t* tile_cnt;
    uint32_t ntiles;
    uint32_t* keys;
    uint32_t* table;
    LayoutCounts* lc;
    uint32_t cap;
    Status* st;
};
// RebinScatter and the extension's Mark in one launch (both need only the scan's
// outputs): workgroups [0, mark_wgs) mark, the others scatter; the emit pass follows.
struct ScatterMarkArgs {
    RebinScatterArgs x;
    MarkArgs m;
    uint32_t mark_wgs;
};

// Kernels of one substep.  The grid-update and G2P2G entries have variants (the boundary
// instantiation; material kind x workgroup size of 256 or 512 threads).
enum FusedKernel : int {
    kFkGrid = 0,      // variant: custom boundary
    kFkG2P2G,         // variant: kind * 2 + (threads == 512)
    kFkRebinCount,
    kFkRebinScan,
    kFkRebinScatter,
    kFkMark,
    kFkEmit,
    kFkScatterMark,
};
size_t FusedArgsSize(int kernel);  // bytes of the kernel's argument struct
// One launch of grid workgroups; args points to the kernel's argument struct.
void LaunchOne(int kernel, int variant, const void* args, uint32_t grid, hipStream_t s);
// A fused launch runs one phase for many scenes: table holds one argument struct per scene
// (FusedArgsSize bytes each), begin[i] is the first workgroup of scene i (begin[0] = 0,
// begin[n] = total), and workgroup w of the launch runs workgroup w - begin[i] of a
// launch of begin[i+1] - begin[i] workgroups for scene i.  The scenes must not share
// buffers.
void LaunchFused(int kernel, int variant, const void* table, const uint32_t* begin, uint32_t n, uint32_t total,
                 hipStream_t s);
// Workgroup counts: RebinCount for about ex_hint exterior blocks, Mark for about a_hint
// particle blocks (the kernels loop over the device counts, so any hint is correct; a
// close one saves idle workgroups); RebinScan must cover the capacity.
uint32_t RebinCountGrid(uint32_t cap, uint32_t ex_hint);
uint32_t RebinScanGrid(uint32_t cap);
uint32_t MarkGrid(uint32_t cap, uint32_t a_hint);
// map tiles (one emit workgroup each) for a domain of nblocks blocks, of
// ExtTileWords(nblocks) words each; the key shift of Mark; the emit launch variant
uint32_t ExtTiles(uint64_t nblocks);
uint32_t ExtTileShift(uint64_t nblocks);
inline int EmitVariant(uint64_t nblocks) { return ExtTileWords(nblocks) == kEmitTileWordsLarge ? 1 : 0; }
uint32_t G2P2GGrid(uint32_t grid);

}  // namespace ckmpm
