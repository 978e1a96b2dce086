// LSD radix sort of (key, value) pairs on the device, 8-bit digits, stable.
// Integer-only, so results are bit-exact and deterministic: it orders the
// Barnes-Hut Morton keys every iteration and the (row, col) edge keys of the
// P symmetrisation.
#pragma once

#include "common.hpp"

namespace tka {

constexpr int kRsBlock = 256;
constexpr int kRsMaxPasses = 8;                       // 64-bit keys
constexpr size_t kRsMaxItems = ((size_t)1 << 30) - 1;  // counts share a word with a 2-bit status flag

constexpr int kRsItems = 8;                           // keys per thread: tile = 2048 keys
// Keys per thread of the counted sort (radix_sort_pairs_counted): its callers
// sort ~1e5 keys per launch, where the tiles' serial ranking and not the
// look-back over more tiles sets the time, so tiles are smaller (1024 keys).
constexpr int kRsItemsCounted = 4;
inline size_t radix_tiles(size_t n, int items = kRsItems)
{
    return (n + (size_t)kRsBlock * items - 1) / ((size_t)kRsBlock * items);
}

// Scratch (uint32 words) for n pairs sorted in tiles of kRsBlock x items
// keys; must be zero-initialised once before the first use (every sort
// leaves its counters at zero).
size_t radix_sort_scratch_words(size_t n, int items = kRsItems);

// Sorts keys[0..n) (only bits [0, end_bit) are examined) carrying vals.
// keys_alt/vals_alt are ping-pong buffers of the same size.  Returns true if
// the sorted result ended in the *_alt buffers.  n <= kRsMaxItems.
template <class K, class V>
bool radix_sort_pairs(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                      hipStream_t stream);

// The same sort without its counting launch, for a caller whose own kernel
// wrote the keys and ran rs_count_tile for every tile (see below).
template <class K, class V>
bool radix_sort_pairs_counted(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                              hipStream_t stream);

// Work that rides along in a sort's passes: pass p launches side_wgs[p]
// workgroups past its tiles, which call side[p].run(w, side_wgs[p]) for
// w = 0 .. side_wgs[p]-1 and take no tile.  The side work must not depend on
// the sort.  Instantiated for AttractSlice<float> and AttractSlice<double>.
struct RsNoSide {
    TKA_DEV void run(int, int) const {}
};
template <class K, class V, class Side>
bool radix_sort_pairs_counted_side(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit,
                                   uint32_t* scratch, const Side* side, const int* side_wgs, hipStream_t stream);

// ---------------------------------------------------------------- device side
// Scratch layout (words): [1] fault, [2, 2+8) tile tickets per pass, [16,
// 16+8) per-pass count readers, then the global digit counts (8 x 256), then
// the tile status words (passes x tiles x 256).
constexpr size_t kRsHeader = 32 + (size_t)kRsMaxPasses * 256;

// The counted sort (32-bit keys, at most 4 passes) keeps its tile status
// words for 4 passes and, after them in the same scratch, every tile's pass-0
// digit counts as 16-bit pairs (tiles x 128 words).  Up to kRsDirectTiles
// tiles pass 0 takes its digit totals and tile offsets from those counts (no
// look-back); the later passes, and every pass of larger sorts, take their
// totals from the global per-pass counts and look back.
constexpr int kRsCountedPasses = 4;
constexpr int kRsDirectTiles = 256;
static_assert(kRsBlock * kRsItemsCounted < 65536, "a tile's digit count fits 16 bits");
TKA_DEV uint32_t* rs_tile_counts(uint32_t* status, int tiles)
{
    return status + (size_t)kRsCountedPasses * 256 * tiles;
}

// Tiles numbered by workgroup index (no arrival tickets) up to this many: all
// of a sort's workgroups then fit on the device at once, so every tile a
// workgroup waits on gets to run.  Larger sorts number tiles in arrival order.
constexpr int kRsMaxIndexedTiles = 512;

struct RsLayout {
    uint32_t* fault;
    uint32_t* tile_ctr;
    uint32_t* rd_ctr;
    uint32_t* gcnt;
    uint32_t* status;
};

inline RsLayout rs_layout(uint32_t* s)
{
    return {s + 1, s + 2, s + 16, s + 32, s + kRsHeader};
}

TKA_DEV uint32_t rs_ld(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
TKA_DEV void rs_st(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

TKA_DEV uint32_t wave_inclusive_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan over the threads of a block (blockDim.x multiple of 64,
// <= 1024).  `sh` needs blockDim.x/64 words.  *total receives the block sum.
TKA_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    __syncthreads();
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int i = 0; i < nw; i++) {
        const uint32_t s = sh[i];
        before += (i < w) ? s : 0u;
        all += s;
    }
    if (total) *total = all;
    return before + inc - v;
}

// Adds the digit histograms (all passes) of one tile's keys to a block's
// LDS histograms h.  Keys that arrive nearly sorted (CSR edges, Morton keys
// of points kept in Morton order) give whole waves one digit: those add once,
// from one lane, instead of 64 atomics serialised on a single bin.
template <class K, int ITEMS>
TKA_DEV void rs_tile_histogram(const K (&key)[ITEMS], const bool (&ok)[ITEMS], int passes, uint32_t (*h)[256])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint64_t act = __ballot(ok[k]);
        for (int p = 0; p < passes; p++) {
            const uint32_t d = (uint32_t)(key[k] >> (8 * p)) & 255u;
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            if (__ballot(ok[k] && d == d0) == act) {
                if (act != 0 && lane == __ffsll((long long)act) - 1) atomicAdd(&h[p][d0], (uint32_t)__popcll(act));
            } else if (ok[k]) {
                atomicAdd(&h[p][d], 1u);
            }
        }
    }
}

// Adds a block's histograms to the global per-pass counts.  Called by all
// kRsBlock threads after a barrier that completes h.  The counts are turned
// into digit bases by each tile of the pass that uses them, and cleared by
// the last tile to read them, ready for the next sort.
TKA_DEV void rs_count_flush(int passes, RsLayout L, uint32_t (*h)[256], int p0 = 0)
{
    const int t = threadIdx.x;
    for (int p = p0; p < passes; p++) {
        const uint32_t c = h[p][t];
        if (c) atomicAdd(&L.gcnt[p * 256 + t], c);
    }
}

// Counting step of the counted sort for one tile, called by all kRsBlock
// threads of the workgroup that holds tile `tile` (keys
// tile*kRsBlock*kRsItemsCounted + k*kRsBlock + threadIdx.x, k <
// kRsItemsCounted; ok[k] false past n), one workgroup per tile.
// Stores the tile's pass-0 digit counts (direct mode) and adds the other
// passes' (all passes' otherwise) to the global counts, and clears the
// tile's status words.  The sort's passes must follow: they clear the global
// counts again.
template <class K>
TKA_DEV void rs_count_tile(const K (&key)[kRsItemsCounted], const bool (&ok)[kRsItemsCounted], int passes, RsLayout L,
                           int tile,
                           int tiles)
{
    __shared__ uint32_t h[kRsCountedPasses][256];
    const int t = threadIdx.x;
    const bool direct = tiles <= kRsDirectTiles;
    for (int p = 0; p < passes; p++) {
        h[p][t] = 0;
        if (p > 0 || !direct) L.status[((size_t)p * tiles + tile) * 256 + t] = 0;  // direct pass 0: no status
    }
    if (tile == 0 && t < kRsMaxPasses) L.tile_ctr[t] = 0;
    __syncthreads();
    rs_tile_histogram<K, kRsItemsCounted>(key, ok, passes, h);
    __syncthreads();
    if (direct && t < 128) rs_tile_counts(L.status, tiles)[(size_t)tile * 128 + t] = h[0][2 * t] | (h[0][2 * t + 1] << 16);
    rs_count_flush(passes, L, h, direct ? 1 : 0);
}

// Nonzero if a sort using this scratch gave up waiting on a predecessor tile
// (its output is then wrong); reads one word, synchronising the stream.
uint32_t radix_sort_fault(const uint32_t* scratch, hipStream_t stream);

// Device-wide exclusive scan of n uint32 counts into out (n+1 entries: out[n]
// is the total).  scratch needs exclusive_scan_scratch_words(n) words.
constexpr int kScanTile = 2048;
inline size_t exclusive_scan_scratch_words(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t stream);

}  // namespace tka
