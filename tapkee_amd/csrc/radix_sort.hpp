// LSD radix sort of (key, value) pairs on the device, 8-bit digits, stable.
// Integer-only, so results are bit-exact and deterministic: it orders the
// Barnes-Hut Morton keys every iteration and the (row, col) edge keys of the
// P symmetrisation.
#pragma once

#include "common.hpp"

namespace tka {

constexpr int kRsBlock = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsBlock * kRsItems;  // 2048 pairs per workgroup
constexpr int kRsFusedMaxTiles = 64;          // scan folded into the scatter up to this many tiles

// Scratch (uint32) needed for n pairs: per-tile digit histograms + offsets.
inline size_t radix_sort_scratch_words(size_t n)
{
    const size_t tiles = (n + kRsTile - 1) / kRsTile;
    return 2 * 256 * (tiles ? tiles : 1);
}

// Sorts keys[0..n) (only bits [0, end_bit) are examined) carrying vals.
// keys_alt/vals_alt are ping-pong buffers of the same size.  Returns true if
// the sorted result ended in the *_alt buffers.
template <class K, class V>
bool radix_sort_pairs(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                      hipStream_t stream);

// Device-wide exclusive scan of n uint32 counts into out (n+1 entries: out[n]
// is the total).  scratch needs exclusive_scan_scratch_words(n) words.
inline size_t exclusive_scan_scratch_words(size_t n) { return (n + kRsTile - 1) / kRsTile + 1; }
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t stream);

}  // namespace tka
