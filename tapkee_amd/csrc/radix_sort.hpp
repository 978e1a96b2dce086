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
inline size_t radix_tiles(size_t n) { return (n + (size_t)kRsBlock * kRsItems - 1) / ((size_t)kRsBlock * kRsItems); }

// Scratch (uint32 words) for n pairs; must be zero-initialised once before
// the first use (every sort leaves its counters at zero).
size_t radix_sort_scratch_words(size_t n);

// Sorts keys[0..n) (only bits [0, end_bit) are examined) carrying vals.
// keys_alt/vals_alt are ping-pong buffers of the same size.  Returns true if
// the sorted result ended in the *_alt buffers.  n <= kRsMaxItems.
template <class K, class V>
bool radix_sort_pairs(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                      hipStream_t stream);

// Nonzero if a sort using this scratch gave up waiting on a predecessor tile
// (its output is then wrong); reads one word, synchronising the stream.
uint32_t radix_sort_fault(const uint32_t* scratch, hipStream_t stream);

// Device-wide exclusive scan of n uint32 counts into out (n+1 entries: out[n]
// is the total).  scratch needs exclusive_scan_scratch_words(n) words.
constexpr int kScanTile = 2048;
inline size_t exclusive_scan_scratch_words(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t stream);

}  // namespace tka
