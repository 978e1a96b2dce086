// Barnes-Hut quadtree on the device, rebuilt every iteration from Morton
// keys (replaces quadtree.hpp's pointer tree and its per-point recursion).
#pragma once

#include "common.hpp"

namespace tka {

template <class T>
struct Vec2;
template <>
struct Vec2<float> {
    typedef float2 type;
};
template <>
struct Vec2<double> {
    typedef double2 type;
};
template <class T>
struct Vec4;
template <>
struct Vec4<float> {
    typedef float4 type;
};
template <>
struct Vec4<double> {
    typedef double4 type;
};

constexpr int kMortonBits = 32;  // per axis: 64-bit keys, leaf cells of 2^-32 of the root box

// Per-block partial statistics of an N x dims array, kBoxStride doubles per
// block: {sum d0, sum d1, sum d2, min x, min y, max x, max y, 0}.
constexpr int kBoxBlock = 256;

// Root cell and derived constants, written by the key kernel each build.
struct RootBox {
    double mean_x, mean_y;  // mean removed from Y by this build
    double hw, hh;          // half-widths (quadtree.hpp:128-129)
    double size0;           // max(hw, hh): half-width used by the acceptance test at level 0
};

// Device storage of one tree build; capacity for N points.
template <class T>
struct BhTree {
    int N = 0;
    DevBuf<uint32_t> keys, keys_alt, perm, perm_alt, sort_scratch;  // high key words, radix-sorted
    DevBuf<uint32_t> klo;        // low key word per point
    DevBuf<uint64_t> k64;        // full keys, sorted
    DevBuf<uint32_t> p2;         // point at each sorted position
    DevBuf<unsigned long long> mask;  // N: levels of internal nodes whose range starts at key j
    DevBuf<int> node_a, node_b;  // per pair (owner only)
    DevBuf<int> node_lvl;        // per pair: level if owner, -1 otherwise
    DevBuf<typename Vec2<T>::type> ys;  // sorted coordinates
    DevBuf<uint32_t> pre;        // N+1 exclusive prefix of (popcount(mask)+1)
    DevBuf<double> sx, sy;       // N+1 exclusive prefix sums of sorted coordinates
    DevBuf<double> part;         // scan partials
    DevBuf<typename Vec4<T>::type> nodes;  // pre-order: com x, com y, count, size^2 (<0: leaf)
    DevBuf<int> aux;             // internal: skip index; leaf: point index
    DevBuf<RootBox> root;
    const uint32_t* sorted_perm = nullptr;  // valid after build
    void alloc(int n, MemBudget* mb);
};

// Bounding-box partials of Y (N x 2, T) -> part (kBoxStride doubles per block).
template <class T>
void bbox_partials_device(const T* Y, int N, double* part, hipStream_t s);
inline int bbox_blocks(int N) { return (int)ceil_div(N, kBoxBlock); }

// Builds the tree for Y (N x 2 row-major, T).  Y is centred in place using the
// mean from `box_part` (nbox blocks of partials of this Y).
template <class T>
void bh_build(T* Y, int N, const double* box_part, int nbox, BhTree<T>& tr, hipStream_t s);

// Repulsive forces: neg (N x 2) = sum over accepted cells of
// count*q^2*(y - com) and zpart[block] = partial sums of count*q
// (quadtree.hpp:317-354).  Returns the number of zpart blocks.
template <class T>
int bh_repulsion(const BhTree<T>& tr, const T* Y, double theta, T* neg, double* zpart,
                 unsigned long long* interactions, hipStream_t s);
inline int bh_zblocks(int N) { return (int)ceil_div(N, 256); }

}  // namespace tka
