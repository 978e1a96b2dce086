// Barnes-Hut octree for 3-D embeddings on the device, rebuilt every
// iteration from Morton keys.  The reference's tree is 2-D only
// (quadtree.hpp:71, QT_NO_DIMS = 2); this is its generalisation to a third
// axis: root box centred on the mean with half-widths
// max(max - mean, mean - min) + 1e-5 per axis, cells halved on every axis,
// eight children (c = 4 zbit + 2 ybit + xbit, a point on a split plane in
// the lower half), a cell used as a summary when it is a leaf or
// max(half-widths)/sqrt(D) < theta.
#pragma once

#include "bh.hpp"
#include "common.hpp"

namespace tka {

constexpr int kMortonBits3 = 21;  // per axis: 63-bit keys, leaf cells of 2^-21 of the root box
constexpr int kBox3Stride = 10;   // doubles per block of 3-D box partials: sums, mins, maxs (3 each), pad
constexpr int kBox3Block = 256;

struct RootBox3 {
    double mean[3];  // mean removed from Y by the build
    double h[3];     // half-widths
    double size0;    // max of the half-widths: the acceptance test's size at level 0
};

// One node of the pre-order tree: centre of mass (leaf: the point), cl =
// point count | level << kCountBits for an internal node (1 for a leaf), link
// = skip index of an internal node or kLeafBit | point for a leaf.  Padded
// to 32 bytes (two vector loads).
template <class T>
struct alignas(32) BhNode3 {
    T x, y, z;
    uint32_t cl;
    uint32_t link;
};

template <class T>
struct BhTree3 {
    int N = 0;
    DevBuf<uint64_t> keys, keys_sorted;      // Morton keys per point; the same in sorted order
    DevBuf<uint32_t> ktop, ktop_alt, perm, perm_alt, sort_scratch;  // top 24 key bits and points, radix-sorted
    DevBuf<uint32_t> p2;                     // point at each sorted position
    const uint64_t* k64 = nullptr;          // sorted keys (valid after build)
    const uint32_t* sorted_perm = nullptr;  // point at each sorted position (valid after build)
    DevBuf<T> ys;                           // sorted coordinates, N x 3
    DevBuf<unsigned long long> mask;        // levels of internal nodes whose range starts at key j
    DevBuf<int> node_a, node_b, node_lvl;   // per pair (owner only)
    DevBuf<uint32_t> pre;                   // N+1 exclusive prefix of (popcount(mask)+1)
    DevBuf<double> ps;                      // 3 x (N+1) exclusive prefix sums of sorted coordinates
    DevBuf<double> part;                    // scan partials: 3 coordinate sums per 256-point tile
    DevBuf<uint32_t> tile_cnt;              // internal nodes whose range starts in each tile
    DevBuf<double> box;                     // bounding-box partials of Y
    DevBuf<BhNode3<T>> nodes;               // pre-order, 2N - 1 at most
    DevBuf<RootBox3> root;
    // Word the Morton sort sets when a tile's bounded look-back gives up
    // (sticky: no later build clears it).
    const uint32_t* sort_fault() const { return sort_scratch.get() + 1; }
    void alloc(int n, MemBudget* mb);
};

// Builds the octree of Y (N x 3 row-major, T).  Y is centred in place (its
// mean removed), as the reference's tree build does on its data copy.
// With `attract`, its rows run in extra workgroups of the node kernels.
template <class T>
void bh3_build(T* Y, int N, BhTree3<T>& tr, hipStream_t s, const AttractSlice<T, 3>* attract = nullptr);

// Repulsive forces: neg (N x 3) = sum over accepted cells of
// count*q^2*(y - com) and zpart[block] = partial sums of count*q; returns
// the number of zpart blocks (bh_zblocks(N)).  `interactions` as for
// bh_repulsion.
template <class T>
int bh3_repulsion(const BhTree3<T>& tr, double theta, T* neg, double* zpart, unsigned long long* interactions,
                  hipStream_t s);

// Synchronises the stream and tells whether any sort of this tree's builds
// gave up (the forces of such builds are not valid).
template <class T>
bool bh3_sort_failed(const BhTree3<T>& tr, hipStream_t s);

// Sets the tree's sort fault word from the host (tests of the error path).
template <class T>
void bh3_inject_sort_fault(BhTree3<T>& tr, hipStream_t s);

}  // namespace tka
