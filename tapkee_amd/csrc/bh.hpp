// Barnes-Hut quadtree on the device, rebuilt every iteration from Morton
// keys (replaces quadtree.hpp's pointer tree and its per-point recursion).
#pragma once

#include "attract.hpp"
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

// Root cell of the points from nb box partials (quadtree.hpp:104-129),
// computed by all BLOCK threads of a workgroup in a fixed order, so every
// workgroup that calls it gets the same bits, and returned to all threads.
// The mean is removed from Y by the key kernel, so the root cell is centred
// at 0.
template <int BLOCK>
TKA_DEV RootBox reduce_root_box(const double* part, int nb, int N)
{
    __shared__ double sh[6][BLOCK / 64];
    __shared__ RootBox out;
    double v[6] = {0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int b = threadIdx.x; b < nb; b += BLOCK) {
        const double* p = part + (size_t)kBoxStride * b;
        v[0] += p[0];
        v[1] += p[1];
        v[2] = fmin(v[2], p[3]);
        v[3] = fmin(v[3], p[4]);
        v[4] = fmax(v[4], p[5]);
        v[5] = fmax(v[5], p[6]);
    }
    v[0] = wave_sum(v[0]);
    v[1] = wave_sum(v[1]);
    v[2] = wave_min(v[2]);
    v[3] = wave_min(v[3]);
    v[4] = wave_max(v[4]);
    v[5] = wave_max(v[5]);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[6] = {0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int ww = 0; ww < BLOCK / 64; ww++) {
            r[0] += sh[0][ww];
            r[1] += sh[1][ww];
            r[2] = fmin(r[2], sh[2][ww]);
            r[3] = fmin(r[3], sh[3][ww]);
            r[4] = fmax(r[4], sh[4][ww]);
            r[5] = fmax(r[5], sh[5][ww]);
        }
        const double mx = r[0] / (double)N, my = r[1] / (double)N;
        const double hw = fmax(r[4] - mx, mx - r[2]) + 1e-5;
        const double hh = fmax(r[5] - my, my - r[3]) + 1e-5;
        out.mean_x = mx;
        out.mean_y = my;
        out.hw = hw;
        out.hh = hh;
        out.size0 = fmax(hw, hh);
    }
    __syncthreads();
    return out;
}

// One node of the pre-order tree, read with one vector load per step of the
// walk.  cl: point count (low kCountBits bits) and quadtree level (high
// bits) of an internal node, 1 for a leaf; link: skip index (end of the
// subtree) of an internal node, kLeafBit | point index for a leaf.  An
// internal node's cell half-width is size0 * 2^-level.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kCountBits = 26;
constexpr int kMaxTreePoints = (1 << kCountBits) - 1;

template <class T>
struct alignas(4 * sizeof(T)) BhNode {
    T x, y;  // centre of mass (leaf: the point)
    uint32_t cl;
    uint32_t link;
};

// Device storage of one tree build; capacity for N points.
template <class T>
struct BhTree {
    int N = 0;
    // Radix passes: the sort orders the top 8*passes key bits, and bh_fix
    // sorts each run of equal top bits by full key.  2 passes (levels 1-8)
    // while every run of equal top 16 bits is short enough for bh_fix's rank
    // sort, 3 (levels 1-12) or 4 when runs are longer (tight clusters).  The
    // choice reads counts that the previous build posted to pinned host
    // memory (no synchronisation; results do not depend on it).
    int passes = 3;
    int forced_passes = 0;          // > 0: this many passes, whatever the counts say
    uint32_t* runs_host = nullptr;  // pinned: keys in long runs at 16 and 24 top bits, last build
    DevBuf<uint32_t> runs;          // the same counts per pairs workgroup
    DevBuf<uint32_t> keys, keys_alt, perm, perm_alt, sort_scratch;  // top key bits, radix-sorted
    DevBuf<uint64_t> kfull;      // full key per point
    DevBuf<uint64_t> k64;        // full keys, sorted
    DevBuf<uint32_t> p2;         // point at each sorted position
    DevBuf<unsigned long long> mask;  // N: levels of internal nodes whose range starts at key j
    DevBuf<int> node_a, node_b;  // per pair (owner only)
    DevBuf<int> node_lvl;        // per pair: level if owner, -1 otherwise
    DevBuf<typename Vec2<T>::type> ys;  // sorted coordinates
    DevBuf<uint32_t> pre;        // N+1 exclusive prefix of (popcount(mask)+1)
    DevBuf<double> sx, sy;       // N+1 exclusive prefix sums of sorted coordinates
    DevBuf<double> part;         // scan partials: coordinate sums per 256-point tile
    DevBuf<uint32_t> tile_cnt;   // internal nodes whose range starts in each 256-point tile
    DevBuf<BhNode<T>> nodes;     // pre-order, 2N - 1 at most
    DevBuf<RootBox> root;
    const uint32_t* sorted_perm = nullptr;  // valid after build
    // Word the Morton sort sets when a tile's bounded look-back gives up
    // (sticky: no later build clears it).
    const uint32_t* sort_fault() const { return sort_scratch.get() + 1; }
    void alloc(int n, MemBudget* mb);
    // The same buffers as views into a caller's region (nothing allocated):
    // no pinned count for the adaptive sort width (3 radix passes up to 2^22
    // points, as a fresh alloc() starts), and the sort's state words must be
    // cleared with clear_sort_state() before the first build.
    void carve(int n, Arena& a);
    void clear_sort_state(hipStream_t s);
    BhTree() = default;
    BhTree(const BhTree&) = delete;
    BhTree& operator=(const BhTree&) = delete;
    ~BhTree()
    {
        if (runs_host) (void)hipHostFree(runs_host);
    }
};

// Bounding-box partials of Y (N x 2, T) -> part (kBoxStride doubles per block).
template <class T>
void bbox_partials_device(const T* Y, int N, double* part, hipStream_t s);
inline int bbox_blocks(int N) { return (int)ceil_div(N, kBoxBlock); }

// Builds the tree for Y (N x 2 row-major, T).  Y is centred in place using the
// mean from `box_part` (nbox blocks of partials of this Y); every key
// workgroup reduces the partials itself, so keep nbox small (hundreds).
template <class T>
void bh_build(T* Y, int N, const double* box_part, int nbox, BhTree<T>& tr, hipStream_t s);

// bh_build in two halves: root cell, centring of Y and keys (after which Y
// is final for this iteration); then sort and nodes.  bh_build_tree can carry
// the attraction rows of `attract` (of this final Y) in extra workgroups of
// its launches, on CUs that the sort and node kernels leave idle.
template <class T>
void bh_build_keys(T* Y, int N, const double* box_part, int nbox, BhTree<T>& tr, hipStream_t s);
template <class T>
void bh_build_tree(const T* Y, int N, BhTree<T>& tr, hipStream_t s, const AttractSlice<T>* attract = nullptr);

// Repulsive forces: neg (N x 2) = sum over accepted cells of
// count*q^2*(y - com) and zpart[block] = partial sums of count*q
// (quadtree.hpp:317-354).  Returns the number of zpart blocks.  With
// `interactions` (1 + blocks words), workgroup b leaves its count of cell and
// leaf interactions in interactions[1 + b]; bh_add_interactions adds them,
// times the number of launches that left them, to interactions[0].
template <class T>
int bh_repulsion(const BhTree<T>& tr, const T* Y, double theta, T* neg, double* zpart,
                 unsigned long long* interactions, hipStream_t s);
void bh_add_interactions(unsigned long long* counts, int nblocks, int launches, hipStream_t s);
// Points per traversal workgroup (2-D and 3-D): one wave, so that the waves
// spread evenly over the CUs; one Z partial per workgroup.
constexpr int kTravBlock = 64;
inline int bh_zblocks(int N) { return (int)ceil_div(N, kTravBlock); }

// Synchronises the stream and tells whether any sort of this tree's builds
// gave up (the forces of such builds are not valid).
template <class T>
bool bh_sort_failed(const BhTree<T>& tr, hipStream_t s);

// Sets the tree's sort fault word from the host, as a sort whose look-back
// gave up would (for tests of the error path: the sort itself stays valid).
template <class T>
void bh_inject_sort_fault(BhTree<T>& tr, hipStream_t s);

}  // namespace tka
