// Barnes-Hut quadtree from Morton keys, and the per-point traversal.
//
// Tree semantics follow quadtree.hpp: the root is the rectangle centred on
// mean(Y) with half-widths max(max - mean, mean - min) + 1e-5 per axis
// (:104-129); a cell splits into NW, NE, SW, SE quadrants with points on a
// split line going west/north (closed intervals tried in that order, :53-64,
// :241-248); leaves hold one point; a cell is used as a summary when it is a
// leaf or max(hh, hw)/sqrt(D) < theta (:335).
//
// Construction (all integer, bit-exact):
//   box    every key workgroup reduces the bounding-box partials, in a fixed
//          order, to the root cell.
//   keys   point -> 32+32 bit Morton key of its leaf cell, y bit above x bit,
//          so digit order 0..3 is NW, NE, SW, SE and sorting keys gives the
//          reference's depth-first child order.  Quantisation is
//          ceil(t * 2^32) - 1, so coordinates on a split line go west/north;
//          its top 16 bits equal ceil(t * 2^16) - 1, so the key's high word
//          is the 16+16 bit key of the level-16 cell.
//   sort   stable LSD radix sort of (top 8*passes key bits, point): 2
//          passes (levels 1-8) while the last build had no run of more than
//          64 keys with equal top 16 bits, 3 (levels 1-12) or 4 when runs
//          are longer, and always 4 beyond 2^22 points.
//   fix    runs of equal top bits (points in one level-8 cell, a few dozen
//          at most in t-SNE embeddings of 1e5 points) are put in full-key
//          order: runs of at most 64 points by each point's rank in the run,
//          longer ones by a wave-wide bitonic sort of (key, point); the
//          result equals a stable sort of the 64-bit keys.
//   pairs  adjacent sorted keys i, i+1 first differ at quadtree level
//          l = clz64(k_i ^ k_i+1)/2; the level-l cell holding both has >= 2
//          non-empty children, i.e. it is a node of the compressed tree.
//          The pair whose keys straddle the end of that cell's first child
//          owns the node; searches give the cell's key range [a,b) (binary
//          searches in an LDS window of keys, a wave-wide 64-way search in
//          global memory beyond it).
//          Chains of one-child cells are skipped: such a cell has the same
//          count and centre of mass as its child and a larger size, so the
//          reference accepts it only if it would accept the child.
//   scan   pre-order positions: the internal nodes whose ranges start at key
//          j (a chain, shallowest first) precede leaf j, so with
//          c_j = popcount(mask_j) + 1, node (a, l) sits at P[a] + #levels < l
//          in mask_a, leaf j at P[j] + popcount(mask_j), and a node's subtree
//          ends at P[b] (its skip pointer).  Centres of mass come from FP64
//          prefix sums of the sorted coordinates.
//   Equal keys (points closer than 2^-32 of the box) form one level-32 node
//   whose children are the individual points.
//
// Traversal is stackless: visit node t; if it is a leaf, interact unless it
// is the point itself and go to t+1; if accepted, interact and jump to its
// skip pointer; otherwise descend to t+1.
#include "bh.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <type_traits>
#include <vector>

#include "radix_sort.hpp"
#include "tree_search.hpp"

namespace tka {
namespace {

constexpr int kB = kMortonBits;
constexpr int kTreeScanTile = 256;  // points per scan tile = per pairs workgroup

// ---------------------------------------------------------------- bbox
template <class T>
__global__ __launch_bounds__(kBoxBlock) void bbox_part(const T* __restrict__ Y, int N, double* __restrict__ part)
{
    __shared__ double sh[6][4];
    const int i = blockIdx.x * kBoxBlock + threadIdx.x;
    double v[6] = {0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    if (i < N) {
        const double x = (double)Y[2 * (size_t)i], y = (double)Y[2 * (size_t)i + 1];
        v[0] = x;
        v[1] = y;
        v[2] = v[4] = x;
        v[3] = v[5] = y;
    }
    v[0] = wave_sum(v[0]);
    v[1] = wave_sum(v[1]);
    v[2] = wave_min(v[2]);
    v[3] = wave_min(v[3]);
    v[4] = wave_max(v[4]);
    v[5] = wave_max(v[5]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[6] = {0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int ww = 0; ww < kBoxBlock / 64; ww++) {
            r[0] += sh[0][ww];
            r[1] += sh[1][ww];
            r[2] = fmin(r[2], sh[2][ww]);
            r[3] = fmin(r[3], sh[3][ww]);
            r[4] = fmax(r[4], sh[4][ww]);
            r[5] = fmax(r[5], sh[5][ww]);
        }
        double* o = part + (size_t)kBoxStride * blockIdx.x;
        o[0] = r[0];
        o[1] = r[1];
        o[2] = 0.0;
        o[3] = r[2];
        o[4] = r[3];
        o[5] = r[4];
        o[6] = r[5];
        o[7] = 0.0;
    }
}

// ---------------------------------------------------------------- keys
TKA_DEV uint32_t spread16(uint32_t v)
{
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Quantise t in (0, 1) to ceil(t * 2^32) - 1 in [0, 2^32 - 1].
TKA_DEV uint32_t quantise(double t)
{
    double c = ceil(ldexp(t, kB)) - 1.0;
    c = c < 0.0 ? 0.0 : (c > 4294967295.0 ? 4294967295.0 : c);
    return (uint32_t)c;
}

// One radix-sort tile of points per workgroup: root cell (every workgroup
// reduces the box partials itself; workgroup 0 stores it), centre, key, then
// the sort's counting step on the top key bits (so the sort needs no counting
// launch).
template <class T>
__global__ __launch_bounds__(kRsBlock) void bh_keys(T* __restrict__ Y, int N, const double* __restrict__ box_part,
                                                    int nbox, RootBox* __restrict__ root, int passes,
                                                    uint32_t* __restrict__ ktop, uint64_t* __restrict__ kfull,
                                                    uint32_t* __restrict__ perm,
                                                    unsigned long long* __restrict__ mask,
                                                    uint32_t* __restrict__ tile_cnt, RsLayout L, int tiles)
{
    {
        const int g = blockIdx.x * kRsBlock + threadIdx.x;
        if (g < (N + kTreeScanTile - 1) / kTreeScanTile) tile_cnt[g] = 0u;
    }
    // the tile's points are requested before the box reduction, so both
    // loads are in flight together
    using V2 = typename Vec2<T>::type;
    const int base = blockIdx.x * kRsBlock * kRsItemsCounted + threadIdx.x;
    V2 yv[kRsItemsCounted];
#pragma unroll
    for (int k = 0; k < kRsItemsCounted; k++) {
        const int i = base + k * kRsBlock;
        yv[k] = i < N ? reinterpret_cast<const V2*>(Y)[i] : V2{};
    }
    const RootBox rb = reduce_root_box<kRsBlock>(box_part, nbox, N);
    if (blockIdx.x == 0 && threadIdx.x == 0) *root = rb;
    const double mx = rb.mean_x, my = rb.mean_y, hw = rb.hw, hh = rb.hh;
    uint32_t hk[kRsItemsCounted];
    bool ok[kRsItemsCounted];
#pragma unroll
    for (int k = 0; k < kRsItemsCounted; k++) {
        const int i = base + k * kRsBlock;
        ok[k] = i < N;
        hk[k] = 0u;
        if (ok[k]) {
            const T x = (T)((double)yv[k].x - mx);
            const T y = (T)((double)yv[k].y - my);
            Y[2 * (size_t)i] = x;
            Y[2 * (size_t)i + 1] = y;
            const uint32_t qx = quantise(((double)x + hw) / (2.0 * hw));
            const uint32_t qy = quantise(((double)y + hh) / (2.0 * hh));
            const uint64_t key = ((uint64_t)((spread16(qy >> 16) << 1) | spread16(qx >> 16)) << 32) |
                                 ((spread16(qy) << 1) | spread16(qx));
            hk[k] = (uint32_t)(key >> (64 - 8 * passes));
            ktop[i] = hk[k];
            kfull[i] = key;
            perm[i] = (uint32_t)i;
            mask[i] = 0ull;
        }
    }
    rs_count_tile<uint32_t>(hk, ok, passes, L, (int)blockIdx.x, tiles);
}

// Full 64-bit keys in sorted order, the final point order and the sorted
// coordinates (run_sort_tile), with attraction rows in extra workgroups.
template <class T>
__global__ __launch_bounds__(256) void bh_fix(const uint32_t* __restrict__ ktop, const uint32_t* __restrict__ perm,
                                              const uint64_t* __restrict__ kfull, const T* __restrict__ Y, int N,
                                              uint64_t* __restrict__ k64, uint32_t* __restrict__ p2,
                                              typename Vec2<T>::type* __restrict__ ys, int nmain,
                                              AttractSlice<T> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    // the sorted coordinates along with the order
    run_sort_tile(ktop, perm, kfull, N, k64, p2, [&](int d, uint32_t p) {
        typename Vec2<T>::type v;
        v.x = Y[2 * (size_t)p];
        v.y = Y[2 * (size_t)p + 1];
        ys[d] = v;
    });
}

// ---------------------------------------------------------------- pairs
// Keys [w0, w0 + kPairWin) of the sorted array are staged in LDS: the
// searches of deep cells end inside it (binary search in LDS).  A search whose
// answer lies outside the window (shallow cells, or a cell cut by the window's
// edge) is done by the whole wave, one lane's search at a time, as a 64-way
// search over global memory: each round 64 probes narrow the range 64-fold,
// so the root's search over N keys takes 3 dependent loads at N = 2^18
// instead of 2 log2 N for a per-thread gallop.
constexpr int kPairWin = 3 * 256;

//
// The workgroup also leaves the scan's tile partials for its 256 points (one
// scan tile): the FP64 sums of their coordinates, and, in tile_cnt, the
// number of internal nodes whose range starts in each tile (the owners'
// counts, gathered per workgroup for the two tiles the window covers).
template <class T>
__global__ __launch_bounds__(256) void bh_pairs(const uint64_t* __restrict__ keys,
                                                const typename Vec2<T>::type* __restrict__ ys, int N,
                                                unsigned long long* __restrict__ mask, int* __restrict__ node_a,
                                                int* __restrict__ node_b, int* __restrict__ node_lvl,
                                                uint32_t* __restrict__ tile_cnt, double* __restrict__ part,
                                                uint32_t* __restrict__ runs, int nmain, AttractSlice<T> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    __shared__ uint64_t win[kPairWin];
    __shared__ uint32_t own_cnt[2];  // owners with a in tile blockIdx.x - 1, blockIdx.x
    __shared__ uint32_t run_cnt[2];  // keys that share their top 16 / 24 bits with the key kFixHalo on
    __shared__ double sh[4];
    const int w0 = (int)blockIdx.x * 256 - 256;
    for (int t = threadIdx.x; t < kPairWin; t += 256) {
        const int g = w0 + t;
        win[t] = (g >= 0 && g < N) ? keys[g] : 0ull;
    }
    if (threadIdx.x < 2) {
        own_cnt[threadIdx.x] = 0u;
        run_cnt[threadIdx.x] = 0u;
    }
    __syncthreads();
    // the window's valid keys: [wl, wh)
    const int wl = max(w0, 0), wh = min(w0 + kPairWin, N);
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool act = i < N - 1;  // pair (i, i+1); every lane stays for the wave's searches
    const uint64_t ki = act ? win[i - w0] : 0ull, kj = act ? win[i + 1 - w0] : 0ull;
    const uint64_t x = ki ^ kj;
    {
        // Radix width of the next build: a key that shares its top 16 (24)
        // bits with the key kFixHalo positions on lies in a run longer than
        // kFixHalo points when 2 (3) radix passes sort the keys.
        // (i + kFixHalo - w0 < kPairWin.)
        const bool in = i + kFixHalo < N;
        const uint64_t y = in ? win[i - w0] ^ win[i + kFixHalo - w0] : ~0ull;
        const uint64_t e16 = __ballot((y >> 48) == 0ull), e24 = __ballot((y >> 40) == 0ull);
        if ((threadIdx.x & 63) == 0) {
            if (e16 != 0ull) atomicAdd(&run_cnt[0], (uint32_t)__popcll(e16));
            if (e24 != 0ull) atomicAdd(&run_cnt[1], (uint32_t)__popcll(e24));
        }
    }
    const int lvl = (x == 0ull) ? kB : (__clzll((long long)x) >> 1);
    const int lowbits = 2 * (kB - lvl);  // 0..64
    const uint64_t lowmask = (lowbits >= 64) ? ~0ull : ((1ull << lowbits) - 1ull);
    const uint64_t lo_key = ki & ~lowmask, hi_key = ki | lowmask;

    // a = first key of the level-lvl cell holding ki (k[a] >= lo_key; k[i] >= lo_key)
    int a = i;
    bool far = false;
    if (act && x != 0ull) {
        far = w0 > 0 && win[0] >= lo_key;  // every window key is in the cell: the start is before it
        if (!far) a = lds_lower_bound(win, w0, wl, i, lo_key, false);
    }
    wave_searches(__ballot(far), keys, 0, w0 + 1, lo_key, false, &a);

    // owner: equal keys -> the first pair of the run; else the pair straddling
    // the end of the cell's first child (ki in the same child as k[a])
    bool owner = false;
    if (act) {
        if (x == 0ull) {
            owner = (i == 0) || (win[i - 1 - w0] != ki);
        } else {
            const int sh = lowbits - 2;
            const uint64_t ka = (a >= wl && a < wh) ? win[a - w0] : keys[a];
            owner = (ki >> sh) == (ka >> sh);
        }
    }
    // b = first index past the cell (k[b] > hi_key), or N
    int b = 0;
    far = false;
    if (owner) {
        far = wh < N && win[wh - 1 - w0] <= hi_key;  // every window key up to its end is in the cell
        if (!far) b = lds_lower_bound(win, w0, i + 1, wh, hi_key, true);
    }
    wave_searches(__ballot(far), keys, wh, N, hi_key, true, &b);

    if (owner) {
        atomicOr(&mask[a], 1ull << lvl);
        node_a[i] = a;
        node_b[i] = b;
        node_lvl[i] = lvl;
        const int ta = a / kTreeScanTile;
        if (ta >= (int)blockIdx.x - 1)
            atomicAdd(&own_cnt[ta - ((int)blockIdx.x - 1)], 1u);
        else
            atomicAdd(&tile_cnt[ta], 1u);
    } else if (act) {
        node_lvl[i] = -1;
    }
    // coordinate sums of this tile's points (fixed order: deterministic)
    const bool pt = i < N;
    const double sx = block_sum(pt ? (double)ys[i].x : 0.0, sh);
    const double sy = block_sum(pt ? (double)ys[i].y : 0.0, sh);  // its barriers also complete own_cnt
    if (threadIdx.x == 0) {
        part[2 * (size_t)blockIdx.x] = sx;
        part[2 * (size_t)blockIdx.x + 1] = sy;
        if (own_cnt[1]) atomicAdd(&tile_cnt[blockIdx.x], own_cnt[1]);
        if (own_cnt[0]) atomicAdd(&tile_cnt[blockIdx.x - 1], own_cnt[0]);
        runs[2 * (size_t)blockIdx.x] = run_cnt[0];
        runs[2 * (size_t)blockIdx.x + 1] = run_cnt[1];
    }
}

// ---------------------------------------------------------------- scan
// Pre-order positions and coordinate prefix sums, one 256-point tile per
// workgroup: the prefix of the earlier tiles from their partials (node
// counts exact in integers, coordinate sums in a fixed order), then the
// tile's own scan.

template <class T>
__global__ __launch_bounds__(256) void bh_scan_write(const unsigned long long* __restrict__ mask,
                                                     const typename Vec2<T>::type* __restrict__ ys, int N,
                                                     const uint32_t* __restrict__ tile_cnt,
                                                     const double* __restrict__ part, uint32_t* __restrict__ pre,
                                                     double* __restrict__ psx, double* __restrict__ psy, int nmain,
                                                     AttractSlice<T> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    __shared__ double sh[4];
    __shared__ double wsum[3][4];
    // prefix of earlier tiles: nodes = points + internal nodes starting there
    uint32_t bc = 0;
    double b1 = 0.0, b2 = 0.0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += 256) {
        bc += (uint32_t)kTreeScanTile + tile_cnt[b];
        b1 += part[2 * (size_t)b];
        b2 += part[2 * (size_t)b + 1];
    }
    const double b0 = block_sum((double)bc, sh);  // integers below 2^53: exact
    b1 = block_sum(b1, sh);
    b2 = block_sum(b2, sh);
    const int j = blockIdx.x * kTreeScanTile + threadIdx.x;
    const bool ok = j < N;
    const double c = ok ? (double)(__popcll(mask[j]) + 1) : 0.0;
    const double x = ok ? (double)ys[j].x : 0.0;
    const double y = ok ? (double)ys[j].y : 0.0;
    const double ic = wave_incl_scan_d(c), ix = wave_incl_scan_d(x), iy = wave_incl_scan_d(y);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 63) {
        wsum[0][w] = ic;
        wsum[1][w] = ix;
        wsum[2][w] = iy;
    }
    __syncthreads();
    double oc = b0, ox = b1, oy = b2;
    for (int ww = 0; ww < w; ww++) {
        oc += wsum[0][ww];
        ox += wsum[1][ww];
        oy += wsum[2][ww];
    }
    if (ok) {
        pre[j] = (uint32_t)(oc + ic - c);
        psx[j] = ox + ix - x;
        psy[j] = oy + iy - y;
        if (j == N - 1) {
            pre[N] = (uint32_t)(oc + ic);
            psx[N] = ox + ix;
            psy[N] = oy + iy;
        }
    }
}

// ---------------------------------------------------------------- emit
template <class T>
__global__ __launch_bounds__(256) void bh_emit(const unsigned long long* __restrict__ mask,
                                               const int* __restrict__ node_a, const int* __restrict__ node_b,
                                               const int* __restrict__ node_lvl,
                                               const typename Vec2<T>::type* __restrict__ ys,
                                               const uint32_t* __restrict__ perm, int N,
                                               const uint32_t* __restrict__ pre, const double* __restrict__ psx,
                                               const double* __restrict__ psy, BhNode<T>* __restrict__ nodes,
                                               const uint32_t* __restrict__ runs, uint32_t* __restrict__ runs_host,
                                               int nmain, AttractSlice<T> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    if (blockIdx.x == 0 && runs_host != nullptr) {  // the pairs kernel's counts, posted to the host for the
                                                     // next build's radix width
        __shared__ double sh[4];
        double c16 = 0.0, c24 = 0.0;
        for (int b = threadIdx.x; b < nmain; b += 256) {
            c16 += (double)runs[2 * (size_t)b];
            c24 += (double)runs[2 * (size_t)b + 1];
        }
        c16 = block_sum(c16, sh);
        c24 = block_sum(c24, sh);
        if (threadIdx.x == 0) {
            __hip_atomic_store(runs_host, (uint32_t)c16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(runs_host + 1, (uint32_t)c24, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    {  // leaf j
        const unsigned long long m = mask[j];
        const uint32_t t = pre[j] + (uint32_t)__popcll(m);
        BhNode<T> nd;
        nd.x = ys[j].x;
        nd.y = ys[j].y;
        nd.cl = 1u;
        nd.link = kLeafBit | perm[j];
        nodes[t] = nd;
    }
    if (j < N - 1) {
        const int lvl = node_lvl[j];
        if (lvl >= 0) {
            const int a = node_a[j], b = node_b[j];
            const unsigned long long m = mask[a];
            const uint32_t t = pre[a] + (uint32_t)__popcll(m & ((1ull << lvl) - 1ull));
            const double cnt = (double)(b - a);
            BhNode<T> nd;
            nd.x = (T)((psx[b] - psx[a]) / cnt);
            nd.y = (T)((psy[b] - psy[a]) / cnt);
            nd.cl = (uint32_t)(b - a) | ((uint32_t)lvl << kCountBits);
            nd.link = pre[b];
            nodes[t] = nd;
        }
    }
}

// ---------------------------------------------------------------- traverse
// One thread per point, in Morton order, so the lanes of a wave walk similar
// paths and their node gathers mostly hit the same cache lines.
//
// The walk is a chain of dependent loads; three things shorten it:
//  - the next node's index is known once the acceptance test is done, so its
//    load (one 16-byte node) is issued before the current node's division
//    and accumulation and arrives under them (a block of four nodes per
//    load, picked from registers, cost more in selects and load issue than
//    it saved: 57 against 44 us at N=100k);
//  - each workgroup first pulls its slice of its XCD's node range into L2
//    (below);
//  - workgroups are dispatched to the 8 XCDs round-robin, so workgroup b is
//    given the points of logical block x*q + min(x, r) + k (x = b % 8,
//    k = b / 8, grid = 8q + r): every XCD walks one contiguous Morton range,
//    whose cells are a small part of the tree, and its L2 holds them.
template <class T>
__global__ __launch_bounds__(kTravBlock) void bh_traverse(const BhNode<T>* __restrict__ nodes,
                                                   const uint32_t* __restrict__ pre,
                                                   const typename Vec2<T>::type* __restrict__ ys,
                                                   const uint32_t* __restrict__ perm, int N, T theta2,
                                                   const RootBox* __restrict__ root, T* __restrict__ neg,
                                                   double* __restrict__ zpart,
                                                   unsigned long long* __restrict__ interactions,
                                                   const uint32_t* __restrict__ sort_fault)
{
    __shared__ double sh[4];
    const int nb = (int)gridDim.x, xcd = (int)blockIdx.x & 7, kx = (int)blockIdx.x >> 3;
    const int lb = xcd * (nb >> 3) + min(xcd, nb & 7) + kx;
    const int j = lb * kTravBlock + threadIdx.x;
    // No walk over a tree whose Morton sort gave up (its nodes may hold
    // entries that no kernel wrote): forces stay unwritten, Z partials are
    // 0, and the host reports the fault word.
    const bool active = j < N && *sort_fault == 0u;
    double zd = 0.0;
    unsigned cnt = 0;
    const int total = (int)pre[N];
    const T size0sq = (T)(root->size0 * root->size0);
    // a cell of level l is accepted when (size0 2^-l)^2 / theta^2 < D: the
    // reference's size/sqrt(D) < theta, squared, with the division by theta^2
    // taken out of the walk (one multiply fewer before each node load)
    const T reach0 = size0sq / theta2;
    const uint32_t p = active ? perm[j] : 0u;
    T px = (T)0, py = (T)0;
    if (active) {
        px = ys[j].x;
        py = ys[j].y;
    }
    // The tree was written by the previous kernel through other XCDs' L2s,
    // so this XCD's L2 starts cold and every first touch of a node line waits
    // on the memory behind it.  The XCD's own points have their leaves and
    // finest cells in one pre-order range, [pre[j0], pre[j1]); its workgroups
    // first fetch one word of every 128-byte line of that range, each a slice,
    // with independent loads that overlap, so the walks' dependent node loads
    // mostly hit L2.  The fetched words are only consumed at the end.
    T warm = (T)0;
    {
        const int q = nb >> 3, r = nb & 7, count = q + (xcd < r ? 1 : 0);
        const int first = xcd * q + min(xcd, r);
        const int j0 = min(N, first * kTravBlock), j1 = min(N, (first + count) * kTravBlock);
        const int a = (int)pre[j0], b = (int)pre[j1];
        const int s0 = a + (int)((long long)(b - a) * kx / count);
        const int s1 = a + (int)((long long)(b - a) * (kx + 1) / count);
        constexpr int kPerLine = 128 / (int)sizeof(BhNode<T>);
        const int t = s0 + kPerLine * (int)threadIdx.x;
        if (t < s1) warm = nodes[t].x;
    }
    T fx = 0, fy = 0, z = 0;
    if (active) {
        BhNode<T> nd = nodes[0];
        int t = 0;
        while (t < total) {
            const T dx = px - nd.x, dy = py - nd.y;
            const T D = dx * dx + dy * dy;
            // leaf or cell by selects, not a branch (lanes at leaves and at
            // cells share every instruction; a branch costs more than the
            // cell test a leaf does not need)
            const int lvl = (int)(nd.cl >> kCountBits);
            const bool cell_acc = ldexp(reach0, -2 * lvl) < D;  // exact scaling by 4^-lvl
            const int sk = (int)nd.link;
            // leaves: sk < 0 (leaf bit 31); a cell's skip pointer lies past
            // it; a max and a select keep the choice off the scalar unit
            const int next = max(cell_acc ? sk : INT_MIN, t + 1);
            // the next node is requested as soon as its index is known, so
            // its load runs under this node's division and accumulation.
            // Unguarded: next <= total <= 2N - 1 (N leaves, at most N - 1
            // internal nodes) lies inside the 2N-node buffer, and a load at
            // next == total ends the walk unused (a guard cost 1 us of 38).
            // A 32-bit byte offset lets the load take the node base from
            // scalar registers.
            static_assert(2ull * kMaxTreePoints * sizeof(BhNode<T>) <= 0xffffffffull, "node offsets fit 32 bits");
            const uint32_t link = nd.link, cl = nd.cl;
            nd = *reinterpret_cast<const BhNode<T>*>(reinterpret_cast<const char*>(nodes) +
                                                     (uint32_t)next * (uint32_t)sizeof(BhNode<T>));
            // nothing of the rest may be scheduled ahead of the load
            __builtin_amdgcn_sched_barrier(0);
            const bool leaf = (link & kLeafBit) != 0u;
            const bool acc = leaf ? (link & ~kLeafBit) != p : cell_acc;
            const uint32_t count = cl & ((1u << kCountBits) - 1u);
            if (acc) {  // a leaf's count is 1: count*q and count*q*q are exactly q and q*q
                // FP32: q = 1/(1 + D) as the hardware reciprocal (1 ulp) with
                // one Newton-Raphson step, r + r(1 - a r); 3 instructions
                // against 11 for the IEEE division (the N=100k swiss-roll
                // embedding after 430 iterations came out bit-identical to the
                // one with the division).  a is clamped to
                // FLT_MAX so that an overflowing D gives q = 0 rather than
                // inf * 0 = NaN in the residual.
                T q;
                if constexpr (std::is_same<T, float>::value) {
                    const float a = fminf(1.0f + D, FLT_MAX);
                    const float r = __builtin_amdgcn_rcpf(a);
                    q = fmaf(r, fmaf(-a, r, 1.0f), r);
                } else {
                    q = (T)1 / ((T)1 + D);
                }
                const T cq = (T)count * q;
                z += cq;
                const T m = cq * q;
                fx += m * dx;
                fy += m * dy;
                cnt++;
            }
            t = next;
        }
    }
    if (active) {
        neg[2 * (size_t)p] = fx;
        neg[2 * (size_t)p + 1] = fy;
        zd = (double)z;
    }
    asm volatile("" ::"v"(warm));  // keeps the warming loads
    zd = block_sum(zd, sh);
    if (threadIdx.x == 0) zpart[lb] = zd;
    if (interactions != nullptr) {  // per workgroup: one atomic per workgroup would serialise 10^3 of them
        const double c = block_sum((double)cnt, sh);
        if (threadIdx.x == 0) interactions[1 + lb] = (unsigned long long)c;
    }
}

}  // namespace

template <class T>
void BhTree<T>::alloc(int n, MemBudget* mb)
{
    require(n <= kMaxTreePoints, TKA_E_UNSUPPORTED, "Barnes-Hut tree: at most 2^26 - 1 points");
    N = n;
    keys.alloc(n, mb);
    keys_alt.alloc(n, mb);
    perm.alloc(n, mb);
    perm_alt.alloc(n, mb);
    kfull.alloc(n, mb);
    passes = n <= (1 << 22) ? 3 : 4;
    runs.alloc(2 * (size_t)ceil_div(n, 256), mb);
    if (!runs_host) {
        void* h = nullptr;
        TKA_HIP(hipHostMalloc(&h, 2 * sizeof(uint32_t), hipHostMallocDefault));
        runs_host = static_cast<uint32_t*>(h);
    }
    // no counts yet: the first build sorts 3 radix passes
    __atomic_store_n(runs_host, 1u, __ATOMIC_RELAXED);
    __atomic_store_n(runs_host + 1, 0u, __ATOMIC_RELAXED);
    k64.alloc(n, mb);
    p2.alloc(n, mb);
    sort_scratch.alloc(radix_sort_scratch_words(n, kRsItemsCounted), mb);
    // on the stream the buffer was allocated on (ordered after the allocation)
    TKA_HIP(hipMemsetAsync(sort_scratch.get(), 0, sort_scratch.size() * sizeof(uint32_t), alloc_stream()));
    mask.alloc(n, mb);
    node_a.alloc(n, mb);
    node_b.alloc(n, mb);
    node_lvl.alloc(n, mb);
    ys.alloc(n, mb);
    pre.alloc(n + 1, mb);
    sx.alloc(n + 1, mb);
    sy.alloc(n + 1, mb);
    part.alloc(2 * (size_t)ceil_div(n, kTreeScanTile), mb);
    tile_cnt.alloc(ceil_div(n, kTreeScanTile), mb);
    nodes.alloc(2 * (size_t)n, mb);  // > the node count (<= 2n - 1): the walk's look-ahead load stays inside
    root.alloc(1, mb);
}

template <class T>
void BhTree<T>::carve(int n, Arena& a)
{
    require(n <= kMaxTreePoints, TKA_E_UNSUPPORTED, "Barnes-Hut tree: at most 2^26 - 1 points");
    N = n;
    a.take(keys, n);
    a.take(keys_alt, n);
    a.take(perm, n);
    a.take(perm_alt, n);
    a.take(kfull, n);
    passes = n <= (1 << 22) ? 3 : 4;
    a.take(runs, 2 * (size_t)ceil_div(n, 256));
    runs_host = nullptr;
    a.take(k64, n);
    a.take(p2, n);
    a.take(sort_scratch, radix_sort_scratch_words(n, kRsItemsCounted));
    a.take(mask, n);
    a.take(node_a, n);
    a.take(node_b, n);
    a.take(node_lvl, n);
    a.take(ys, n);
    a.take(pre, n + 1);
    a.take(sx, n + 1);
    a.take(sy, n + 1);
    a.take(part, 2 * (size_t)ceil_div(n, kTreeScanTile));
    a.take(tile_cnt, ceil_div(n, kTreeScanTile));
    a.take(nodes, 2 * (size_t)n);
    a.take(root, 1);
}

static __global__ __launch_bounds__(256) void fill_words(uint32_t* __restrict__ p, int n, uint32_t v)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = v;
}

template <class T>
void BhTree<T>::clear_sort_state(hipStream_t s)
{
    // the sort keeps its per-pass counts and tickets cleared after each
    // sort and clears the tile status words while counting; the fault word
    // and those counters must start at zero (a fill kernel, not a memset
    // node, so that the clearing can be captured in a graph)
    const int n = (int)kRsHeader;
    hipLaunchKernelGGL(fill_words, dim3(ceil_div(n, 256)), dim3(256), 0, s, sort_scratch.get(), n, 0u);
    TKA_LAUNCH_CHECK();
}

template <class T>
void bbox_partials_device(const T* Y, int N, double* part, hipStream_t s)
{
    hipLaunchKernelGGL(bbox_part<T>, dim3(bbox_blocks(N)), dim3(kBoxBlock), 0, s, Y, N, part);
    TKA_LAUNCH_CHECK();
}

template <class T>
void bh_build_keys(T* Y, int N, const double* box_part, int nbox, BhTree<T>& tr, hipStream_t s)
{
    require(N >= 2 && tr.N >= N, TKA_E_INTERNAL, "tree storage too small");
    if (N <= (1 << 22) && tr.runs_host) {
        // 2 passes while no run of equal top 16 bits is longer than kFixHalo
        // (the rank sort takes every run); 3 while runs of equal top 24 bits
        // leave fewer than kWaveRun - kFixHalo keys in long runs (no run is
        // then longer than a wave sorts in LDS); else 4
        const uint32_t c16 = __atomic_load_n(tr.runs_host, __ATOMIC_RELAXED);
        const uint32_t c24 = __atomic_load_n(tr.runs_host + 1, __ATOMIC_RELAXED);
        tr.passes = c16 == 0u ? 2 : (c24 < (uint32_t)(kWaveRun - kFixHalo) ? 3 : 4);
    }
    if (tr.forced_passes > 0) tr.passes = tr.forced_passes;
    const int tiles = (int)radix_tiles((size_t)N, kRsItemsCounted);
    hipLaunchKernelGGL(bh_keys<T>, dim3(tiles), dim3(kRsBlock), 0, s, Y, N, box_part, nbox, tr.root.get(),
                       tr.passes, tr.keys.get(), tr.kfull.get(), tr.perm.get(), tr.mask.get(), tr.tile_cnt.get(),
                       rs_layout(tr.sort_scratch.get()), tiles);
    TKA_LAUNCH_CHECK();
}

template <class T>
void bh_build(T* Y, int N, const double* box_part, int nbox, BhTree<T>& tr, hipStream_t s)
{
    bh_build_keys<T>(Y, N, box_part, nbox, tr, s);
    bh_build_tree<T>(Y, N, tr, s);
}

template <class T>
void bh_build_tree(const T* Y, int N, BhTree<T>& tr, hipStream_t s, const AttractSlice<T>* attract)
{
    const unsigned g = ceil_div(N, 256);
    // The attraction rows are split over the sort passes and the node kernels
    // in proportion to their durations at N = 100k (a pass 12 us; pairs 10,
    // scan 5, emit 4), one row group per extra workgroup, so each launch's
    // extra workgroups finish within it on the CUs its own workgroups leave
    // idle.  bh_fix takes none: its 48 KB of LDS per workgroup would hold
    // back the extra workgroups.  Other splits (passes 8 or 16; node kernels
    // 14/7/6) measured the same within 0.5%.
    const int passes = N > 1 ? tr.passes : 0;
    constexpr int kWPass = 12, kWPass0 = 8, kWFix = 0, kWPairs = 10, kWScan = 5, kWEmit = 4;
    std::vector<int> weight(passes, kWPass);
    // pass 0 without look-back (few tiles) is the shorter one
    if (passes > 0 && radix_tiles((size_t)N, kRsItemsCounted) <= (size_t)kRsDirectTiles) weight[0] = kWPass0;
    weight.insert(weight.end(), {kWFix, kWPairs, kWScan, kWEmit});
    const int nk = (int)weight.size();
    std::vector<AttractSlice<T>> slice(nk);
    std::vector<int> wgs(nk, 0);
    if (attract) {
        int wsum = 0, wacc = 0;
        for (int w : weight) wsum += w;
        const int r0 = attract->row0, rows = attract->row1 - attract->row0;
        for (int k = 0; k < nk; k++) {
            slice[k] = *attract;
            slice[k].row0 = r0 + (int)((int64_t)rows * wacc / wsum);
            wacc += weight[k];
            slice[k].row1 = r0 + (int)((int64_t)rows * wacc / wsum);
            const int groups = (int)ceil_div((size_t)std::max(0, slice[k].row1 - slice[k].row0), kRowsPerBlock);
            wgs[k] = groups;
        }
    }
    bool alt = false;
    if (passes > 0)
        alt = radix_sort_pairs_counted_side<uint32_t, uint32_t, AttractSlice<T>>(
            tr.keys.get(), tr.perm.get(), tr.keys_alt.get(), tr.perm_alt.get(), N, 8 * passes, tr.sort_scratch.get(),
            slice.data(), wgs.data(), s);
    const uint32_t* hk = alt ? tr.keys_alt.get() : tr.keys.get();
    const uint32_t* hp = alt ? tr.perm_alt.get() : tr.perm.get();
    const int kf = passes, kp = passes + 1, ks = passes + 2, ke = passes + 3;
    hipLaunchKernelGGL(bh_fix<T>, dim3(g + wgs[kf]), dim3(256), 0, s, hk, hp, tr.kfull.get(), Y, N, tr.k64.get(),
                       tr.p2.get(), tr.ys.get(), (int)g, slice[kf]);
    TKA_LAUNCH_CHECK();
    const uint32_t* p = tr.p2.get();
    tr.sorted_perm = p;
    static_assert(kTreeScanTile == 256, "one scan tile per pairs workgroup");
    hipLaunchKernelGGL(bh_pairs<T>, dim3(g + wgs[kp]), dim3(256), 0, s, tr.k64.get(), tr.ys.get(), N, tr.mask.get(),
                       tr.node_a.get(), tr.node_b.get(), tr.node_lvl.get(), tr.tile_cnt.get(), tr.part.get(),
                       tr.runs.get(), (int)g, slice[kp]);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(bh_scan_write<T>, dim3(g + wgs[ks]), dim3(256), 0, s, tr.mask.get(), tr.ys.get(), N,
                       tr.tile_cnt.get(), tr.part.get(), tr.pre.get(), tr.sx.get(), tr.sy.get(), (int)g, slice[ks]);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(bh_emit<T>, dim3(g + wgs[ke]), dim3(256), 0, s, tr.mask.get(), tr.node_a.get(),
                       tr.node_b.get(), tr.node_lvl.get(), tr.ys.get(), p, N, tr.pre.get(), tr.sx.get(), tr.sy.get(),
                       tr.nodes.get(), tr.runs.get(), tr.runs_host, (int)g, slice[ke]);
    TKA_LAUNCH_CHECK();
}

template <class T>
int bh_repulsion(const BhTree<T>& tr, const T* Y, double theta, T* neg, double* zpart,
                 unsigned long long* interactions, hipStream_t s)
{
    (void)Y;
    const int g = bh_zblocks(tr.N);
    hipLaunchKernelGGL((bh_traverse<T>), dim3(g), dim3(kTravBlock), 0, s, tr.nodes.get(), tr.pre.get(), tr.ys.get(),
                       tr.sorted_perm, tr.N, (T)(theta * theta), tr.root.get(), neg, zpart, interactions,
                       tr.sort_fault());
    TKA_LAUNCH_CHECK();
    return g;
}

static __global__ __launch_bounds__(256) void add_interactions(unsigned long long* counts, int n, int times)
{
    __shared__ unsigned long long sh[256];
    unsigned long long c = 0;
    for (int i = threadIdx.x; i < n; i += 256) c += counts[1 + i];
    sh[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < 256; i++) t += sh[i];
        counts[0] += t * (unsigned long long)times;
    }
}

void bh_add_interactions(unsigned long long* counts, int nblocks, int launches, hipStream_t s)
{
    hipLaunchKernelGGL(add_interactions, dim3(1), dim3(256), 0, s, counts, nblocks, launches);
    TKA_LAUNCH_CHECK();
}

template <class T>
bool bh_sort_failed(const BhTree<T>& tr, hipStream_t s)
{
    return radix_sort_fault(tr.sort_scratch.get(), s) != 0u;
}

template <class T>
void bh_inject_sort_fault(BhTree<T>& tr, hipStream_t s)
{
    const uint32_t one = 1u;
    TKA_HIP(hipMemcpyAsync(tr.sort_scratch.get() + 1, &one, sizeof(one), hipMemcpyHostToDevice, s));
    TKA_HIP(hipStreamSynchronize(s));
}

template struct BhTree<float>;
template struct BhTree<double>;
template bool bh_sort_failed<float>(const BhTree<float>&, hipStream_t);
template bool bh_sort_failed<double>(const BhTree<double>&, hipStream_t);
template void bh_inject_sort_fault<float>(BhTree<float>&, hipStream_t);
template void bh_inject_sort_fault<double>(BhTree<double>&, hipStream_t);
template void bbox_partials_device<float>(const float*, int, double*, hipStream_t);
template void bbox_partials_device<double>(const double*, int, double*, hipStream_t);
template void bh_build<float>(float*, int, const double*, int, BhTree<float>&, hipStream_t);
template void bh_build<double>(double*, int, const double*, int, BhTree<double>&, hipStream_t);
template void bh_build_keys<float>(float*, int, const double*, int, BhTree<float>&, hipStream_t);
template void bh_build_keys<double>(double*, int, const double*, int, BhTree<double>&, hipStream_t);
template void bh_build_tree<float>(const float*, int, BhTree<float>&, hipStream_t, const AttractSlice<float>*);
template void bh_build_tree<double>(const double*, int, BhTree<double>&, hipStream_t, const AttractSlice<double>*);
template int bh_repulsion<float>(const BhTree<float>&, const float*, double, float*, double*, unsigned long long*,
                                 hipStream_t);
template int bh_repulsion<double>(const BhTree<double>&, const double*, double, double*, double*,
                                  unsigned long long*, hipStream_t);

}  // namespace tka
