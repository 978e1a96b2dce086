// Barnes-Hut octree (3-D embeddings) from Morton keys, and the per-point
// traversal.  The construction follows the quadtree's (bh.hip) with a third
// axis:
//   box    bounding-box partials per 256 points, reduced in a fixed order by
//          one workgroup to the root cell (mean, half-widths
//          max(max - mean, mean - min) + 1e-5 per axis: quadtree.hpp:104-129
//          with three axes).
//   keys   point -> 63-bit Morton key of its level-21 cell, left-aligned in
//          64 bits, digit 4 zbit + 2 ybit + xbit per level, so sorting keys
//          gives the octree's depth-first child order; quantisation
//          ceil(t * 2^21) - 1 puts a point on a split plane in the lower half
//          (the closed-interval containment tried in child order).
//   sort   stable LSD radix sort of the full keys (8 passes).
//   pairs  adjacent sorted keys i, i+1 first differ at octree level
//          l = clz64(k_i ^ k_i+1) / 3; the level-l cell holding both has
//          >= 2 non-empty children: a node of the compressed tree, owned by
//          the pair that straddles the end of its first child.  Chains of
//          one-child cells are skipped (same count and centre of mass as the
//          child, a larger size: accepted only if the child would be).
//   scan   pre-order positions and FP64 prefix sums of the sorted
//          coordinates, as in the quadtree.
//   Equal keys (points closer than 2^-21 of the box on every axis) form one
//   level-21 node whose children are the individual points.
// Traversal is the quadtree's stackless walk: visit node t; a leaf interacts
// unless it is the point itself, then t+1; an accepted cell interacts and
// jumps to its skip pointer; otherwise t+1.
#include "bh3.hpp"

#include <cfloat>
#include <climits>
#include <type_traits>

#include "radix_sort.hpp"
#include "tree_search.hpp"

namespace tka {
namespace {

constexpr int kB3 = kMortonBits3;
constexpr int kTile3 = 256;  // points per scan tile = per pairs workgroup
constexpr int kPairWin3 = 3 * 256;

// ---------------------------------------------------------------- box
template <class T>
__global__ __launch_bounds__(kBox3Block) void bbox3_part(const T* __restrict__ Y, int N, double* __restrict__ part)
{
    __shared__ double sh[9][kBox3Block / 64];
    const int i = blockIdx.x * kBox3Block + threadIdx.x;
    double v[9] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    if (i < N) {
        for (int d = 0; d < 3; d++) {
            const double c = (double)Y[3 * (size_t)i + d];
            v[d] = c;
            v[3 + d] = c;
            v[6 + d] = c;
        }
    }
    for (int d = 0; d < 3; d++) {
        v[d] = wave_sum(v[d]);
        v[3 + d] = wave_min(v[3 + d]);
        v[6 + d] = wave_max(v[6 + d]);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 9; k++) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[9] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int ww = 0; ww < kBox3Block / 64; ww++)
            for (int d = 0; d < 3; d++) {
                r[d] += sh[d][ww];
                r[3 + d] = fmin(r[3 + d], sh[3 + d][ww]);
                r[6 + d] = fmax(r[6 + d], sh[6 + d][ww]);
            }
        double* o = part + (size_t)kBox3Stride * blockIdx.x;
        for (int k = 0; k < 9; k++) o[k] = r[k];
        o[9] = 0.0;
    }
}

// Root cell from the nb partials, one workgroup, fixed order.
__global__ __launch_bounds__(256) void bh3_root(const double* __restrict__ part, int nb, int N,
                                                RootBox3* __restrict__ root)
{
    __shared__ double sh[9][4];
    double v[9] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int b = threadIdx.x; b < nb; b += 256) {
        const double* p = part + (size_t)kBox3Stride * b;
        for (int d = 0; d < 3; d++) {
            v[d] += p[d];
            v[3 + d] = fmin(v[3 + d], p[3 + d]);
            v[6 + d] = fmax(v[6 + d], p[6 + d]);
        }
    }
    for (int d = 0; d < 3; d++) {
        v[d] = wave_sum(v[d]);
        v[3 + d] = wave_min(v[3 + d]);
        v[6 + d] = wave_max(v[6 + d]);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 9; k++) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[9] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int ww = 0; ww < 4; ww++)
            for (int d = 0; d < 3; d++) {
                r[d] += sh[d][ww];
                r[3 + d] = fmin(r[3 + d], sh[3 + d][ww]);
                r[6 + d] = fmax(r[6 + d], sh[6 + d][ww]);
            }
        RootBox3 rb;
        double s0 = 0.0;
        for (int d = 0; d < 3; d++) {
            rb.mean[d] = r[d] / (double)N;
            rb.h[d] = fmax(r[6 + d] - rb.mean[d], rb.mean[d] - r[3 + d]) + 1e-5;
            s0 = fmax(s0, rb.h[d]);
        }
        rb.size0 = s0;
        *root = rb;
    }
}

// ---------------------------------------------------------------- keys
// 21 bits of v to every third bit of a 64-bit word (bit k -> bit 3k).
TKA_DEV uint64_t spread21(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Quantise t in (0, 1) to ceil(t * 2^21) - 1 in [0, 2^21 - 1].
TKA_DEV uint32_t quantise21(double t)
{
    double c = ceil(ldexp(t, kB3)) - 1.0;
    c = c < 0.0 ? 0.0 : (c > 2097151.0 ? 2097151.0 : c);
    return (uint32_t)c;
}

template <class T>
__global__ __launch_bounds__(256) void bh3_keys(T* __restrict__ Y, int N, const RootBox3* __restrict__ root,
                                                uint64_t* __restrict__ keys, uint32_t* __restrict__ ktop,
                                                uint32_t* __restrict__ perm, unsigned long long* __restrict__ mask,
                                                uint32_t* __restrict__ tile_cnt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < (N + kTile3 - 1) / kTile3) tile_cnt[i] = 0u;
    if (i >= N) return;
    const RootBox3 rb = *root;
    uint64_t key = 0ull;
    for (int d = 0; d < 3; d++) {
        const T c = (T)((double)Y[3 * (size_t)i + d] - rb.mean[d]);
        Y[3 * (size_t)i + d] = c;
        key |= spread21(quantise21(((double)c + rb.h[d]) / (2.0 * rb.h[d]))) << d;
    }
    keys[i] = key << 1;
    ktop[i] = (uint32_t)(key >> 39);  // the top 24 bits of the left-aligned key: levels 1-8
    perm[i] = (uint32_t)i;
    mask[i] = 0ull;
}

// Full keys in sorted order, the final point order and the sorted
// coordinates: the radix sort ordered the points by their top 24 key bits,
// and each run of equal top bits is put in (full key, point) order
// (run_sort_tile), which equals a stable sort of the full keys.
template <class T>
__global__ __launch_bounds__(256) void bh3_runs(const uint32_t* __restrict__ ktop, const uint32_t* __restrict__ perm,
                                                const uint64_t* __restrict__ keys, const T* __restrict__ Y, int N,
                                                uint64_t* __restrict__ k64, uint32_t* __restrict__ p2,
                                                T* __restrict__ ys)
{
    run_sort_tile(ktop, perm, keys, N, k64, p2, [&](int d, uint32_t p) {
        for (int c = 0; c < 3; c++) ys[3 * (size_t)d + c] = Y[3 * (size_t)p + c];
    });
}

// ---------------------------------------------------------------- pairs
// Keys [w0, w0 + kPairWin3) of the sorted array are staged in LDS for the
// searches of deep cells; a search whose answer lies outside the window is
// done by the whole wave over global memory (tree_search.hpp).  Also leaves
// the scan's tile partials: the FP64 sums of the tile's coordinates and the
// number of internal nodes whose range starts in each tile.
template <class T>
__global__ __launch_bounds__(256) void bh3_pairs(const uint64_t* __restrict__ keys, const T* __restrict__ ys, int N,
                                                 unsigned long long* __restrict__ mask, int* __restrict__ node_a,
                                                 int* __restrict__ node_b, int* __restrict__ node_lvl,
                                                 uint32_t* __restrict__ tile_cnt, double* __restrict__ part,
                                                 int nmain, AttractSlice<T, 3> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    __shared__ uint64_t win[kPairWin3];
    __shared__ uint32_t own_cnt[2];  // owners with a in tile blockIdx.x - 1, blockIdx.x
    __shared__ double sh[4];
    const int w0 = (int)blockIdx.x * 256 - 256;
    for (int t = threadIdx.x; t < kPairWin3; t += 256) {
        const int g = w0 + t;
        win[t] = (g >= 0 && g < N) ? keys[g] : 0ull;
    }
    if (threadIdx.x < 2) own_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int wl = max(w0, 0), wh = min(w0 + kPairWin3, N);
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool act = i < N - 1;  // pair (i, i+1); every lane stays for the wave's searches
    const uint64_t ki = act ? win[i - w0] : 0ull, kj = act ? win[i + 1 - w0] : 0ull;
    const uint64_t x = ki ^ kj;
    const int lvl = (x == 0ull) ? kB3 : (__clzll((long long)x) / 3);
    const int lowbits = 64 - 3 * lvl;  // the cell's key bits below its prefix: 1..64
    const uint64_t lowmask = (lowbits >= 64) ? ~0ull : ((1ull << lowbits) - 1ull);
    const uint64_t lo_key = ki & ~lowmask, hi_key = ki | lowmask;

    int a = i;
    bool far = false;
    if (act && x != 0ull) {
        far = w0 > 0 && win[0] >= lo_key;
        if (!far) a = lds_lower_bound(win, w0, wl, i, lo_key, false);
    }
    wave_searches(__ballot(far), keys, 0, w0 + 1, lo_key, false, &a);

    bool owner = false;
    if (act) {
        if (x == 0ull) {
            owner = (i == 0) || (win[i - 1 - w0] != ki);
        } else {
            const int sh3 = lowbits - 3;  // the child digit of the level-lvl cell
            const uint64_t ka = (a >= wl && a < wh) ? win[a - w0] : keys[a];
            owner = (ki >> sh3) == (ka >> sh3);
        }
    }
    int b = 0;
    far = false;
    if (owner) {
        far = wh < N && win[wh - 1 - w0] <= hi_key;
        if (!far) b = lds_lower_bound(win, w0, i + 1, wh, hi_key, true);
    }
    wave_searches(__ballot(far), keys, wh, N, hi_key, true, &b);

    if (owner) {
        atomicOr(&mask[a], 1ull << lvl);
        node_a[i] = a;
        node_b[i] = b;
        node_lvl[i] = lvl;
        const int ta = a / kTile3;
        if (ta >= (int)blockIdx.x - 1)
            atomicAdd(&own_cnt[ta - ((int)blockIdx.x - 1)], 1u);
        else
            atomicAdd(&tile_cnt[ta], 1u);
    } else if (act) {
        node_lvl[i] = -1;
    }
    const bool pt = i < N;
    double s3[3];
    for (int d = 0; d < 3; d++) s3[d] = block_sum(pt ? (double)ys[3 * (size_t)i + d] : 0.0, sh);
    if (threadIdx.x == 0) {  // block_sum's barriers also completed own_cnt
        for (int d = 0; d < 3; d++) part[3 * (size_t)blockIdx.x + d] = s3[d];
        if (own_cnt[1]) atomicAdd(&tile_cnt[blockIdx.x], own_cnt[1]);
        if (own_cnt[0]) atomicAdd(&tile_cnt[blockIdx.x - 1], own_cnt[0]);
    }
}

// ---------------------------------------------------------------- scan
template <class T>
__global__ __launch_bounds__(256) void bh3_scan(const unsigned long long* __restrict__ mask, const T* __restrict__ ys,
                                                int N, const uint32_t* __restrict__ tile_cnt,
                                                const double* __restrict__ part, uint32_t* __restrict__ pre,
                                                double* __restrict__ ps, int nmain, AttractSlice<T, 3> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    __shared__ double sh[4];
    __shared__ double wsum[4][4];
    uint32_t bc = 0;
    double bs[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += 256) {
        bc += (uint32_t)kTile3 + tile_cnt[b];
        for (int d = 0; d < 3; d++) bs[d] += part[3 * (size_t)b + d];
    }
    const double b0 = block_sum((double)bc, sh);  // integers below 2^53: exact
    for (int d = 0; d < 3; d++) bs[d] = block_sum(bs[d], sh);
    const int j = blockIdx.x * kTile3 + threadIdx.x;
    const bool ok = j < N;
    const double c = ok ? (double)(__popcll(mask[j]) + 1) : 0.0;
    double v[3], iv[3];
    for (int d = 0; d < 3; d++) {
        v[d] = ok ? (double)ys[3 * (size_t)j + d] : 0.0;
        iv[d] = wave_incl_scan_d(v[d]);
    }
    const double ic = wave_incl_scan_d(c);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 63) {
        wsum[0][w] = ic;
        for (int d = 0; d < 3; d++) wsum[1 + d][w] = iv[d];
    }
    __syncthreads();
    double oc = b0, o[3] = {bs[0], bs[1], bs[2]};
    for (int ww = 0; ww < w; ww++) {
        oc += wsum[0][ww];
        for (int d = 0; d < 3; d++) o[d] += wsum[1 + d][ww];
    }
    if (ok) {
        pre[j] = (uint32_t)(oc + ic - c);
        for (int d = 0; d < 3; d++) ps[(size_t)d * (N + 1) + j] = o[d] + iv[d] - v[d];
        if (j == N - 1) {
            pre[N] = (uint32_t)(oc + ic);
            for (int d = 0; d < 3; d++) ps[(size_t)d * (N + 1) + N] = o[d] + iv[d];
        }
    }
}

// ---------------------------------------------------------------- emit
template <class T>
__global__ __launch_bounds__(256) void bh3_emit(const unsigned long long* __restrict__ mask,
                                                const int* __restrict__ node_a, const int* __restrict__ node_b,
                                                const int* __restrict__ node_lvl, const T* __restrict__ ys,
                                                const uint32_t* __restrict__ perm, int N,
                                                const uint32_t* __restrict__ pre, const double* __restrict__ ps,
                                                BhNode3<T>* __restrict__ nodes, int nmain, AttractSlice<T, 3> side)
{
    if ((int)blockIdx.x >= nmain) {  // attraction rows riding in this launch
        side.run((int)blockIdx.x - nmain, (int)gridDim.x - nmain);
        return;
    }
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    {  // leaf j
        const unsigned long long m = mask[j];
        const uint32_t t = pre[j] + (uint32_t)__popcll(m);
        BhNode3<T> nd;
        nd.x = ys[3 * (size_t)j];
        nd.y = ys[3 * (size_t)j + 1];
        nd.z = ys[3 * (size_t)j + 2];
        nd.cl = 1u;
        nd.link = kLeafBit | min(perm[j], (uint32_t)(N - 1));
        nodes[t] = nd;
    }
    if (j < N - 1) {
        const int lvl = node_lvl[j];
        if (lvl >= 0) {
            const int a = node_a[j], b = node_b[j];
            const unsigned long long m = mask[a];
            const uint32_t t = pre[a] + (uint32_t)__popcll(m & ((1ull << lvl) - 1ull));
            const double cnt = (double)(b - a);
            const size_t s = (size_t)N + 1;
            BhNode3<T> nd;
            nd.x = (T)((ps[b] - ps[a]) / cnt);
            nd.y = (T)((ps[s + b] - ps[s + a]) / cnt);
            nd.z = (T)((ps[2 * s + b] - ps[2 * s + a]) / cnt);
            nd.cl = (uint32_t)(b - a) | ((uint32_t)lvl << kCountBits);
            nd.link = pre[b];
            nodes[t] = nd;
        }
    }
}

// ---------------------------------------------------------------- traverse
// One thread per point, in Morton order; workgroup b walks the points of
// logical block x*q + min(x, r) + k (x = b % 8, k = b / 8, grid = 8q + r),
// so every XCD walks one contiguous Morton range and its L2 holds the cells.
template <class T>
__global__ __launch_bounds__(kTravBlock) void bh3_traverse(const BhNode3<T>* __restrict__ nodes,
                                                    const uint32_t* __restrict__ pre, const T* __restrict__ ys,
                                                    const uint32_t* __restrict__ perm, int N, T theta2,
                                                    const RootBox3* __restrict__ root, T* __restrict__ neg,
                                                    double* __restrict__ zpart,
                                                    unsigned long long* __restrict__ interactions,
                                                    const uint32_t* __restrict__ sort_fault)
{
    __shared__ double sh[4];
    const int nb = (int)gridDim.x, xcd = (int)blockIdx.x & 7, kx = (int)blockIdx.x >> 3;
    const int lb = xcd * (nb >> 3) + min(xcd, nb & 7) + kx;
    const int j = lb * kTravBlock + threadIdx.x;
    // no walk over a tree whose Morton sort gave up: forces stay unwritten,
    // Z partials are 0, and the host reports the fault word
    const bool active = j < N && *sort_fault == 0u;
    const int total = (int)pre[N];
    const T size0sq = (T)(root->size0 * root->size0);
    const T reach0 = size0sq / theta2;  // a level-l cell is accepted when reach0 * 4^-l < D
    const uint32_t p = active ? min(perm[j], (uint32_t)(N - 1)) : 0u;
    T px = 0, py = 0, pz = 0;
    if (active) {
        px = ys[3 * (size_t)j];
        py = ys[3 * (size_t)j + 1];
        pz = ys[3 * (size_t)j + 2];
    }
    T fx = 0, fy = 0, fz = 0, z = 0;
    unsigned cnt = 0;
    if (active) {
        BhNode3<T> nd = nodes[0];
        int t = 0;
        while (t < total) {
            const T dx = px - nd.x, dy = py - nd.y, dz = pz - nd.z;
            const T D = dx * dx + dy * dy + dz * dz;
            const int lvl = (int)(nd.cl >> kCountBits);
            const bool cell_acc = ldexp(reach0, -2 * lvl) < D;
            const int sk = (int)nd.link;
            // leaves have bit 31 set (sk < 0): next = t + 1 for them
            const int next = max(cell_acc ? sk : INT_MIN, t + 1);
            const uint32_t link = nd.link, cl = nd.cl;
            // next <= total <= 2N - 1 lies inside the 2N-node buffer; a load
            // at next == total ends the walk unused
            nd = nodes[next];
            const bool leaf = (link & kLeafBit) != 0u;
            const bool acc = leaf ? (link & ~kLeafBit) != p : cell_acc;
            const uint32_t count = cl & ((1u << kCountBits) - 1u);
            if (acc) {
                T q;
                if constexpr (std::is_same<T, float>::value) {
                    // hardware reciprocal (1 ulp) and one Newton-Raphson step;
                    // a is clamped so that an overflowing D gives q = 0
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
                fz += m * dz;
                cnt++;
            }
            t = next;
        }
    }
    double zd = 0.0;
    if (active) {
        neg[3 * (size_t)p] = fx;
        neg[3 * (size_t)p + 1] = fy;
        neg[3 * (size_t)p + 2] = fz;
        zd = (double)z;
    }
    zd = block_sum(zd, sh);
    if (threadIdx.x == 0) zpart[lb] = zd;
    if (interactions != nullptr) {  // per workgroup, as the 2-D walk
        const double c = block_sum((double)cnt, sh);
        if (threadIdx.x == 0) interactions[1 + lb] = (unsigned long long)c;
    }
}

}  // namespace

template <class T>
void BhTree3<T>::alloc(int n, MemBudget* mb)
{
    require(n >= 2 && n <= kMaxTreePoints, TKA_E_UNSUPPORTED, "Barnes-Hut octree: 2 to 2^26 - 1 points");
    N = n;
    keys.alloc(n, mb);
    keys_sorted.alloc(n, mb);
    ktop.alloc(n, mb);
    ktop_alt.alloc(n, mb);
    perm.alloc(n, mb);
    perm_alt.alloc(n, mb);
    p2.alloc(n, mb);
    sort_scratch.alloc(radix_sort_scratch_words(n), mb);
    // on the stream the buffer was allocated on (ordered after the allocation)
    TKA_HIP(hipMemsetAsync(sort_scratch.get(), 0, sort_scratch.size() * sizeof(uint32_t), alloc_stream()));
    ys.alloc(3 * (size_t)n, mb);
    mask.alloc(n, mb);
    node_a.alloc(n, mb);
    node_b.alloc(n, mb);
    node_lvl.alloc(n, mb);
    pre.alloc((size_t)n + 1, mb);
    ps.alloc(3 * ((size_t)n + 1), mb);
    part.alloc(3 * (size_t)ceil_div(n, kTile3), mb);
    tile_cnt.alloc(ceil_div(n, kTile3), mb);
    box.alloc((size_t)kBox3Stride * ceil_div(n, kBox3Block), mb);
    nodes.alloc(2 * (size_t)n, mb);  // > the node count (<= 2n - 1): the walk's look-ahead load stays inside
    root.alloc(1, mb);
}

template <class T>
void bh3_build(T* Y, int N, BhTree3<T>& tr, hipStream_t s, const AttractSlice<T, 3>* attract)
{
    require(N >= 2 && tr.N >= N, TKA_E_INTERNAL, "octree storage too small");
    const int nb = (int)ceil_div(N, kBox3Block);
    const unsigned g = ceil_div(N, 256);
    static_assert(kTile3 == 256 && kBox3Block == 256, "one tile and one box block per 256 points");
    hipLaunchKernelGGL(bbox3_part<T>, dim3(nb), dim3(kBox3Block), 0, s, (const T*)Y, N, tr.box.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(bh3_root, dim3(1), dim3(256), 0, s, (const double*)tr.box.get(), nb, N, tr.root.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(bh3_keys<T>, dim3(g), dim3(256), 0, s, Y, N, (const RootBox3*)tr.root.get(), tr.keys.get(),
                       tr.ktop.get(), tr.perm.get(), tr.mask.get(), tr.tile_cnt.get());
    TKA_LAUNCH_CHECK();
    // 3 radix passes over the top 24 key bits (octree levels 1-8), then the
    // runs of equal top bits by full key
    const bool alt = radix_sort_pairs<uint32_t, uint32_t>(tr.ktop.get(), tr.perm.get(), tr.ktop_alt.get(),
                                                          tr.perm_alt.get(), N, 24, tr.sort_scratch.get(), s);
    hipLaunchKernelGGL(bh3_runs<T>, dim3(g), dim3(256), 0, s, alt ? tr.ktop_alt.get() : tr.ktop.get(),
                       alt ? tr.perm_alt.get() : tr.perm.get(), (const uint64_t*)tr.keys.get(), (const T*)Y, N,
                       tr.keys_sorted.get(), tr.p2.get(), tr.ys.get());
    TKA_LAUNCH_CHECK();
    tr.k64 = tr.keys_sorted.get();
    tr.sorted_perm = tr.p2.get();
    // The attraction rows ride in extra workgroups of the node kernels (Y is
    // centred by the key kernel), split in proportion to the kernels'
    // durations; the run sort takes none (its LDS would hold them back).
    constexpr int kW[3] = {10, 5, 5};  // pairs, scan, emit
    AttractSlice<T, 3> slice[3];
    int wgs[3] = {0, 0, 0};
    if (attract) {
        const int r0 = attract->row0, rows = attract->row1 - attract->row0;
        int wacc = 0;
        for (int k = 0; k < 3; k++) {
            slice[k] = *attract;
            slice[k].row0 = r0 + (int)((int64_t)rows * wacc / 20);
            wacc += kW[k];
            slice[k].row1 = r0 + (int)((int64_t)rows * wacc / 20);
            wgs[k] = (int)ceil_div((size_t)std::max(0, slice[k].row1 - slice[k].row0), kRowsPerBlock);
        }
    }
    hipLaunchKernelGGL(bh3_pairs<T>, dim3(g + wgs[0]), dim3(256), 0, s, tr.k64, (const T*)tr.ys.get(), N,
                       tr.mask.get(), tr.node_a.get(), tr.node_b.get(), tr.node_lvl.get(), tr.tile_cnt.get(),
                       tr.part.get(), (int)g, slice[0]);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(bh3_scan<T>, dim3(g + wgs[1]), dim3(256), 0, s, (const unsigned long long*)tr.mask.get(),
                       (const T*)tr.ys.get(), N, (const uint32_t*)tr.tile_cnt.get(), (const double*)tr.part.get(),
                       tr.pre.get(), tr.ps.get(), (int)g, slice[1]);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(bh3_emit<T>, dim3(g + wgs[2]), dim3(256), 0, s, (const unsigned long long*)tr.mask.get(),
                       (const int*)tr.node_a.get(), (const int*)tr.node_b.get(), (const int*)tr.node_lvl.get(),
                       (const T*)tr.ys.get(), tr.sorted_perm, N, (const uint32_t*)tr.pre.get(),
                       (const double*)tr.ps.get(), tr.nodes.get(), (int)g, slice[2]);
    TKA_LAUNCH_CHECK();
}

template <class T>
int bh3_repulsion(const BhTree3<T>& tr, double theta, T* neg, double* zpart, unsigned long long* interactions,
                  hipStream_t s)
{
    const int g = bh_zblocks(tr.N);
    hipLaunchKernelGGL((bh3_traverse<T>), dim3(g), dim3(kTravBlock), 0, s, (const BhNode3<T>*)tr.nodes.get(),
                       (const uint32_t*)tr.pre.get(), (const T*)tr.ys.get(), tr.sorted_perm, tr.N,
                       (T)(theta * theta), (const RootBox3*)tr.root.get(), neg, zpart, interactions, tr.sort_fault());
    TKA_LAUNCH_CHECK();
    return g;
}

template <class T>
bool bh3_sort_failed(const BhTree3<T>& tr, hipStream_t s)
{
    return radix_sort_fault(tr.sort_scratch.get(), s) != 0u;
}

template <class T>
void bh3_inject_sort_fault(BhTree3<T>& tr, hipStream_t s)
{
    const uint32_t one = 1u;
    TKA_HIP(hipMemcpyAsync(tr.sort_scratch.get() + 1, &one, sizeof(one), hipMemcpyHostToDevice, s));
    TKA_HIP(hipStreamSynchronize(s));
}

template struct BhTree3<float>;
template struct BhTree3<double>;
template void bh3_build<float>(float*, int, BhTree3<float>&, hipStream_t, const AttractSlice<float, 3>*);
template void bh3_build<double>(double*, int, BhTree3<double>&, hipStream_t, const AttractSlice<double, 3>*);
template int bh3_repulsion<float>(const BhTree3<float>&, double, float*, double*, unsigned long long*, hipStream_t);
template int bh3_repulsion<double>(const BhTree3<double>&, double, double*, double*, unsigned long long*,
                                   hipStream_t);
template bool bh3_sort_failed<float>(const BhTree3<float>&, hipStream_t);
template bool bh3_sort_failed<double>(const BhTree3<double>&, hipStream_t);
template void bh3_inject_sort_fault<float>(BhTree3<float>&, hipStream_t);
template void bh3_inject_sort_fault<double>(BhTree3<double>&, hipStream_t);

}  // namespace tka
