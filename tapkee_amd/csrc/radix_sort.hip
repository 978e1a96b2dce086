// Radix sort and exclusive scan kernels (wave64, LDS-ranked, stable).
//
// Sort = one counting launch + one launch per 8-bit digit ("one sweep"):
//   rs_count     every tile histograms all digits of its keys at once and adds
//                them to global per-pass counts; it also clears the tile
//                status words.  A kernel that produces the keys can do this
//                itself (rs_count_tile).
//   rs_onesweep  turns the pass's global counts into digit bases (a block
//                scan, in every tile: no launch or arrival round trip between
//                counting and scattering), ranks its keys by the pass's digit
//                (wave ballots + per-wave LDS counters), publishes its
//                per-digit counts, sums its predecessors' counts by decoupled
//                look-back (Merrill & Garland 2016; Adinets & Merrill 2022)
//                and scatters.  The last tile to read the counts clears them.
//                Up to kRsMaxIndexedTiles tiles, tile = workgroup index: the
//                whole grid fits on the device, so a tile only waits on tiles
//                that run.  Beyond, tiles are numbered in the order their
//                workgroups start, so a tile only waits on tiles that are
//                already running.  The wait is bounded anyway; a tile that
//                gives up sets a fault word the host can read.
#include "radix_sort.hpp"

#include "attract.hpp"

namespace tka {
namespace {

constexpr uint32_t kFlagAgg = 1u << 30;   // tile's own count published
constexpr uint32_t kFlagInc = 2u << 30;   // inclusive prefix through this tile published
constexpr uint32_t kValMask = (1u << 30) - 1;
constexpr int kSpinLimit = 1 << 22;
constexpr int kLookback = 4;              // predecessor tiles read per look-back round

// Standalone counting launch: every block histograms a strided set of tiles
// in LDS and adds to the global counts once (one atomic per pass and digit
// per tile would pile thousands of tiles onto the same 256*passes counters).
constexpr int kRsCountBlocks = 256;

template <class K>
__global__ __launch_bounds__(kRsBlock) void rs_count(const K* __restrict__ keys, size_t n, int passes, RsLayout L,
                                                     int tiles)
{
    __shared__ uint32_t h[kRsMaxPasses][256];
    const int t = threadIdx.x;
    for (int p = 0; p < passes; p++) h[p][t] = 0;
    if (blockIdx.x == 0 && t < kRsMaxPasses) L.tile_ctr[t] = 0;
    __syncthreads();
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        for (int p = 0; p < passes; p++) L.status[((size_t)p * tiles + tile) * 256 + t] = 0;
        K key[kRsItems];
        bool ok[kRsItems];
        const size_t base = (size_t)tile * kRsBlock * kRsItems;
#pragma unroll
        for (int k = 0; k < kRsItems; k++) {
            const size_t i = base + (size_t)k * kRsBlock + t;
            ok[k] = i < n;
            key[k] = ok[k] ? keys[i] : K(0);
        }
        rs_tile_histogram<K, kRsItems>(key, ok, passes, h);
    }
    __syncthreads();
    rs_count_flush(passes, L, h);
}

// Pass 0 of the counted sort in direct mode: the digit totals and the counts
// in the tiles before `below`, from the per-tile counts (tiles x 128 words,
// two 16-bit counts each).  A wave reads two rows per load (a half-wave per
// row, eight digits per lane), wave w rows 2w, 2w + 1, 2w + 8, ... with all
// its loads in flight at once; the waves' sums meet in LDS.  Returns the
// total of digit threadIdx.x and its count before `below` in *before.  Called
// by all kRsBlock threads; tiles <= kRsDirectTiles.
constexpr int kSumWaves = kRsBlock / 64;
constexpr int kSumLoads = kRsDirectTiles / (2 * kSumWaves);
TKA_DEV uint32_t rs_direct_sums(const uint32_t* counts, int tiles, int below, uint32_t* before,
                                uint32_t (*sh)[256])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, q = lane & 31;
    const uint4* c4 = reinterpret_cast<const uint4*>(counts);
    uint32_t tot[8], bef[8];
#pragma unroll
    for (int c = 0; c < 8; c++) tot[c] = bef[c] = 0u;
    constexpr int kBatch = 8;
    for (int k0 = 0; k0 < kSumLoads; k0 += kBatch) {
        if (2 * (k0 * kSumWaves + w) >= tiles) break;
        uint4 v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int r = 2 * ((k0 + k) * kSumWaves + w) + half;
            v[k] = r < tiles ? c4[(size_t)r * 32 + q] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int r = 2 * ((k0 + k) * kSumWaves + w) + half;
            const uint32_t u[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            const bool b = r < below;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t lo = u[c] & 0xffffu, hi = u[c] >> 16;
                tot[2 * c] += lo;
                tot[2 * c + 1] += hi;
                bef[2 * c] += b ? lo : 0u;
                bef[2 * c + 1] += b ? hi : 0u;
            }
        }
    }
    // digit 8q + c of this half-wave
#pragma unroll
    for (int c = 0; c < 8; c++) {
        tot[c] += __shfl_xor(tot[c], 32, 64);
        bef[c] += __shfl_xor(bef[c], 32, 64);
    }
    if (half == 0) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
            sh[2 * w][8 * q + c] = tot[c];
            sh[2 * w + 1][8 * q + c] = bef[c];
        }
    }
    __syncthreads();
    uint32_t t = 0u, b = 0u;
#pragma unroll
    for (int ww = 0; ww < kSumWaves; ww++) {
        t += sh[2 * ww][threadIdx.x];
        b += sh[2 * ww + 1][threadIdx.x];
    }
    *before = b;
    return t;
}

// Stable scatter of one tile.  Item k of lane l in wave w sits at tile
// position w*64*ITEMS + k*64 + l, so ranking items in (w, k, l) order keeps
// the input order among equal digits.
// Workgroups past the pass's tiles run the side work and take no tile.
template <class K, class V, int ITEMS, class Side>
__global__ __launch_bounds__(kRsBlock) void rs_onesweep(const K* __restrict__ kin, const V* __restrict__ vin,
                                                        K* __restrict__ kout, V* __restrict__ vout, size_t n,
                                                        int pass, RsLayout L, int tiles, int ticketed, int direct_mode,
                                                        Side side)
{
    if ((int)blockIdx.x >= tiles) {
        side.run((int)blockIdx.x - tiles, (int)gridDim.x - tiles);
        return;
    }
    constexpr int kWaves = kRsBlock / 64;
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t base[256];
    __shared__ uint32_t scan_sh[kWaves];
    __shared__ uint32_t sum_sh[2 * kSumWaves][256];
    __shared__ uint32_t tile_sh;
    __shared__ int last_sh;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // direct mode (counted sort, few tiles): pass 0's totals and tile
    // offsets from the per-tile counts, without look-back
    const bool direct = direct_mode != 0 && pass == 0;
    // the global count of digit threadIdx.x, in flight with the keys
    const uint32_t gc = direct ? 0u : L.gcnt[pass * 256 + threadIdx.x];
    for (int t = threadIdx.x; t < kWaves * 256; t += kRsBlock) (&wcnt[0][0])[t] = 0;
    if (ticketed) {
        if (threadIdx.x == 0) tile_sh = atomicAdd(&L.tile_ctr[pass], 1u);
        __syncthreads();
    }
    const uint32_t tile = ticketed ? tile_sh : blockIdx.x;
    const int shift = 8 * pass;

    K key[ITEMS];
    V val[ITEMS];
    uint32_t rank[ITEMS];
    int dig[ITEMS];
    const size_t tile0 = (size_t)tile * kRsBlock * ITEMS + (size_t)w * 64 * ITEMS + lane;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t i = tile0 + (size_t)k * 64;
        const bool ok = i < n;
        key[k] = ok ? kin[i] : K(0);
        val[k] = ok ? vin[i] : V(0);
    }
    uint32_t total = gc, before0 = 0u;
    if (direct)
        total = rs_direct_sums(rs_tile_counts(L.status, tiles), tiles, (int)tile, &before0, sum_sh);
    // digit bases; the barrier inside also completes the clearing of wcnt
    const uint32_t gbase = block_exclusive_scan(total, scan_sh, nullptr);
    // every thread has its count now: register this tile as a reader (the
    // result is needed only at the end)
    uint32_t readers = 0;
    if (!direct && threadIdx.x == 0) readers = atomicAdd(&L.rd_ctr[pass], 1u);
#pragma unroll
    for (int k = 0; k < ITEMS; k++) dig[k] = (tile0 + (size_t)k * 64 < n) ? (int)((uint32_t)(key[k] >> shift) & 255u) : -1;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int d = dig[k];
        const bool ok = d >= 0;
        uint64_t m = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t prior = ok ? wcnt[w][d] : 0u;
        const uint32_t below = (uint32_t)__popcll(m & lt);
        rank[k] = prior + below;
        __builtin_amdgcn_wave_barrier();
        if (ok && below == 0) wcnt[w][d] = prior + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t cnt = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = cnt;
            cnt += c;
        }
        uint32_t* st = L.status + (size_t)pass * tiles * 256 + d;
        uint32_t excl = 0;
        if (direct) {
            excl = before0;
        } else if (tile == 0) {
            rs_st(st, kFlagInc | cnt);
        } else {
            rs_st(st + (size_t)tile * 256, kFlagAgg | cnt);
            // Look back over a window of kLookback predecessors per round
            // (independent loads in flight together), nearest first: add
            // published counts until an inclusive prefix; stop the round at
            // a tile that has not published yet and read it again.
            int j = (int)tile - 1, spins = 0;
            bool done = false;
            while (!done) {
                uint32_t v[kLookback];
#pragma unroll
                for (int q = 0; q < kLookback; q++) v[q] = (j - q >= 0) ? rs_ld(st + (size_t)(j - q) * 256) : 0u;
                // branch-free over the window (keeps v[] in registers); tile 0
                // is always inclusive, so entries past it are never taken
                int used = 0;
                bool stop = false;
#pragma unroll
                for (int q = 0; q < kLookback; q++) {
                    const uint32_t f = v[q] & ~kValMask;
                    const bool take = !stop && f != 0;
                    excl += take ? (v[q] & kValMask) : 0u;
                    used = take ? q + 1 : used;
                    done = done || (take && f == kFlagInc);
                    stop = stop || f == 0 || f == kFlagInc;
                }
                if (done) break;
                j -= used;
                if (used == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(L.fault, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            rs_st(st + (size_t)tile * 256, kFlagInc | (excl + cnt));
        }
        base[d] = gbase + excl;
        if (threadIdx.x == 0) last_sh = !direct && readers == (uint32_t)(tiles - 1);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int d = dig[k];
        if (d >= 0) {
            const uint32_t pos = base[d] + wcnt[w][d] + rank[k];
            kout[pos] = key[k];
            vout[pos] = val[k];
        }
    }
    // every tile has read the counts: clear them for the next sort
    if (last_sh) {
        L.gcnt[pass * 256 + threadIdx.x] = 0;
        if (threadIdx.x == 0) L.rd_ctr[pass] = 0;
    }
}

// ---------------------------------------------------------------- scan
constexpr int kScanItems = kScanTile / kRsBlock;

__global__ __launch_bounds__(kRsBlock) void scan_tile_sums(const uint32_t* __restrict__ in, size_t n,
                                                           uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sh[kRsBlock / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of L words by one 1024-thread block; `in` and `out` may
// alias (each thread reads its chunk before writing it).
__global__ __launch_bounds__(1024) void scan_single(const uint32_t* in, uint32_t* out, size_t L)
{
    __shared__ uint32_t sh[16];
    const size_t per = (L + 1023) / 1024;
    const size_t beg = threadIdx.x * per;
    const size_t end = beg + per < L ? beg + per : L;
    uint32_t s = 0;
    for (size_t i = beg; i < end; i++) s += in[i];
    uint32_t run = block_exclusive_scan(s, sh, nullptr);
    for (size_t i = beg; i < end; i++) {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kRsBlock) void scan_tiles(const uint32_t* __restrict__ in, size_t n,
                                                       const uint32_t* __restrict__ tile_off,
                                                       uint32_t* __restrict__ out, int write_total)
{
    __shared__ uint32_t sh[kRsBlock / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t run = block_exclusive_scan(s, sh, nullptr) + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    const unsigned last = (unsigned)((n + kScanTile - 1) / kScanTile) - 1;
    if (write_total && blockIdx.x == last && threadIdx.x == kRsBlock - 1) out[n] = run;
}

void scan_impl(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, bool write_total, hipStream_t s)
{
    const int tiles = (int)ceil_div(n, kScanTile);
    hipLaunchKernelGGL(scan_tile_sums, dim3(tiles), dim3(kRsBlock), 0, s, in, n, scratch);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_single, dim3(1), dim3(1024), 0, s, scratch, scratch, (size_t)tiles);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_tiles, dim3(tiles), dim3(kRsBlock), 0, s, in, n, scratch, out, write_total ? 1 : 0);
    TKA_LAUNCH_CHECK();
}

template <class K, class V, int ITEMS, class Side = RsNoSide>
bool sort_impl(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch, bool counted,
               hipStream_t s, const Side* side = nullptr, const int* side_wgs = nullptr)
{
    const int tiles = (int)ceil_div(n, (size_t)kRsBlock * ITEMS);
    const int passes = (end_bit + 7) / 8;
    const RsLayout L = rs_layout(scratch);
    require(counted || ITEMS == kRsItems, TKA_E_INTERNAL, "radix sort: the counting launch uses kRsItems tiles");
    require(!counted || passes <= kRsCountedPasses, TKA_E_INTERNAL, "radix sort: the counted sort takes 32-bit keys");
    // the key kernel's rs_count_tile chose the same mode from the same tile count
    const int direct = counted && tiles <= kRsDirectTiles ? 1 : 0;
    if (!counted) {
        hipLaunchKernelGGL((rs_count<K>), dim3(std::min(tiles, kRsCountBlocks)), dim3(kRsBlock), 0, s, keys, n, passes, L,
                           tiles);
        TKA_LAUNCH_CHECK();
    }
    K* ka = keys;
    V* va = vals;
    K* kb = keys_alt;
    V* vb = vals_alt;
    for (int p = 0; p < passes; p++) {
        const int extra = side ? side_wgs[p] : 0;
        hipLaunchKernelGGL((rs_onesweep<K, V, ITEMS, Side>), dim3(tiles + extra), dim3(kRsBlock), 0, s, ka, va, kb,
                           vb, n, p, L, tiles, tiles > kRsMaxIndexedTiles ? 1 : 0, direct, side ? side[p] : Side{});
        TKA_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    return (passes & 1) != 0;
}

}  // namespace

size_t radix_sort_scratch_words(size_t n, int items)
{
    return kRsHeader + (size_t)kRsMaxPasses * 256 * radix_tiles(n ? n : 1, items);
}

template <class K, class V>
bool radix_sort_pairs(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                      hipStream_t stream)
{
    if (n <= 1 || end_bit <= 0) return false;
    require(n <= kRsMaxItems, TKA_E_UNSUPPORTED, "radix sort: more than 2^30 - 1 items");
    require(end_bit <= 8 * (int)sizeof(K), TKA_E_INTERNAL, "radix sort: end_bit beyond the key width");
    return sort_impl<K, V, kRsItems>(keys, vals, keys_alt, vals_alt, n, end_bit, scratch, false, stream);
}

template <class K, class V>
bool radix_sort_pairs_counted(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                              hipStream_t stream)
{
    if (n <= 1 || end_bit <= 0) return false;
    require(n <= kRsMaxItems, TKA_E_UNSUPPORTED, "radix sort: more than 2^30 - 1 items");
    require(end_bit <= 8 * (int)sizeof(K), TKA_E_INTERNAL, "radix sort: end_bit beyond the key width");
    return sort_impl<K, V, kRsItemsCounted>(keys, vals, keys_alt, vals_alt, n, end_bit, scratch, true, stream);
}

template bool radix_sort_pairs_counted<uint32_t, uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, size_t, int,
                                                           uint32_t*, hipStream_t);

template <class K, class V, class Side>
bool radix_sort_pairs_counted_side(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit,
                                   uint32_t* scratch, const Side* side, const int* side_wgs, hipStream_t stream)
{
    if (n <= 1 || end_bit <= 0) return false;
    require(n <= kRsMaxItems, TKA_E_UNSUPPORTED, "radix sort: more than 2^30 - 1 items");
    require(end_bit <= 8 * (int)sizeof(K), TKA_E_INTERNAL, "radix sort: end_bit beyond the key width");
    return sort_impl<K, V, kRsItemsCounted, Side>(keys, vals, keys_alt, vals_alt, n, end_bit, scratch, true, stream,
                                                  side, side_wgs);
}

template bool radix_sort_pairs_counted_side<uint32_t, uint32_t, AttractSlice<float>>(
    uint32_t*, uint32_t*, uint32_t*, uint32_t*, size_t, int, uint32_t*, const AttractSlice<float>*, const int*,
    hipStream_t);
template bool radix_sort_pairs_counted_side<uint32_t, uint32_t, AttractSlice<double>>(
    uint32_t*, uint32_t*, uint32_t*, uint32_t*, size_t, int, uint32_t*, const AttractSlice<double>*, const int*,
    hipStream_t);

template bool radix_sort_pairs<uint32_t, uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, size_t, int, uint32_t*,
                                                   hipStream_t);
template bool radix_sort_pairs<uint64_t, uint64_t>(uint64_t*, uint64_t*, uint64_t*, uint64_t*, size_t, int, uint32_t*,
                                                   hipStream_t);
template bool radix_sort_pairs<uint64_t, uint32_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, size_t, int, uint32_t*,
                                                   hipStream_t);

uint32_t radix_sort_fault(const uint32_t* scratch, hipStream_t stream)
{
    uint32_t f = 0;
    read_to_host(&f, scratch + 1, sizeof(f), stream);
    return f;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t stream)
{
    if (n == 0) {
        TKA_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), stream));
        return;
    }
    scan_impl(in, out, n, scratch, true, stream);
}

}  // namespace tka
