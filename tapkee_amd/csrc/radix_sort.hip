// Radix sort and exclusive scan kernels (wave64, LDS-ranked, stable).
#include "radix_sort.hpp"

namespace tka {
namespace {

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

template <class K>
__global__ __launch_bounds__(kRsBlock) void rs_hist(const K* __restrict__ keys, size_t n, int shift,
                                                    uint32_t* __restrict__ hist, int tiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
#pragma unroll
    for (int k = 0; k < kRsItems; k++) {
        const size_t i = base + (size_t)k * kRsBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of L words by one 1024-thread block (digit-major histogram).
// `in` and `out` may alias (each thread reads its chunk before writing it).
__global__ __launch_bounds__(1024) void rs_scan_single(const uint32_t* in, uint32_t* out, size_t L)
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

// Stable scatter of one tile.  Item k of lane l in wave w sits at tile
// position w*64*kRsItems + k*64 + l, so ranking items in (w, k, l) order keeps
// the input order among equal digits.
template <class K, class V, bool kFused>
__global__ __launch_bounds__(kRsBlock) void rs_scatter(const K* __restrict__ kin, const V* __restrict__ vin,
                                                       K* __restrict__ kout, V* __restrict__ vout, size_t n, int shift,
                                                       const uint32_t* __restrict__ hist,
                                                       const uint32_t* __restrict__ offs, int tiles)
{
    constexpr int kWaves = kRsBlock / 64;
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t base[256];
    __shared__ uint32_t sh[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = threadIdx.x; t < kWaves * 256; t += kRsBlock) (&wcnt[0][0])[t] = 0;
    {
        const int d = threadIdx.x;
        if (kFused) {
            uint32_t tot = 0, before = 0;
            for (int b = 0; b < tiles; b++) {
                const uint32_t v = hist[(size_t)d * tiles + b];
                tot += v;
                before += (b < (int)blockIdx.x) ? v : 0u;
            }
            const uint32_t ex = block_exclusive_scan(tot, sh, nullptr);
            base[d] = ex + before;
        } else {
            base[d] = offs[(size_t)d * tiles + blockIdx.x];
        }
    }
    __syncthreads();

    K key[kRsItems];
    V val[kRsItems];
    uint32_t rank[kRsItems];
    int dig[kRsItems];
    const size_t tile0 = (size_t)blockIdx.x * kRsTile + (size_t)w * 64 * kRsItems + lane;
#pragma unroll
    for (int k = 0; k < kRsItems; k++) {
        const size_t i = tile0 + (size_t)k * 64;
        const bool ok = i < n;
        key[k] = ok ? kin[i] : K(0);
        val[k] = ok ? vin[i] : V(0);
        dig[k] = ok ? (int)((uint32_t)(key[k] >> shift) & 255u) : -1;
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < kRsItems; k++) {
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
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRsItems; k++) {
        const int d = dig[k];
        if (d >= 0) {
            const uint32_t pos = base[d] + wcnt[w][d] + rank[k];
            kout[pos] = key[k];
            vout[pos] = val[k];
        }
    }
}

// ---------------------------------------------------------------- scan
__global__ __launch_bounds__(kRsBlock) void scan_tile_sums(const uint32_t* __restrict__ in, size_t n,
                                                           uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sh[kRsBlock / 64];
    const size_t base = (size_t)blockIdx.x * kRsTile;
    uint32_t s = 0;
    for (int k = 0; k < kRsItems; k++) {
        const size_t i = base + (size_t)threadIdx.x * kRsItems + k;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kRsBlock) void scan_tiles(const uint32_t* __restrict__ in, size_t n,
                                                       const uint32_t* __restrict__ tile_off,
                                                       uint32_t* __restrict__ out)
{
    __shared__ uint32_t sh[kRsBlock / 64];
    const size_t base = (size_t)blockIdx.x * kRsTile + (size_t)threadIdx.x * kRsItems;
    uint32_t v[kRsItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kRsItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t run = block_exclusive_scan(s, sh, nullptr) + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kRsItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    const unsigned last = (unsigned)((n + kRsTile - 1) / kRsTile) - 1;
    if (blockIdx.x == last && threadIdx.x == kRsBlock - 1) out[n] = run;
}

}  // namespace

template <class K, class V>
bool radix_sort_pairs(K* keys, V* vals, K* keys_alt, V* vals_alt, size_t n, int end_bit, uint32_t* scratch,
                      hipStream_t stream)
{
    if (n <= 1) return false;
    const int tiles = (int)ceil_div(n, kRsTile);
    uint32_t* hist = scratch;
    uint32_t* offs = scratch + (size_t)256 * tiles;
    K* ka = keys;
    V* va = vals;
    K* kb = keys_alt;
    V* vb = vals_alt;
    bool in_alt = false;
    for (int shift = 0; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(rs_hist<K>, dim3(tiles), dim3(kRsBlock), 0, stream, ka, n, shift, hist, tiles);
        TKA_LAUNCH_CHECK();
        if (tiles <= kRsFusedMaxTiles) {
            hipLaunchKernelGGL((rs_scatter<K, V, true>), dim3(tiles), dim3(kRsBlock), 0, stream, ka, va, kb, vb, n,
                               shift, hist, nullptr, tiles);
        } else {
            hipLaunchKernelGGL(rs_scan_single, dim3(1), dim3(1024), 0, stream, hist, offs, (size_t)256 * tiles);
            TKA_LAUNCH_CHECK();
            hipLaunchKernelGGL((rs_scatter<K, V, false>), dim3(tiles), dim3(kRsBlock), 0, stream, ka, va, kb, vb, n,
                               shift, hist, offs, tiles);
        }
        TKA_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
        in_alt = !in_alt;
    }
    return in_alt;
}

template bool radix_sort_pairs<uint32_t, uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, size_t, int, uint32_t*,
                                                   hipStream_t);
template bool radix_sort_pairs<uint64_t, uint64_t>(uint64_t*, uint64_t*, uint64_t*, uint64_t*, size_t, int, uint32_t*,
                                                   hipStream_t);
template bool radix_sort_pairs<uint64_t, uint32_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, size_t, int, uint32_t*,
                                                   hipStream_t);

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t stream)
{
    if (n == 0) {
        TKA_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), stream));
        return;
    }
    const int tiles = (int)ceil_div(n, kRsTile);
    uint32_t* sums = scratch;  // tiles words; scanned in place into sums (exclusive)
    hipLaunchKernelGGL(scan_tile_sums, dim3(tiles), dim3(kRsBlock), 0, stream, in, n, sums);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(rs_scan_single, dim3(1), dim3(1024), 0, stream, sums, sums, (size_t)tiles);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_tiles, dim3(tiles), dim3(kRsBlock), 0, stream, in, n, sums, out);
    TKA_LAUNCH_CHECK();
}

}  // namespace tka
