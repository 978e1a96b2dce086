// Searches and scans shared by the Barnes-Hut tree builds (quadtree and
// octree): lower bounds over sorted Morton keys, in an LDS window or by a
// whole wave over global memory, and a wave-wide inclusive scan.
#pragma once

#include "common.hpp"

namespace tka {

// Smallest g in [lo, hi) with win[g - w0] >= t (strict: > t), or hi; the
// range lies inside the window.
TKA_DEV int lds_lower_bound(const uint64_t* win, int w0, int lo, int hi, uint64_t t, bool strict)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t v = win[mid - w0];
        if (strict ? v > t : v >= t)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// The same over global keys k[lo, hi), by all 64 lanes of the wave for one
// (wave-uniform) query.
TKA_DEV int wave_lower_bound(const uint64_t* __restrict__ k, int lo, int hi, uint64_t t, bool strict)
{
    const int lane = threadIdx.x & 63;
    while (lo < hi) {
        const int step = (hi - lo + 63) >> 6;
        const int q = lo + lane * step;
        const bool in = q < hi;
        const uint64_t v = in ? k[q] : 0ull;
        const uint64_t bal = __ballot(in && (strict ? v > t : v >= t));
        if (bal == 0ull) {
            lo += ((hi - 1 - lo) / step) * step + 1;  // past the last probe
        } else {
            const int f = __ffsll((long long)bal) - 1;
            if (f == 0) return lo;
            hi = lo + f * step;
            lo += (f - 1) * step + 1;
        }
    }
    return hi;
}

// Runs the searches flagged in `need` (one bit per lane; query t of that lane
// over [lo, hi), lo and hi per lane) with wave_lower_bound; each flagged lane
// receives its result in *res.
TKA_DEV void wave_searches(uint64_t need, const uint64_t* __restrict__ k, int lo, int hi, uint64_t t, bool strict,
                           int* res)
{
    const int lane = threadIdx.x & 63;
    while (need != 0ull) {
        const int l = __ffsll((long long)need) - 1;
        need &= need - 1ull;
        const int ql = __shfl(lo, l, 64), qh = __shfl(hi, l, 64);
        const uint64_t qt = ((uint64_t)(uint32_t)__shfl((int)(t >> 32), l, 64) << 32) |
                            (uint32_t)__shfl((int)(uint32_t)t, l, 64);
        const int r = wave_lower_bound(k, ql, qh, qt, strict);
        if (lane == l) *res = r;
    }
}

TKA_DEV double wave_incl_scan_d(double v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Full 64-bit keys in sorted order and the final point order (the
// Barnes-Hut trees' run sort).  The radix sort ordered the points by the top bits of their
// keys, keeping equal top bits in point order; each run of equal top bits is
// then put in (full key, point) order, which equals a stable sort by full key:
//  - a run of at most kFixHalo points: each of its points goes to the run's
//    start plus its rank among the run's (key, point) pairs, counted in LDS.
//    A workgroup stages its 256 positions and kFixHalo more on either side,
//    which hold the whole run of each of its points, and places its own
//    points (a run that crosses workgroups is ranked alike by each);
//  - a longer run by the whole wave holding its first point, as a bitonic
//    sort in the wave's LDS (at most kWaveRun points) or in place in the
//    outputs in global memory.
// The bitonic network compares every pair (a, b), a < b, into ascending order
// (the first step of each merge compares mirror positions), so a run of any
// length sorts as if padded with +infinity: comparisons with b past the run
// are skipped.
constexpr int kFixHalo = 64;
constexpr int kFixWin = 256 + 2 * kFixHalo;
constexpr int kWaveRun = 1024;

TKA_DEV void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

TKA_DEV void wave_sync_global()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <bool kGlobal>
TKA_DEV void wave_bitonic(uint64_t* K, uint32_t* P, int len)
{
    const int lane = threadIdx.x & 63;
    int n = 1;
    while (n < len) n <<= 1;
    auto cmpx = [&](int a, int b) {
        if (b >= len) return;
        const uint64_t ka = K[a], kb = K[b];
        const uint32_t pa = P[a], pb = P[b];
        if (kb < ka || (kb == ka && pb < pa)) {
            K[a] = kb;
            K[b] = ka;
            P[a] = pb;
            P[b] = pa;
        }
    };
    for (int size = 2; size <= n; size <<= 1) {
        const int half = size >> 1;
        for (int t = lane; t < (n >> 1); t += 64) {
            const int blk = (t / half) * size, x = t % half;
            cmpx(blk + x, blk + size - 1 - x);
        }
        if (kGlobal) wave_sync_global(); else wave_sync_lds();
        for (int d = size >> 2; d >= 1; d >>= 1) {
            for (int t = lane; t < (n >> 1); t += 64) {
                const int a = (t / d) * 2 * d + (t % d);
                cmpx(a, a + d);
            }
            if (kGlobal) wave_sync_global(); else wave_sync_lds();
        }
    }
}

// The run sort of one 256-position tile, called by all 256 threads of the
// workgroup that holds tile blockIdx.x: ktop and perm are the radix-sorted
// top key bits and points, kfull the full key of each point; writes the
// sorted full keys k64 and points p2, and calls emit(position, point) once
// for every position it writes.
template <class Emit>
TKA_DEV void run_sort_tile(const uint32_t* __restrict__ ktop, const uint32_t* __restrict__ perm,
                           const uint64_t* __restrict__ kfull, int N, uint64_t* __restrict__ k64,
                           uint32_t* __restrict__ p2, Emit emit)
{
    __shared__ uint64_t sk[4][kWaveRun];
    __shared__ uint32_t sp[4][kWaveRun];
    __shared__ uint64_t wk[kFixWin];  // window: full keys, top bits, points
    __shared__ uint32_t wt[kFixWin], wp[kFixWin];
    __shared__ uint64_t wstart[kFixWin / 64];  // bit k % 64 of word k / 64: a run starts at window position k
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int w0 = (int)blockIdx.x * 256 - kFixHalo;  // window position 0
    const int lo = max(0, -w0), hi = min(kFixWin, N - w0);  // valid window positions
    // Sorted point indices are clamped to N-1 when read: a sort whose tile
    // look-back gave up (its fault word set) may leave positions unwritten,
    // and every gather and scatter downstream then still stays in bounds.
    const uint32_t pmax = (uint32_t)(N - 1);
    static_assert(kFixWin <= 512 && kFixWin % 64 == 0, "two window positions per thread, whole words of flags");
    {
        // two positions per thread, both points requested before their keys
        const int t1 = (int)threadIdx.x + 256;
        const int g0 = w0 + (int)threadIdx.x, g1 = w0 + t1;
        const bool in0 = g0 >= 0 && g0 < N, in1 = t1 < kFixWin && g1 >= 0 && g1 < N;
        const uint32_t p0 = in0 ? min(perm[g0], pmax) : 0u, p1 = in1 ? min(perm[g1], pmax) : 0u;
        const uint32_t h0 = in0 ? ktop[g0] : 0u, h1 = in1 ? ktop[g1] : 0u;
        const uint64_t f0 = in0 ? kfull[p0] : 0ull, f1 = in1 ? kfull[p1] : 0ull;
        wt[threadIdx.x] = h0;
        wp[threadIdx.x] = p0;
        wk[threadIdx.x] = f0;
        if (t1 < kFixWin) {
            wt[t1] = h1;
            wp[t1] = p1;
            wk[t1] = f1;
        }
    }
    __syncthreads();
    // run starts (every position outside [lo, hi) counts as one, so no run
    // reaches past the valid keys); the second positions are whole waves
    for (int k = (int)threadIdx.x; k < kFixWin; k += 256) {
        const bool f = k <= lo || k >= hi || wt[k] != wt[k - 1];
        const uint64_t m = __ballot(f);
        if (lane == 0) wstart[k >> 6] = m;
    }
    __syncthreads();
    const int j = (int)blockIdx.x * 256 + (int)threadIdx.x;
    const int lj = (int)threadIdx.x + kFixHalo;
    bool start = false;  // j starts a run longer than kFixHalo
    if (j < N) {
        // The run around j from the start bits: its start is the last one at
        // or before lj, its end the first one after lj, each looked for in
        // lj's word and the next one out (a run of at most kFixHalo = 64
        // points has both there; one not found means a longer run).
        const int wi = lj >> 6, bi = lj & 63;
        const uint64_t upto = (2ull << bi) - 1ull;  // bits 0..bi (all for bi = 63)
        const uint64_t mb = wstart[wi] & upto, ma = wstart[wi] & ~upto;
        int s = -1, e = INT_MAX;
        if (mb != 0ull) {
            s = wi * 64 + 63 - __clzll((long long)mb);
        } else {
            const uint64_t m2 = wstart[wi - 1];  // wi >= 1: lj >= kFixHalo
            if (m2 != 0ull) s = (wi - 1) * 64 + 63 - __clzll((long long)m2);
        }
        if (ma != 0ull) {
            e = wi * 64 + __ffsll((long long)ma) - 1;
        } else {
            const uint64_t m2 = wstart[wi + 1];  // wi + 1 < kFixWin / 64: lj < 256 + kFixHalo
            if (m2 != 0ull) e = (wi + 1) * 64 + __ffsll((long long)m2) - 1;
        }
        if (s >= 0 && e != INT_MAX && e - s <= kFixHalo) {
            const uint64_t kj = wk[lj];
            const uint32_t pj = wp[lj];
            int r = 0;
            for (int k = s; k < e; k++) {
                const uint64_t kk = wk[k];
                r += (kk < kj || (kk == kj && wp[k] < pj)) ? 1 : 0;
            }
            const int d = w0 + s + r;
            k64[d] = kj;
            p2[d] = pj;
            emit(d, pj);
        } else {
            start = ((wstart[wi] >> bi) & 1ull) != 0ull;
        }
    }
    // long runs, one at a time, by the whole wave; the end by a gallop, then
    // a bisection, over the sorted top bits
    int e = j + 1;
    if (start) {
        const uint32_t h = wt[lj];
        int lo2 = e, hi2 = N, step = 2;  // ktop[lo2] == h (the run is longer than 1); hi2: first index known past it
        while (j + step < N) {
            if (ktop[j + step] != h) {
                hi2 = j + step;
                break;
            }
            lo2 = j + step;
            step <<= 1;
        }
        while (hi2 - lo2 > 1) {
            const int mid = (lo2 + hi2) >> 1;
            if (ktop[mid] == h)
                lo2 = mid;
            else
                hi2 = mid;
        }
        e = hi2;
    }
    uint64_t longs = __ballot(start);
    while (longs != 0ull) {
        const int l = __ffsll((long long)longs) - 1;
        longs &= longs - 1ull;
        const int rj = __shfl(j, l, 64), len = __shfl(e, l, 64) - rj;
        if (len <= kWaveRun) {
            uint64_t* K = sk[w];
            uint32_t* P = sp[w];
            for (int i = lane; i < len; i += 64) {
                const uint32_t p = min(perm[rj + i], pmax);
                P[i] = p;
                K[i] = kfull[p];
            }
            wave_sync_lds();
            wave_bitonic<false>(K, P, len);
            for (int i = lane; i < len; i += 64) {
                const uint32_t p = P[i];
                k64[rj + i] = K[i];
                p2[rj + i] = p;
                emit(rj + i, p);
            }
            wave_sync_lds();  // the next run reuses the LDS
        } else {
            uint64_t* K = k64 + rj;
            uint32_t* P = p2 + rj;
            for (int i = lane; i < len; i += 64) {
                const uint32_t p = min(perm[rj + i], pmax);
                P[i] = p;
                K[i] = kfull[p];
            }
            wave_sync_global();
            wave_bitonic<true>(K, P, len);  // ends with a fence: P is complete
            for (int i = lane; i < len; i += 64) emit(rj + i, P[i]);
        }
    }
}

}  // namespace tka
