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

}  // namespace tka
