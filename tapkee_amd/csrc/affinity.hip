// Perplexity calibration (FP64) and symmetrisation of P.
#include "affinity.hpp"

#include <cfloat>

#include "radix_sort.hpp"

namespace tka {
namespace {

constexpr int kBisectMaxIter = 200;  // tsne.hpp:658
constexpr double kBisectTol = 1e-5;  // tsne.hpp:653

// Bisection state update, tsne.hpp:676-698 (identical in :563-585).
struct Bisect {
    double beta = 1.0, min_beta = -DBL_MAX, max_beta = DBL_MAX;
    TKA_DEV bool step(double H, double logp)
    {
        const double Hdiff = H - logp;
        if (Hdiff < kBisectTol && -Hdiff < kBisectTol) return true;
        if (Hdiff > 0) {
            min_beta = beta;
            if (max_beta == DBL_MAX || max_beta == -DBL_MAX)
                beta *= 2.0;
            else
                beta = (beta + max_beta) / 2.0;
        } else {
            max_beta = beta;
            if (min_beta == -DBL_MAX || min_beta == DBL_MAX)
                beta /= 2.0;
            else
                beta = (beta + min_beta) / 2.0;
        }
        return false;
    }
};

// One wave per kNN row; K <= 512 values live in registers (8 per lane).
__global__ __launch_bounds__(256) void perplexity_knn(const double* __restrict__ dist, int N, int K, double logp,
                                                      double* __restrict__ val, double* __restrict__ beta_out,
                                                      int* __restrict__ iters_out, int* __restrict__ unconv)
{
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= N) return;
    double d[8], p[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int m = r * 64 + lane;
        d[r] = m < K ? dist[(size_t)row * K + m] : 0.0;
        p[r] = 0.0;
    }
    Bisect b;
    bool found = false;
    int iter = 0;
    double sum_P = DBL_MIN, last_beta = b.beta;
    while (!found && iter < kBisectMaxIter) {
        const double beta = b.beta;
        last_beta = beta;
        double s = 0.0, h = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int m = r * 64 + lane;
            if (m < K) {
                p[r] = exp(-beta * d[r]);
                s += p[r];
                h += beta * (d[r] * p[r]);
            }
        }
        s = wave_sum(s);
        h = wave_sum(h);
        sum_P = DBL_MIN + s;
        const double H = (h / sum_P) + log(sum_P);
        found = b.step(H, logp);
        iter++;
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int m = r * 64 + lane;
        if (m < K) val[(size_t)row * K + m] = p[r] / sum_P;
    }
    if (lane == 0) {
        if (beta_out) beta_out[row] = last_beta;
        if (iters_out) iters_out[row] = iter;
        if (!found) atomicAdd(unconv, 1);
    }
}

// Rows of K <= 16 R values, four per wave: 16 lanes per row, R values per
// lane.  Each evaluation's sums are 16-lane butterflies, so every lane of a
// row holds the same bits and steps the bisection identically; a row whose
// bisection has ended keeps its values while the wave's other rows go on.
// Four rows share each exp instruction sequence and each reduction step,
// where one row per wave left a 90-value row's second sequence on 26 lanes
// and paid a 64-lane reduction per row.
template <int R>
__global__ __launch_bounds__(256) void perplexity_knn_packed(const double* __restrict__ dist, int N, int K,
                                                             double logp, double* __restrict__ val,
                                                             double* __restrict__ beta_out,
                                                             int* __restrict__ iters_out, int* __restrict__ unconv)
{
    const int lane = threadIdx.x & 63, sub = lane & 15;
    const int row = blockIdx.x * 16 + (threadIdx.x >> 4);
    const bool live = row < N;
    double d[R], p[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int m = r * 16 + sub;
        d[r] = (live && m < K) ? dist[(size_t)row * K + m] : 0.0;
        p[r] = 0.0;
    }
    Bisect b;
    bool found = !live;
    int iter = 0;
    double sum_P = DBL_MIN, last_beta = b.beta;
    while (__ballot(!found && iter < kBisectMaxIter) != 0ull) {
        const bool go = !found && iter < kBisectMaxIter;
        double s = 0.0, h = 0.0;
        const double beta = b.beta;
        if (go) {
            last_beta = beta;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int m = r * 16 + sub;
                if (m < K) {
                    p[r] = exp(-beta * d[r]);
                    s += p[r];
                    h += beta * (d[r] * p[r]);
                }
            }
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            s += __shfl_xor(s, o, 16);
            h += __shfl_xor(h, o, 16);
        }
        if (go) {
            sum_P = DBL_MIN + s;
            const double H = (h / sum_P) + log(sum_P);
            found = b.step(H, logp);
            iter++;
        }
    }
    if (!live) return;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int m = r * 16 + sub;
        if (m < K) val[(size_t)row * K + m] = p[r] / sum_P;
    }
    if (sub == 0) {
        if (beta_out) beta_out[row] = last_beta;
        if (iters_out) iters_out[row] = iter;
        if (!found) atomicAdd(unconv, 1);
    }
}

// Rows of K > 512 values (up to 32 per thread, K <= 8192): one workgroup per
// row.  Each sum is a 64-lane butterfly per wave, then the four wave sums
// added in wave order by every thread, so all threads hold the same bits
// and step the bisection identically.
__global__ __launch_bounds__(256) void perplexity_knn_block(const double* __restrict__ dist, int N, int K, double logp,
                                                            double* __restrict__ val, double* __restrict__ beta_out,
                                                            int* __restrict__ iters_out, int* __restrict__ unconv)
{
    constexpr int R = 32;
    __shared__ double shs[4], shh[4];
    const int row = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double d[R], p[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int m = r * 256 + threadIdx.x;
        d[r] = m < K ? dist[(size_t)row * K + m] : 0.0;
        p[r] = 0.0;
    }
    Bisect b;
    bool found = false;
    int iter = 0;
    double sum_P = DBL_MIN, last_beta = b.beta;
    while (!found && iter < kBisectMaxIter) {
        const double beta = b.beta;
        last_beta = beta;
        double s = 0.0, h = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int m = r * 256 + threadIdx.x;
            if (m < K) {
                p[r] = exp(-beta * d[r]);
                s += p[r];
                h += beta * (d[r] * p[r]);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            s += __shfl_xor(s, o, 64);
            h += __shfl_xor(h, o, 64);
        }
        __syncthreads();  // the previous evaluation's sums have been read
        if (lane == 0) {
            shs[w] = s;
            shh[w] = h;
        }
        __syncthreads();
        s = ((shs[0] + shs[1]) + shs[2]) + shs[3];
        h = ((shh[0] + shh[1]) + shh[2]) + shh[3];
        sum_P = DBL_MIN + s;
        const double H = (h / sum_P) + log(sum_P);
        found = b.step(H, logp);
        iter++;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int m = r * 256 + threadIdx.x;
        if (m < K) val[(size_t)row * K + m] = p[r] / sum_P;
    }
    if (threadIdx.x == 0) {
        if (beta_out) beta_out[row] = last_beta;
        if (iters_out) iters_out[row] = iter;
        if (!found) atomicAdd(unconv, 1);
    }
}

// DD[n][m] for the dense theta = 0 path: textbook sum of squared differences
// (the order of the reference's euclidean_distance), or the reference's
// -2<x_n, x_m> (tsne.hpp:948-950) in compat mode.
__global__ __launch_bounds__(256) void sqdist_dense(const double* __restrict__ X, int N, int D, int compat,
                                                    double* __restrict__ DD)
{
    const size_t total = (size_t)N * N;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / N), m = (int)(e % N);
        const double* a = X + (size_t)n * D;
        const double* b = X + (size_t)m * D;
        double acc = 0.0;
        if (compat) {
            for (int d = 0; d < D; d++) acc = dadd_sep(acc, dmul_sep(a[d], b[d]));
            acc = -2.0 * acc;
        } else {
            for (int d = 0; d < D; d++) {
                const double t = dadd_sep(a[d], -b[d]);
                acc = dadd_sep(acc, dmul_sep(t, t));
            }
        }
        DD[e] = acc;
    }
}

// One block per dense row; the diagonal is DBL_MIN in every evaluation
// (tsne.hpp:548-550).
__global__ __launch_bounds__(256) void perplexity_dense(const double* __restrict__ DD, int N, double logp,
                                                        double* __restrict__ P, int* __restrict__ unconv)
{
    __shared__ double sh[4];
    const int n = blockIdx.x;
    const double* dr = DD + (size_t)n * N;
    double* pr = P + (size_t)n * N;
    Bisect b;
    bool found = false;
    int iter = 0;
    double sum_P = DBL_MIN;
    while (!found && iter < kBisectMaxIter) {
        const double beta = b.beta;
        double s = 0.0, h = 0.0;
        for (int m = threadIdx.x; m < N; m += 256) {
            const double d = dr[m];
            const double e = (m == n) ? DBL_MIN : exp(-beta * d);
            pr[m] = e;
            s += e;
            h += beta * (d * e);
        }
        s = block_sum(s, sh);
        h = block_sum(h, sh);
        sum_P = DBL_MIN + s;
        const double H = (h / sum_P) + log(sum_P);
        found = b.step(H, logp);
        iter++;
    }
    __syncthreads();
    for (int m = threadIdx.x; m < N; m += 256) pr[m] /= sum_P;
    if (threadIdx.x == 0 && !found) atomicAdd(unconv, 1);
}

// P[n][m] = P[m][n] = P[n][m] + P[m][n] for n < m (tsne.hpp:100-107).
__global__ __launch_bounds__(256) void symmetrize_dense(double* __restrict__ P, int N)
{
    const size_t total = (size_t)N * N;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / N), m = (int)(e % N);
        if (m <= n) continue;
        const double v = P[(size_t)n * N + m] + P[(size_t)m * N + n];
        P[(size_t)n * N + m] = v;
        P[(size_t)m * N + n] = v;
    }
}

__global__ __launch_bounds__(256) void sum_partials(const double* __restrict__ v, int64_t n, double* __restrict__ part)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += v[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void sum_final(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(256) void divide_by(double* __restrict__ v, int64_t n, const double* __restrict__ s)
{
    const double d = *s;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) v[i] /= d;
}

// Both orientations of every kNN edge, keyed (row << nb) | col.
__global__ __launch_bounds__(256) void sym_keys(const int* __restrict__ idx, const double* __restrict__ val, int N,
                                                int K, int nb, uint64_t* __restrict__ keys,
                                                uint64_t* __restrict__ vals)
{
    const size_t total = (size_t)N * K;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const uint64_t i = e / K, j = (uint32_t)idx[e];
        const uint64_t v = (uint64_t)__double_as_longlong(val[e]);
        keys[2 * e] = (i << nb) | j;
        vals[2 * e] = v;
        keys[2 * e + 1] = (j << nb) | i;
        vals[2 * e + 1] = v;
    }
}

__global__ __launch_bounds__(256) void sym_flags(const uint64_t* __restrict__ keys, size_t M, uint32_t* __restrict__ f)
{
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < M; e += (size_t)gridDim.x * 256)
        f[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

// Unique (row, col) -> CSR entry with value (p_rc + p_cr)/2, where an absent
// orientation contributes nothing (tsne.hpp:275-307).
__global__ __launch_bounds__(256) void sym_write(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals,
                                                 size_t M, const uint32_t* __restrict__ pos, int nb, int N,
                                                 int* __restrict__ row_ptr, int* __restrict__ col,
                                                 double* __restrict__ sval)
{
    const uint64_t cmask = (1ull << nb) - 1;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < M; e += (size_t)gridDim.x * 256) {
        const uint64_t k = keys[e];
        if (e > 0 && keys[e - 1] == k) continue;
        const uint32_t u = pos[e];
        const int r = (int)(k >> nb);
        double v = __longlong_as_double((long long)vals[e]);
        if (e + 1 < M && keys[e + 1] == k) v = v + __longlong_as_double((long long)vals[e + 1]);
        col[u] = (int)(k & cmask);
        sval[u] = v / 2.0;
        if (e == 0 || (int)(keys[e - 1] >> nb) != r) row_ptr[r] = (int)u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) row_ptr[N] = (int)pos[M];
}

// ---- symmetrisation by mutual-edge search (K <= kSymFastK)
// The union pattern E u E^T of a kNN graph is E (every row's own K entries)
// plus, in row j, the entries (j, i) of the edges i -> j whose reverse
// j -> i is not in E.  With every row's neighbours sorted by index, one
// binary search per edge decides whether its reverse exists; the reverse-only
// entries (a few percent of the edges on manifold data) are counted per
// target row and sorted by (row, column) (radix sort of those edges only),
// and each output row is the merge of its own sorted entries with its
// reverse-only ones.  Positions come from counts and binary searches, so the
// CSR is the same whatever order the atomics ran in: columns ascending in
// every row, values as tsne.hpp:275-307.
constexpr int kSymFastK = 256;

// Ascending wave bitonic sort of 64*R (key, payload) pairs at position
// r*64 + lane, ordered by key then payload.
template <int R>
TKA_DEV void wave_sort_u32_pairs(uint32_t (&k)[R], uint32_t (&v)[R])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int w = 2; w <= 64 * R; w <<= 1) {
#pragma unroll
        for (int j = w >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int jr = j >> 6;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if ((r & jr) == 0) {
                        const int r2 = r | jr;
                        const bool up = (((r * 64 + lane) & w) == 0);
                        const bool gt = k[r] > k[r2] || (k[r] == k[r2] && v[r] > v[r2]);
                        const bool sw = up ? gt : !gt;
                        const uint32_t k0 = k[r], k1 = k[r2], v0 = v[r], v1 = v[r2];
                        k[r] = sw ? k1 : k0;
                        k[r2] = sw ? k0 : k1;
                        v[r] = sw ? v1 : v0;
                        v[r2] = sw ? v0 : v1;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int p = r * 64 + lane;
                    const bool up = (p & w) == 0;
                    const bool lower = (p & j) == 0;
                    const uint32_t ok = __shfl_xor(k[r], j, 64), ov = __shfl_xor(v[r], j, 64);
                    const bool oless = ok < k[r] || (ok == k[r] && ov < v[r]);
                    const bool take = (lower == up) ? oless : !oless;
                    k[r] = take ? ok : k[r];
                    v[r] = take ? ov : v[r];
                }
            }
        }
    }
}

// Every row's neighbours sorted by index, with their values.  One wave per row.
template <int R>
__global__ __launch_bounds__(256) void sym_sort_rows(const int* __restrict__ idx, const double* __restrict__ val,
                                                     int N, int K, int* __restrict__ scol, double* __restrict__ sval)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const size_t base = (size_t)i * K;
    uint32_t key[R], pos[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int p = r * 64 + lane;
        key[r] = p < K ? (uint32_t)idx[base + p] : 0xffffffffu;
        pos[r] = (uint32_t)p;
    }
    wave_sort_u32_pairs<R>(key, pos);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int p = r * 64 + lane;
        if (p < K) {
            scol[base + p] = (int)key[r];
            sval[base + p] = val[base + pos[r]];
        }
    }
}

// First position in the sorted row a[0, n) whose value is >= x.
TKA_DEV int lower_bound_i(const int* __restrict__ a, int n, int x)
{
    int lo = 0;
    while (n > 0) {
        const int h = n >> 1;
        if (a[lo + h] < x) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

// Work slot of a workgroup when rows are visited in a locality order:
// workgroups go to the eight XCDs in turn, so each XCD takes a contiguous
// eighth of the slots and its L2 holds the rows its neighbourhoods touch.
// grid = 8 * per; returns -1 past the last of g workgroups.
TKA_DEV int xcd_block(int g)
{
    const int per = (int)(gridDim.x >> 3);
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    return b < g ? b : -1;
}
inline unsigned xcd_grid(unsigned g) { return (g + 7) / 8 * 8; }

// For every edge i -> j: the position of i in row j (mpos), or -1 with the
// reverse-only entry (j, i) counted for row j.  Rows in `order` (a locality
// order of the points, or the identity when null), so that the binary
// searches of nearby rows hit the same rows j in L2.
__global__ __launch_bounds__(256) void sym_mutual(const int* __restrict__ scol, int N, int K,
                                                  const uint32_t* __restrict__ order, int g, int* __restrict__ mpos,
                                                  uint32_t* __restrict__ ecnt)
{
    const int b = xcd_block(g);
    if (b < 0) return;
    const size_t t = (size_t)b * 256 + threadIdx.x;
    if (t < (size_t)N * K) {
        const int w = (int)(t / K), p = (int)(t % K);
        const int i = order ? (int)order[w] : w;
        const size_t e = (size_t)i * K + p;
        const int j = scol[e];
        const int* rj = scol + (size_t)j * K;
        const int q = lower_bound_i(rj, K, i);
        const bool found = q < K && rj[q] == i;
        mpos[e] = found ? q : -1;
        if (!found) atomicAdd(&ecnt[j], 1u);
    }
}

// Output row lengths: K own entries plus the reverse-only ones.
__global__ __launch_bounds__(256) void sym_lengths(const uint32_t* __restrict__ ecnt, int N, int K,
                                                   uint32_t* __restrict__ len)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) len[i] = (uint32_t)K + ecnt[i];
}

// Edges of row i without a reverse edge (one wave per row).
template <int R>
__global__ __launch_bounds__(256) void sym_count_oneway(const int* __restrict__ mpos, int N, int K,
                                                        uint32_t* __restrict__ n1)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    int c = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int p = r * 64 + lane;
        c += __popcll(__ballot(p < K && mpos[(size_t)i * K + p] < 0));
    }
    if (lane == 0) n1[i] = (uint32_t)c;
}

// Reverse-only entries (j, i) as keys (j << nb) | i with their value p_ij,
// row i's at base[i] + their rank in the row (one wave per row).
template <int R>
__global__ __launch_bounds__(256) void sym_gather_extras(const int* __restrict__ scol, const double* __restrict__ sval,
                                                         const int* __restrict__ mpos, int N, int K, int nb,
                                                         const uint32_t* __restrict__ base, uint64_t* __restrict__ ekey,
                                                         uint64_t* __restrict__ eval)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t at = base[i];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int p = r * 64 + lane;
        const size_t e = (size_t)i * K + p;
        const bool one = p < K && mpos[e] < 0;
        const uint64_t m = __ballot(one);
        if (one) {
            const uint32_t slot = at + (uint32_t)__popcll(m & lt);
            ekey[slot] = ((uint64_t)(uint32_t)scol[e] << nb) | (uint32_t)i;
            eval[slot] = (uint64_t)__double_as_longlong(sval[e]);
        }
        at += (uint32_t)__popcll(m);
    }
}

// First position in the sorted keys a[0, n) whose value is >= x.
TKA_DEV int lower_bound_u64(const uint64_t* __restrict__ a, int n, uint64_t x)
{
    int lo = 0;
    while (n > 0) {
        const int h = n >> 1;
        if (a[lo + h] < x) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

// One wave per output row: own entries (p_ij + p_ji)/2 or p_ij/2, reverse-only
// entries p_ji/2, merged by column: an entry's position is its index in its
// own sorted list plus the number of entries of the other list before it
// (a binary search).  The row's reverse-only entries are the segment
// [row_ptr[i] - i K, + ecnt[i]) of the sorted keys.  Rows in `order` (see
// sym_mutual).
__global__ __launch_bounds__(256) void sym_write_rows(const int* __restrict__ scol, const double* __restrict__ sval,
                                                      const int* __restrict__ mpos, int N, int K, int nb,
                                                      const uint32_t* __restrict__ order, int g,
                                                      const int* __restrict__ row_ptr, const uint32_t* __restrict__ ecnt,
                                                      const uint64_t* __restrict__ ekey,
                                                      const uint64_t* __restrict__ eval, int* __restrict__ col,
                                                      double* __restrict__ pval)
{
    const int lane = threadIdx.x & 63;
    const int b = xcd_block(g);
    if (b < 0) return;
    const int slot = b * 4 + (threadIdx.x >> 6);
    if (slot >= N) return;
    const int i = order ? (int)order[slot] : slot;
    const size_t own = (size_t)i * K;
    const int ne = (int)ecnt[i];
    const size_t eb = (size_t)(row_ptr[i] - (int64_t)i * K);
    const size_t ob = (size_t)row_ptr[i];
    const uint64_t* ek = ekey + eb;
    const uint64_t rowkey = (uint64_t)i << nb;
    for (int p = lane; p < K; p += 64) {
        const int c = scol[own + p];
        const int q = mpos[own + p];
        const double v = q >= 0 ? (sval[own + p] + sval[(size_t)c * K + q]) / 2.0 : sval[own + p] / 2.0;
        const int at = p + lower_bound_u64(ek, ne, rowkey | (uint32_t)c);
        col[ob + at] = c;
        pval[ob + at] = v;
    }
    const uint64_t cmask = (1ull << nb) - 1;
    for (int e = lane; e < ne; e += 64) {
        const int x = (int)(ek[e] & cmask);
        const int at = e + lower_bound_i(scol + own, K, x);
        col[ob + at] = x;
        pval[ob + at] = __longlong_as_double((long long)eval[eb + e]) / 2.0;
    }
}

}  // namespace

int perplexity_knn_device(const double* dist, int N, int K, double perplexity, double* val, double* beta, int* iters,
                          MemBudget* mb, hipStream_t s)
{
    require(K <= 8192, TKA_E_WRONG_PARAMETER, "perplexity calibration supports K <= 8192");
    DevBuf<int> unconv(1, mb);
    TKA_HIP(hipMemsetAsync(unconv.get(), 0, sizeof(int), s));
    const double logp = std::log(perplexity);
    if (K <= 128) {
        const dim3 g(ceil_div(N, 16));
#define TKA_PK(R) \
    case R: hipLaunchKernelGGL(perplexity_knn_packed<R>, g, dim3(256), 0, s, dist, N, K, logp, val, beta, iters, \
                               unconv.get()); break
        switch ((K + 15) / 16) {
            TKA_PK(1); TKA_PK(2); TKA_PK(3); TKA_PK(4); TKA_PK(5); TKA_PK(6); TKA_PK(7); TKA_PK(8);
        }
#undef TKA_PK
    } else if (K <= 512) {
        hipLaunchKernelGGL(perplexity_knn, dim3(ceil_div(N, 4)), dim3(256), 0, s, dist, N, K, logp, val, beta, iters,
                           unconv.get());
    } else {
        hipLaunchKernelGGL(perplexity_knn_block, dim3(N), dim3(256), 0, s, dist, N, K, logp, val, beta, iters,
                           unconv.get());
    }
    TKA_LAUNCH_CHECK();
    int h = 0;
    read_to_host(&h, unconv.get(), sizeof(int), s);
    return h;
}

void sqdist_dense_device(const double* X, int N, int D, int dd_mode, double* DD, hipStream_t s)
{
    const size_t total = (size_t)N * N;
    hipLaunchKernelGGL(sqdist_dense, dim3(std::min<size_t>(65535, ceil_div(total, 256))), dim3(256), 0, s, X, N, D,
                       dd_mode == TKA_DD_TAPKEE_COMPAT ? 1 : 0, DD);
    TKA_LAUNCH_CHECK();
}

int perplexity_dense_device(const double* DD, int N, double perplexity, double* P, MemBudget* mb, hipStream_t s)
{
    DevBuf<int> unconv(1, mb);
    TKA_HIP(hipMemsetAsync(unconv.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(perplexity_dense, dim3(N), dim3(256), 0, s, DD, N, std::log(perplexity), P, unconv.get());
    TKA_LAUNCH_CHECK();
    int h = 0;
    read_to_host(&h, unconv.get(), sizeof(int), s);
    return h;
}

void normalize_sum_device(double* v, int64_t n, MemBudget* mb, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
    DevBuf<double> part(nb, mb), tot(1, mb);
    hipLaunchKernelGGL(sum_partials, dim3(nb), dim3(256), 0, s, v, n, part.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, s, part.get(), nb, tot.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(divide_by, dim3(nb), dim3(256), 0, s, v, n, tot.get());
    TKA_LAUNCH_CHECK();
}

void symmetrize_dense_device(double* P, int N, MemBudget* mb, hipStream_t s)
{
    const size_t total = (size_t)N * N;
    hipLaunchKernelGGL(symmetrize_dense, dim3(std::min<size_t>(65535, ceil_div(total, 256))), dim3(256), 0, s, P, N);
    TKA_LAUNCH_CHECK();
    normalize_sum_device(P, (int64_t)total, mb, s);
}

// K beyond kSymFastK: both orientations of every edge sorted by (row, col)
// (34-bit keys at N = 100k), duplicates merged.
static void symmetrize_knn_sorted(const int* idx, const double* val, int N, int K, Csr& out, MemBudget* mb,
                                  hipStream_t s)
{
    int nb = 1;
    while ((1ll << nb) < (long long)N) nb++;
    const size_t M = 2 * (size_t)N * K;
    DevBuf<uint64_t> k0(M, mb), v0(M, mb), k1(M, mb), v1(M, mb);
    DevBuf<uint32_t> scratch(radix_sort_scratch_words(M), mb);
    TKA_HIP(hipMemsetAsync(scratch.get(), 0, scratch.size() * sizeof(uint32_t), s));
    const unsigned g = std::min<size_t>(65535, ceil_div((size_t)N * K, 256));
    hipLaunchKernelGGL(sym_keys, dim3(g), dim3(256), 0, s, idx, val, N, K, nb, k0.get(), v0.get());
    TKA_LAUNCH_CHECK();
    const bool alt = radix_sort_pairs<uint64_t, uint64_t>(k0.get(), v0.get(), k1.get(), v1.get(), M, 2 * nb,
                                                          scratch.get(), s);
    require(radix_sort_fault(scratch.get(), s) == 0, TKA_E_INTERNAL, "radix sort look-back gave up");
    const uint64_t* ks = alt ? k1.get() : k0.get();
    const uint64_t* vs = alt ? v1.get() : v0.get();
    DevBuf<uint32_t> flags(M, mb), pos(M + 1, mb), sscr(exclusive_scan_scratch_words(M), mb);
    const unsigned gm = std::min<size_t>(65535, ceil_div(M, 256));
    hipLaunchKernelGGL(sym_flags, dim3(gm), dim3(256), 0, s, ks, M, flags.get());
    TKA_LAUNCH_CHECK();
    exclusive_scan_u32(flags.get(), pos.get(), M, sscr.get(), s);
    uint32_t nnz = 0;
    read_to_host(&nnz, pos.get() + M, sizeof(uint32_t), s);
    out.nnz = nnz;
    out.row.alloc(N + 1, mb);
    out.col.alloc(nnz, mb);
    out.val.alloc(nnz, mb);
    hipLaunchKernelGGL(sym_write, dim3(gm), dim3(256), 0, s, ks, vs, M, pos.get(), nb, N, out.row.get(),
                       out.col.get(), out.val.get());
    TKA_LAUNCH_CHECK();
}

void symmetrize_knn_device(const int* idx, const double* val, int N, int K, Csr& out, MemBudget* mb, hipStream_t s,
                           const uint32_t* order)
{
    if (K > kSymFastK) {
        symmetrize_knn_sorted(idx, val, N, K, out, mb, s);
        return;
    }
    const size_t NK = (size_t)N * K;
    int nb = 1;
    while ((1ll << nb) < (long long)N) nb++;
    DevBuf<int> scol(NK, mb), mpos(NK, mb);
    DevBuf<double> sval(NK, mb);
    DevBuf<uint32_t> cnt(N, mb);  // reverse-only entries per row
    DevBuf<uint32_t> len(N, mb), sscr(exclusive_scan_scratch_words(N), mb);
    TKA_HIP(hipMemsetAsync(cnt.get(), 0, cnt.size() * sizeof(uint32_t), s));
    const unsigned gr = ceil_div(N, 4);
    if (K <= 64)
        hipLaunchKernelGGL(sym_sort_rows<1>, dim3(gr), dim3(256), 0, s, idx, val, N, K, scol.get(), sval.get());
    else if (K <= 128)
        hipLaunchKernelGGL(sym_sort_rows<2>, dim3(gr), dim3(256), 0, s, idx, val, N, K, scol.get(), sval.get());
    else
        hipLaunchKernelGGL(sym_sort_rows<4>, dim3(gr), dim3(256), 0, s, idx, val, N, K, scol.get(), sval.get());
    TKA_LAUNCH_CHECK();
    const unsigned gm = ceil_div(NK, 256);
    hipLaunchKernelGGL(sym_mutual, dim3(xcd_grid(gm)), dim3(256), 0, s, scol.get(), N, K, order, (int)gm, mpos.get(),
                       cnt.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(sym_lengths, dim3(ceil_div(N, 256)), dim3(256), 0, s, cnt.get(), N, K, len.get());
    TKA_LAUNCH_CHECK();
    out.row.alloc(N + 1, mb);
    exclusive_scan_u32(len.get(), reinterpret_cast<uint32_t*>(out.row.get()), N, sscr.get(), s);
    int nnz = 0;
    read_to_host(&nnz, out.row.get() + N, sizeof(int), s);
    require(nnz >= (int64_t)NK && (size_t)nnz <= 2 * NK, TKA_E_INTERNAL, "symmetrisation: bad entry count");
    out.nnz = nnz;
    out.col.alloc(nnz, mb);
    out.val.alloc(nnz, mb);
    const size_t M = (size_t)nnz - NK;  // reverse-only entries
    DevBuf<uint64_t> ek0(std::max<size_t>(M, 1), mb), ev0(std::max<size_t>(M, 1), mb);
    DevBuf<uint64_t> ek1(std::max<size_t>(M, 1), mb), ev1(std::max<size_t>(M, 1), mb);
    DevBuf<uint32_t> escr(radix_sort_scratch_words(std::max<size_t>(M, 1)), mb);
    TKA_HIP(hipMemsetAsync(escr.get(), 0, escr.size() * sizeof(uint32_t), s));
    DevBuf<uint32_t> n1(N, mb), base((size_t)N + 1, mb);
#define TKA_SYM_R(R)                                                                                                \
    hipLaunchKernelGGL(sym_count_oneway<R>, dim3(gr), dim3(256), 0, s, mpos.get(), N, K, n1.get());              \
    TKA_LAUNCH_CHECK();                                                                                             \
    exclusive_scan_u32(n1.get(), base.get(), N, sscr.get(), s);                                                     \
    hipLaunchKernelGGL(sym_gather_extras<R>, dim3(gr), dim3(256), 0, s, scol.get(), sval.get(), mpos.get(), N, K, \
                       nb, base.get(), ek0.get(), ev0.get());
    if (K <= 64) {
        TKA_SYM_R(1);
    } else if (K <= 128) {
        TKA_SYM_R(2);
    } else {
        TKA_SYM_R(4);
    }
#undef TKA_SYM_R
    TKA_LAUNCH_CHECK();
    bool alt = false;
    if (M > 0) alt = radix_sort_pairs<uint64_t, uint64_t>(ek0.get(), ev0.get(), ek1.get(), ev1.get(), M, 2 * nb,
                                                          escr.get(), s);
    // a sort that gave up leaves keys that only misplace entries within
    // their rows (every write stays in the row); it is reported below
    hipLaunchKernelGGL(sym_write_rows, dim3(xcd_grid(gr)), dim3(256), 0, s, scol.get(), sval.get(), mpos.get(), N, K,
                       nb, order, (int)gr, out.row.get(), cnt.get(), alt ? ek1.get() : ek0.get(),
                       alt ? ev1.get() : ev0.get(), out.col.get(), out.val.get());
    TKA_LAUNCH_CHECK();
    require(M == 0 || radix_sort_fault(escr.get(), s) == 0, TKA_E_INTERNAL, "radix sort look-back gave up");
}

}  // namespace tka
