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
    TKA_HIP(hipMemcpyAsync(&h, unconv.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
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
    TKA_HIP(hipMemcpyAsync(&h, unconv.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
    return h;
}

double normalize_sum_device(double* v, int64_t n, MemBudget* mb, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
    DevBuf<double> part(nb, mb), tot(1, mb);
    hipLaunchKernelGGL(sum_partials, dim3(nb), dim3(256), 0, s, v, n, part.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, s, part.get(), nb, tot.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(divide_by, dim3(nb), dim3(256), 0, s, v, n, tot.get());
    TKA_LAUNCH_CHECK();
    double h = 0.0;
    TKA_HIP(hipMemcpyAsync(&h, tot.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
    return h;
}

void symmetrize_dense_device(double* P, int N, MemBudget* mb, hipStream_t s)
{
    const size_t total = (size_t)N * N;
    hipLaunchKernelGGL(symmetrize_dense, dim3(std::min<size_t>(65535, ceil_div(total, 256))), dim3(256), 0, s, P, N);
    TKA_LAUNCH_CHECK();
    normalize_sum_device(P, (int64_t)total, mb, s);
}

void symmetrize_knn_device(const int* idx, const double* val, int N, int K, Csr& out, MemBudget* mb, hipStream_t s)
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
    TKA_HIP(hipMemcpyAsync(&nnz, pos.get() + M, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
    out.nnz = nnz;
    out.row.alloc(N + 1, mb);
    out.col.alloc(nnz, mb);
    out.val.alloc(nnz, mb);
    hipLaunchKernelGGL(sym_write, dim3(gm), dim3(256), 0, s, ks, vs, M, pos.get(), nb, N, out.row.get(),
                       out.col.get(), out.val.get());
    TKA_LAUNCH_CHECK();
}

}  // namespace tka
