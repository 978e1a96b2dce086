// Attractive forces + optimiser step (fused), exact repulsion, KL.
#include "forces.hpp"

#include <type_traits>

#include <cfloat>

#include "attract.hpp"
#include "radix_sort.hpp"

namespace tka {
namespace {

template <class T>
TKA_DEV T sgn(T x)
{
    return x == (T)0 ? (T)0 : (x < (T)0 ? (T)-1 : (T)1);  // tsne.hpp:53-56
}

template <int kBlock>
TKA_DEV double z_total(const double* __restrict__ zpart, int nz, double* sh)
{
    // the order of a loop over b = threadIdx.x, + kBlock, ..., with kU loads
    // in flight at a time (a loop that waits on each load makes as many
    // memory round trips as partials per thread; the block size is a
    // constant, as reading blockDim would wait on the rows' loads)
    constexpr int kU = 8;
    double z = 0.0;
    for (int b0 = threadIdx.x; b0 < nz; b0 += kU * kBlock) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int b = b0 + u * kBlock;
            v[u] = b < nz ? zpart[b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; u++) z += v[u];
    }
    return block_sum(z, sh);
}

// Attraction of row i (all kG lanes of the group call this).
template <class T, int DIMS>
TKA_DEV void row_attraction(const T* __restrict__ Y, const int* __restrict__ row_ptr, const int* __restrict__ col,
                            const T* __restrict__ val, T exag, int i, const T (&yi)[DIMS], T (&acc)[DIMS])
{
    const int gl = threadIdx.x % kG;
#pragma unroll
    for (int d = 0; d < DIMS; d++) acc[d] = (T)0;
    const int e1 = row_ptr[i + 1];
    for (int e = row_ptr[i] + gl; e < e1; e += kG) {
        const int j = col[e];
        T diff[DIMS];
        T D = (T)0;
#pragma unroll
        for (int d = 0; d < DIMS; d++) {
            diff[d] = yi[d] - Y[(size_t)j * DIMS + d];
            D += diff[d] * diff[d];
        }
        const T m = (exag * val[e]) / ((T)1 + D);
#pragma unroll
        for (int d = 0; d < DIMS; d++) acc[d] += m * diff[d];
    }
#pragma unroll
    for (int d = 0; d < DIMS; d++) acc[d] = group_sum(acc[d]);
}

template <class T, int DIMS>
__global__ __launch_bounds__(256) void tsne_step(const T* __restrict__ Y, T* __restrict__ Ynext, T* __restrict__ uY,
                                                 T* __restrict__ gains, const T* __restrict__ neg,
                                                 const T* __restrict__ pos_pre, const int* __restrict__ row_ptr,
                                                 const int* __restrict__ col, const T* __restrict__ val,
                                                 const double* __restrict__ zpart, int nz, int N, T momentum, T eta,
                                                 T exag, double* __restrict__ box_part, T* __restrict__ dC_out)
{
    __shared__ double sh[4];
    __shared__ double sbox[7][4];
    const double Z = z_total<256>(zpart, nz, sh);
    const int g = threadIdx.x / kG, gl = threadIdx.x % kG;
    const int i = blockIdx.x * kRowsPerBlock + g;
    T yi[DIMS], acc[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        yi[d] = (i < N) ? Y[(size_t)i * DIMS + d] : (T)0;
        acc[d] = (T)0;
    }
    if (row_ptr != nullptr && i < N) row_attraction<T, DIMS>(Y, row_ptr, col, val, exag, i, yi, acc);
    // statistics of Ynext: sums per dim (<= 3), min/max of the first two
    double bx[7] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    if (gl == 0 && i < N) {
        T yn[DIMS];
#pragma unroll
        for (int d = 0; d < DIMS; d++) {
            const size_t o = (size_t)i * DIMS + d;
            const T pos = (row_ptr != nullptr) ? acc[d] : pos_pre[o];
            const T dC = pos - (T)((double)neg[o] / Z);
            if (dC_out) dC_out[o] = dC;
            T gn = gains[o];
            T u = uY[o];
            gn = (sgn(dC) != sgn(u)) ? (gn + (T)0.2) : (gn * (T)0.8);
            if (gn < (T)0.01) gn = (T)0.01;
            u = momentum * u - eta * gn * dC;
            gains[o] = gn;
            uY[o] = u;
            yn[d] = yi[d] + u;
            Ynext[o] = yn[d];
            if (d < 3) bx[d] = (double)yn[d];
        }
        bx[3] = bx[5] = (double)yn[0];
        bx[4] = bx[6] = (double)yn[DIMS > 1 ? 1 : 0];
    }
    if (box_part != nullptr) {
#pragma unroll
        for (int k = 0; k < 3; k++) bx[k] = wave_sum(bx[k]);
        bx[3] = wave_min(bx[3]);
        bx[4] = wave_min(bx[4]);
        bx[5] = wave_max(bx[5]);
        bx[6] = wave_max(bx[6]);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < 7; k++) sbox[k][w] = bx[k];
        __syncthreads();
        if (threadIdx.x == 0) {
            double r[7] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
            for (int ww = 0; ww < 4; ww++) {
                for (int k = 0; k < 3; k++) r[k] += sbox[k][ww];
                r[3] = fmin(r[3], sbox[3][ww]);
                r[4] = fmin(r[4], sbox[4][ww]);
                r[5] = fmax(r[5], sbox[5][ww]);
                r[6] = fmax(r[6], sbox[6][ww]);
            }
            double* o = box_part + (size_t)kBoxStride * blockIdx.x;
            for (int k = 0; k < 7; k++) o[k] = r[k];
            o[7] = 0.0;
        }
    }
}

// Attraction alone (quadtree.hpp:357-384): one row group of kG lanes per row,
// a grid-stride loop over row groups (attract_row).
template <class T, int DIMS>
__global__ __launch_bounds__(256) void attract(const T* __restrict__ Y, const int* __restrict__ row_ptr,
                                               const int* __restrict__ col, const T* __restrict__ val, T exag,
                                               int N, T* __restrict__ pos)
{
    const int g = threadIdx.x / kG, gl = threadIdx.x % kG;
    for (int rb = blockIdx.x; rb * kRowsPerBlock < N; rb += gridDim.x) {
        const int i = rb * kRowsPerBlock + g;
        if (i < N)  // whole row groups skip together; group sums stay inside a group
            attract_row<T, DIMS>(Y, row_ptr, col, val, exag, i, gl, pos);
    }
}

// The optimiser step from precomputed attraction and repulsion, one thread
// per row (tsne.hpp:152-167 and :344-347); also the box partials of Ynext.
// Workgroups of 256 spread the rows over 4x more CUs than 1024 would (the
// launch is latency-bound); the key kernel reduces the N/256 box partials.
constexpr int kUpdBlock = 256;

static_assert(kUpdBlock == 256, "z_total's block size");
template <class T, int DIMS>
__global__ __launch_bounds__(kUpdBlock) void tsne_update(const T* __restrict__ Y, T* __restrict__ Ynext,
                                                         T* __restrict__ uY, T* __restrict__ gains,
                                                         const T* __restrict__ neg, const T* __restrict__ pos,
                                                         const double* __restrict__ zpart, int nz, int N, T momentum,
                                                         T eta, double* __restrict__ box_part, T* __restrict__ dC_out)
{
    constexpr int kWaves = kUpdBlock / 64;
    __shared__ double sh[kWaves];
    __shared__ double sbox[7][kWaves];
    const int i = blockIdx.x * kUpdBlock + threadIdx.x;
    // The row's operands are requested before the Z reduction, so its loads
    // and theirs are in flight together.
    // rows past N read row N-1 (unconditional loads, all in flight
    // together) and are not written
    T rp[DIMS], rn[DIMS], rg[DIMS], ru[DIMS], ry[DIMS];
    const size_t ic = (size_t)min(i, N - 1);
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        const size_t o = ic * DIMS + d;
        rp[d] = pos[o];
        rn[d] = neg[o];
        rg[d] = gains[o];
        ru[d] = uY[o];
        ry[d] = Y[o];
    }
    const double Z = z_total<256>(zpart, nz, sh);
    double bx[7] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    if (i < N) {
        T yn[DIMS];
#pragma unroll
        for (int d = 0; d < DIMS; d++) {
            const size_t o = (size_t)i * DIMS + d;
            const T dC = rp[d] - (T)((double)rn[d] / Z);
            if (dC_out) dC_out[o] = dC;
            T gn = rg[d];
            T u = ru[d];
            gn = (sgn(dC) != sgn(u)) ? (gn + (T)0.2) : (gn * (T)0.8);
            if (gn < (T)0.01) gn = (T)0.01;
            u = momentum * u - eta * gn * dC;
            gains[o] = gn;
            uY[o] = u;
            yn[d] = ry[d] + u;
            Ynext[o] = yn[d];
            if (d < 3) bx[d] = (double)yn[d];
        }
        bx[3] = bx[5] = (double)yn[0];
        bx[4] = bx[6] = (double)yn[DIMS > 1 ? 1 : 0];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) bx[k] = wave_sum(bx[k]);
    bx[3] = wave_min(bx[3]);
    bx[4] = wave_min(bx[4]);
    bx[5] = wave_max(bx[5]);
    bx[6] = wave_max(bx[6]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 7; k++) sbox[k][w] = bx[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[7] = {0.0, 0.0, 0.0, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int ww = 0; ww < kWaves; ww++) {
            for (int k = 0; k < 3; k++) r[k] += sbox[k][ww];
            r[3] = fmin(r[3], sbox[3][ww]);
            r[4] = fmin(r[4], sbox[4][ww]);
            r[5] = fmax(r[5], sbox[5][ww]);
            r[6] = fmax(r[6], sbox[6][ww]);
        }
        double* o = box_part + (size_t)kBoxStride * blockIdx.x;
        for (int k = 0; k < 7; k++) o[k] = r[k];
        o[7] = 0.0;
    }
}

// zeroMean (tsne.hpp:488-518) from box partials: Y -= sum/N per dimension.
template <class T, int DIMS>
__global__ __launch_bounds__(256) void center_rows(T* __restrict__ Y, int N, const double* __restrict__ part, int nb)
{
    __shared__ double sh[4];
    double mean[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) v += part[(size_t)kBoxStride * b + d];
        mean[d] = block_sum(v, sh) / (double)N;
    }
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) {
#pragma unroll
        for (int d = 0; d < DIMS; d++) Y[(size_t)i * DIMS + d] = (T)((double)Y[(size_t)i * DIMS + d] - mean[d]);
    }
}

// Per-block sums of each dimension (<= 3) in the box-partial layout.
template <class T, int DIMS>
__global__ __launch_bounds__(256) void sum_rows(const T* __restrict__ Y, int N, double* __restrict__ part)
{
    __shared__ double sh[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        const double v = block_sum(i < N ? (double)Y[(size_t)i * DIMS + d] : 0.0, sh);
        if (threadIdx.x == 0) part[(size_t)kBoxStride * blockIdx.x + d] = v;
    }
}

template <class T, int DIMS>
__global__ __launch_bounds__(256) void grad_combine(const T* __restrict__ Y, const T* __restrict__ neg,
                                                    const T* __restrict__ pos_pre, const int* __restrict__ row_ptr,
                                                    const int* __restrict__ col, const T* __restrict__ val, T exag,
                                                    const double* __restrict__ zpart, int nz, int N,
                                                    T* __restrict__ dC)
{
    __shared__ double sh[4];
    const double Z = z_total<256>(zpart, nz, sh);
    const int g = threadIdx.x / kG, gl = threadIdx.x % kG;
    const int i = blockIdx.x * kRowsPerBlock + g;
    T yi[DIMS], acc[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        yi[d] = (i < N) ? Y[(size_t)i * DIMS + d] : (T)0;
        acc[d] = (T)0;
    }
    if (row_ptr != nullptr && i < N) row_attraction<T, DIMS>(Y, row_ptr, col, val, exag, i, yi, acc);
    if (gl == 0 && i < N) {
#pragma unroll
        for (int d = 0; d < DIMS; d++) {
            const size_t o = (size_t)i * DIMS + d;
            const T pos = (row_ptr != nullptr) ? acc[d] : pos_pre[o];
            dC[o] = pos - (T)((double)neg[o] / Z);
        }
    }
}

// 1/(1 + D).  FP32: the hardware reciprocal (1 ulp) with one Newton-Raphson
// step, r + r(1 - a r) (3 instructions against 11 for the IEEE division); a
// is clamped to FLT_MAX so that an overflowing D gives 0, not inf * 0 = NaN.
// FP64: the division.
template <class T>
TKA_DEV T recip_one_plus(T D)
{
    if constexpr (std::is_same<T, float>::value) {
        const float a = fminf(1.0f + D, FLT_MAX);
        const float r = __builtin_amdgcn_rcpf(a);
        return fmaf(r, fmaf(-a, r, 1.0f), r);
    } else {
        return (T)1 / ((T)1 + D);
    }
}

// Exact all-pairs repulsion, tiled in two dimensions so that small N still
// fills the device: workgroup (x, y) takes rows 256x .. 256x+255 (one per
// thread) against the columns of slice y, streamed through LDS 256 at a time.
// Each slice leaves its per-row sums in part (slice-major, N x DIMS each) and
// its Z partial in zpart[y * gridDim.x + x]; exact_reduce adds the slices in
// order, so results do not depend on scheduling.
template <class T, int DIMS, bool kCompat, bool kDense>
__global__ __launch_bounds__(256) void exact_tiles(const T* __restrict__ Y, int N, int cols, const T* __restrict__ P,
                                                   T exag, T* __restrict__ part_rep, T* __restrict__ part_att,
                                                   double* __restrict__ zpart)
{
    __shared__ T sY[256 * DIMS];
    __shared__ double sh[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * cols, c1 = min(N, c0 + cols);
    T yi[DIMS], rep[DIMS], att[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        yi[d] = (i < N) ? Y[(size_t)i * DIMS + d] : (T)0;
        rep[d] = att[d] = (T)0;
    }
    T z = (T)0;
    for (int j0 = c0; j0 < c1; j0 += 256) {
        const int jn = min(256, c1 - j0);
        __syncthreads();
        for (int e = threadIdx.x; e < jn * DIMS; e += 256) sY[e] = Y[(size_t)j0 * DIMS + e];
        __syncthreads();
        if (i < N) {
            for (int jj = 0; jj < jn; jj++) {
                const int j = j0 + jj;
                if (j == i) continue;
                T diff[DIMS];
                T D = (T)0;
#pragma unroll
                for (int d = 0; d < DIMS; d++) {
                    diff[d] = yi[d] - sY[jj * DIMS + d];
                    if (kCompat)
                        D += yi[d] * sY[jj * DIMS + d];
                    else
                        D += diff[d] * diff[d];
                }
                if (kCompat) D = (T)-2 * D;
                const T q = recip_one_plus(D);
                z += q;
                const T q2 = q * q;
                T pq = (T)0;
                if (kDense) pq = exag * P[(size_t)j * N + i] * q;  // P symmetric: P[i][j] == P[j][i]
#pragma unroll
                for (int d = 0; d < DIMS; d++) {
                    rep[d] += q2 * diff[d];
                    if (kDense) att[d] += pq * diff[d];
                }
            }
        }
    }
    if (i < N) {
        const size_t o = ((size_t)blockIdx.y * N + i) * DIMS;
#pragma unroll
        for (int d = 0; d < DIMS; d++) {
            part_rep[o + d] = rep[d];
            if (kDense) part_att[o + d] = att[d];
        }
    }
    const double zs = block_sum((double)z, sh);
    if (threadIdx.x == 0) zpart[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = zs;
}

// neg (and pos) = the slices' per-row sums, added in slice order.
template <class T, bool kDense>
__global__ __launch_bounds__(256) void exact_reduce(const T* __restrict__ part_rep, const T* __restrict__ part_att,
                                                    size_t nd, int slices, T* __restrict__ neg, T* __restrict__ pos)
{
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < nd; e += (size_t)gridDim.x * 256) {
        T r = part_rep[e], a = kDense ? part_att[e] : (T)0;
        for (int sl = 1; sl < slices; sl++) {
            r += part_rep[(size_t)sl * nd + e];
            if (kDense) a += part_att[(size_t)sl * nd + e];
        }
        neg[e] = r;
        if (kDense) pos[e] = a;
    }
}

// FP64 all-pairs sum of 1/(1+d^2) (textbook) or 1/(1-2<y_i,y_j>) (compat),
// on the exact_tiles grid: zpart[y * gridDim.x + x] holds the sum over rows
// of block x and the columns of slice y.
template <class T, int DIMS, bool kCompat>
__global__ __launch_bounds__(256) void pair_z(const T* __restrict__ Y, int N, int cols, double* __restrict__ zpart)
{
    __shared__ double sY[256 * DIMS];
    __shared__ double sh[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * cols, c1 = min(N, c0 + cols);
    double yi[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; d++) yi[d] = (i < N) ? (double)Y[(size_t)i * DIMS + d] : 0.0;
    double z = 0.0;
    for (int j0 = c0; j0 < c1; j0 += 256) {
        const int jn = min(256, c1 - j0);
        __syncthreads();
        for (int e = threadIdx.x; e < jn * DIMS; e += 256) sY[e] = (double)Y[(size_t)j0 * DIMS + e];
        __syncthreads();
        if (i < N) {
            for (int jj = 0; jj < jn; jj++) {
                if (j0 + jj == i) continue;
                double D = 0.0;
#pragma unroll
                for (int d = 0; d < DIMS; d++) {
                    if (kCompat) {
                        D += yi[d] * sY[jj * DIMS + d];
                    } else {
                        const double t = yi[d] - sY[jj * DIMS + d];
                        D += t * t;
                    }
                }
                if (kCompat) D = -2.0 * D;
                z += 1.0 / (1.0 + D);
            }
        }
    }
    const double zs = block_sum(z, sh);
    if (threadIdx.x == 0) zpart[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = zs;
}

// mode 0: p log((p + FLT_MIN)/(q/Z + FLT_MIN))   (tsne.hpp:481-482)
// mode 1: p log(p Z (1 + d^2)) for p > 0          (exact normaliser)
template <class T, int DIMS, int kMode>
__global__ __launch_bounds__(256) void kl_edges(const T* __restrict__ Y, const int* __restrict__ row_ptr,
                                                const int* __restrict__ col, const T* __restrict__ val, double exag,
                                                double Z, int N, double* __restrict__ part)
{
    __shared__ double sh[4];
    const int g = threadIdx.x / kG, gl = threadIdx.x % kG;
    const int i = blockIdx.x * kRowsPerBlock + g;
    double c = 0.0;
    if (i < N) {
        for (int e = row_ptr[i] + gl; e < row_ptr[i + 1]; e += kG) {
            const int j = col[e];
            double D = 0.0;
#pragma unroll
            for (int d = 0; d < DIMS; d++) {
                const double t = (double)Y[(size_t)i * DIMS + d] - (double)Y[(size_t)j * DIMS + d];
                D += t * t;
            }
            const double p = exag * (double)val[e];
            if (kMode == 0) {
                const double Q = (1.0 / (1.0 + D)) / Z;
                c += p * log((p + (double)FLT_MIN) / (Q + (double)FLT_MIN));
            } else if (p > 0.0) {
                c += p * log(p * Z * (1.0 + D));
            }
        }
    }
    c = block_sum(c, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = c;
}

// Dense KL, tsne.hpp:435-446: Q_nm = q_nm / sumQ, Q_nn = DBL_MIN / sumQ.
template <class T, int DIMS, bool kCompat>
__global__ __launch_bounds__(256) void kl_dense(const T* __restrict__ Y, const T* __restrict__ P, double exag,
                                                int N, double sumQ, double* __restrict__ part)
{
    __shared__ double sh[4];
    double c = 0.0;
    const size_t total = (size_t)N * N;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / N), m = (int)(e % N);
        double Q;
        if (n == m) {
            Q = DBL_MIN / sumQ;
        } else {
            double D = 0.0;
#pragma unroll
            for (int d = 0; d < DIMS; d++) {
                const double a = (double)Y[(size_t)n * DIMS + d], b = (double)Y[(size_t)m * DIMS + d];
                if (kCompat)
                    D += a * b;
                else
                    D += (a - b) * (a - b);
            }
            if (kCompat) D = -2.0 * D;
            Q = (1.0 / (1.0 + D)) / sumQ;
        }
        const double p = exag * (double)P[e];
        c += p * log((p + 1e-9) / (Q + 1e-9));
    }
    c = block_sum(c, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void sum_fixed(const double* __restrict__ v, int n, double* __restrict__ out)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += v[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------- renumbering
// The renumbering's permutation: the tree's sorted point order, or the
// identity when that tree's sort gave up (its order may repeat or miss
// points, which would leave holes in the inverse).
__global__ __launch_bounds__(256) void rn_checked_perm(const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ sort_fault, int N,
                                                       uint32_t* __restrict__ out)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < N) out[j] = *sort_fault == 0u ? perm[j] : (uint32_t)j;
}

__global__ __launch_bounds__(256) void rn_inverse(const uint32_t* __restrict__ perm, int N, int* __restrict__ inv)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < N) inv[perm[j]] = j;
}

template <class T>
__global__ __launch_bounds__(256) void rn_gather(const uint32_t* __restrict__ perm, int N, int dims,
                                                 const T* __restrict__ src, T* __restrict__ dst)
{
    const size_t total = (size_t)N * dims;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t j = e / dims, d = e % dims;
        dst[e] = src[(size_t)perm[j] * dims + d];
    }
}

__global__ __launch_bounds__(256) void rn_lengths(const uint32_t* __restrict__ perm, const int* __restrict__ row,
                                                  int N, uint32_t* __restrict__ len)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < N) {
        const uint32_t p = perm[j];
        len[j] = (uint32_t)(row[p + 1] - row[p]);
    }
}

template <class T>
__global__ __launch_bounds__(256) void rn_copy_rows(const uint32_t* __restrict__ perm, const int* __restrict__ inv,
                                                    const int* __restrict__ row, const int* __restrict__ col,
                                                    const T* __restrict__ val, const uint32_t* __restrict__ start,
                                                    int N, int* __restrict__ row2, int* __restrict__ col2,
                                                    T* __restrict__ val2)
{
    const int g = threadIdx.x / kG, gl = threadIdx.x % kG;
    const int j = blockIdx.x * kRowsPerBlock + g;
    if (j >= N) return;
    const uint32_t p = perm[j];
    const int b = row[p], n = row[p + 1] - b;
    const uint32_t o = start[j];
    for (int k = gl; k < n; k += kG) {
        col2[o + k] = inv[col[b + k]];
        val2[o + k] = val[b + k];
    }
    if (gl == 0) {
        row2[j] = (int)o;
        if (j == N - 1) row2[N] = (int)start[N];
    }
}

template <class T, int DIMS, bool kCompat, bool kDense>
void launch_exact(const T* Y, int N, const T* P, double exag, T* neg, T* pos, double* zpart, T* work, hipStream_t s)
{
    const ExactGrid g = exact_grid(N);
    const size_t nd = (size_t)N * DIMS;
    T* part_att = kDense ? work + (size_t)g.slices * nd : nullptr;
    hipLaunchKernelGGL((exact_tiles<T, DIMS, kCompat, kDense>), dim3(g.row_blocks, g.slices), dim3(256), 0, s, Y, N,
                       g.cols, P, (T)exag, work, part_att, zpart);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL((exact_reduce<T, kDense>), dim3(std::min<size_t>(2048, ceil_div(nd, 256))), dim3(256), 0, s,
                       work, part_att, nd, g.slices, neg, pos);
    TKA_LAUNCH_CHECK();
}

template <class T, int DIMS>
void dispatch_exact(const T* Y, int N, int dd_mode, const T* P, double exag, T* neg, T* pos, double* zpart, T* work,
                    hipStream_t s)
{
    const bool compat = dd_mode == TKA_DD_TAPKEE_COMPAT;
    if (P) {
        if (compat)
            launch_exact<T, DIMS, true, true>(Y, N, P, exag, neg, pos, zpart, work, s);
        else
            launch_exact<T, DIMS, false, true>(Y, N, P, exag, neg, pos, zpart, work, s);
    } else {
        if (compat)
            launch_exact<T, DIMS, true, false>(Y, N, P, exag, neg, pos, zpart, work, s);
        else
            launch_exact<T, DIMS, false, false>(Y, N, P, exag, neg, pos, zpart, work, s);
    }
}

}  // namespace

int tsne_step_blocks(int N) { return (int)ceil_div(N, kRowsPerBlock); }

size_t renumber_scan_scratch_words(int N)
{
    return (size_t)N + (size_t)N + 1 + exclusive_scan_scratch_words(N) + (size_t)N;
}

template <class T>
void renumber_points_device(int N, int dims, const uint32_t* sorted, const uint32_t* sort_fault, T* const* state,
                            int nstate, T* tmp, const int* row, const int* col, const T* val, int* row2, int* col2,
                            T* val2, int* orig, int* orig_tmp, int* inv, uint32_t* scan_scratch, hipStream_t s)
{
    const unsigned g = ceil_div(N, 256);
    const size_t nd = (size_t)N * dims;
    const unsigned gn = (unsigned)std::min<size_t>(8192, ceil_div(nd, 256));
    uint32_t* perm = scan_scratch + renumber_scan_scratch_words(N) - N;
    hipLaunchKernelGGL(rn_checked_perm, dim3(g), dim3(256), 0, s, sorted, sort_fault, N, perm);
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(rn_inverse, dim3(g), dim3(256), 0, s, perm, N, inv);
    TKA_LAUNCH_CHECK();
    for (int k = 0; k < nstate; k++) {
        hipLaunchKernelGGL((rn_gather<T>), dim3(gn), dim3(256), 0, s, perm, N, dims, state[k], tmp);
        TKA_LAUNCH_CHECK();
        TKA_HIP(hipMemcpyAsync(state[k], tmp, nd * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL((rn_gather<int>), dim3(g), dim3(256), 0, s, perm, N, 1, orig, orig_tmp);
    TKA_LAUNCH_CHECK();
    TKA_HIP(hipMemcpyAsync(orig, orig_tmp, (size_t)N * sizeof(int), hipMemcpyDeviceToDevice, s));
    uint32_t* len = scan_scratch;
    uint32_t* start = scan_scratch + N;
    uint32_t* tmpw = start + N + 1;
    hipLaunchKernelGGL(rn_lengths, dim3(g), dim3(256), 0, s, perm, row, N, len);
    TKA_LAUNCH_CHECK();
    exclusive_scan_u32(len, start, N, tmpw, s);
    hipLaunchKernelGGL((rn_copy_rows<T>), dim3(tsne_step_blocks(N)), dim3(256), 0, s, perm, inv, row, col, val, start,
                       N, row2, col2, val2);
    TKA_LAUNCH_CHECK();
}

template <class T>
__global__ __launch_bounds__(256) void unpermute_rows(int N, int dims, const T* __restrict__ Y,
                                                      const int* __restrict__ orig, T* __restrict__ out)
{
    const size_t total = (size_t)N * dims;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t j = e / dims, d = e % dims;
        out[(size_t)orig[j] * dims + d] = Y[e];
    }
}

template <class T>
void unpermute_rows_device(int N, int dims, const T* Y, const int* orig, T* out, hipStream_t s)
{
    const size_t nd = (size_t)N * dims;
    hipLaunchKernelGGL((unpermute_rows<T>), dim3((unsigned)std::min<size_t>(8192, ceil_div(nd, 256))), dim3(256), 0,
                       s, N, dims, Y, orig, out);
    TKA_LAUNCH_CHECK();
}

template <class T>
void center_device(int dims, T* Y, int N, const double* box_part, int nb, hipStream_t s)
{
    const unsigned g = ceil_div(N, 256);
    switch (dims) {
        case 1: hipLaunchKernelGGL((center_rows<T, 1>), dim3(g), dim3(256), 0, s, Y, N, box_part, nb); break;
        case 2: hipLaunchKernelGGL((center_rows<T, 2>), dim3(g), dim3(256), 0, s, Y, N, box_part, nb); break;
        case 3: hipLaunchKernelGGL((center_rows<T, 3>), dim3(g), dim3(256), 0, s, Y, N, box_part, nb); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
    TKA_LAUNCH_CHECK();
}

template <class T>
void center_fresh_device(int dims, T* Y, int N, MemBudget* mb, hipStream_t s)
{
    const unsigned g = ceil_div(N, 256);
    DevBuf<double> part((size_t)kBoxStride * g, mb);
    switch (dims) {
        case 1: hipLaunchKernelGGL((sum_rows<T, 1>), dim3(g), dim3(256), 0, s, Y, N, part.get()); break;
        case 2: hipLaunchKernelGGL((sum_rows<T, 2>), dim3(g), dim3(256), 0, s, Y, N, part.get()); break;
        case 3: hipLaunchKernelGGL((sum_rows<T, 3>), dim3(g), dim3(256), 0, s, Y, N, part.get()); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
    TKA_LAUNCH_CHECK();
    center_device<T>(dims, Y, N, part.get(), (int)g, s);
    TKA_HIP(hipStreamSynchronize(s));  // part is released on return
}

template <class T>
int tsne_step_device(int dims, const T* Y, T* Ynext, T* uY, T* gains, const T* neg, const T* pos_pre,
                     const int* row_ptr, const int* col, const T* val, const double* zpart, int nz, int N,
                     StepParams sp, double* box_part, T* dC_out, hipStream_t s)
{
    const int nb = tsne_step_blocks(N);
#define TKA_STEP(DM)                                                                                              \
    hipLaunchKernelGGL((tsne_step<T, DM>), dim3(nb), dim3(256), 0, s, Y, Ynext, uY, gains, neg, pos_pre, row_ptr, \
                       col, val, zpart, nz, N, (T)sp.momentum, (T)sp.eta, (T)sp.exag, box_part, dC_out)
    switch (dims) {
        case 1: TKA_STEP(1); break;
        case 2: TKA_STEP(2); break;
        case 3: TKA_STEP(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_STEP
    TKA_LAUNCH_CHECK();
    return nb;
}

int tsne_update_blocks(int N) { return (int)ceil_div(N, kUpdBlock); }

template <class T>
void attraction_device(int dims, const T* Y, const int* row_ptr, const int* col, const T* val, double exag, int N,
                       T* pos, hipStream_t s)
{
    const int nb = tsne_step_blocks(N);  // one row group per kG lanes: full width, alone on the device
#define TKA_ATT(DM) \
    hipLaunchKernelGGL((attract<T, DM>), dim3(nb), dim3(256), 0, s, Y, row_ptr, col, val, (T)exag, N, pos)
    switch (dims) {
        case 1: TKA_ATT(1); break;
        case 2: TKA_ATT(2); break;
        case 3: TKA_ATT(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_ATT
    TKA_LAUNCH_CHECK();
}

template <class T>
int tsne_update_device(int dims, const T* Y, T* Ynext, T* uY, T* gains, const T* neg, const T* pos,
                       const double* zpart, int nz, int N, StepParams sp, double* box_part, T* dC_out, hipStream_t s)
{
    const int nb = tsne_update_blocks(N);
#define TKA_UPD(DM)                                                                                                  \
    hipLaunchKernelGGL((tsne_update<T, DM>), dim3(nb), dim3(kUpdBlock), 0, s, Y, Ynext, uY, gains, neg, pos, zpart, \
                       nz, N, (T)sp.momentum, (T)sp.eta, box_part, dC_out)
    switch (dims) {
        case 1: TKA_UPD(1); break;
        case 2: TKA_UPD(2); break;
        case 3: TKA_UPD(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_UPD
    TKA_LAUNCH_CHECK();
    return nb;
}

template <class T>
void gradient_combine_device(int dims, const T* Y, const T* neg, const T* pos_pre, const int* row_ptr,
                             const int* col, const T* val, double exag, const double* zpart, int nz, int N, T* dC,
                             hipStream_t s)
{
    const int nb = tsne_step_blocks(N);
#define TKA_GC(DM)                                                                                                    \
    hipLaunchKernelGGL((grad_combine<T, DM>), dim3(nb), dim3(256), 0, s, Y, neg, pos_pre, row_ptr, col, val, (T)exag, \
                       zpart, nz, N, dC)
    switch (dims) {
        case 1: TKA_GC(1); break;
        case 2: TKA_GC(2); break;
        case 3: TKA_GC(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_GC
    TKA_LAUNCH_CHECK();
}

template <class T>
int exact_forces_device(int dims, const T* Y, int N, int dd_mode, const T* P, double exag, T* neg, T* pos,
                        double* zpart, T* work, hipStream_t s)
{
    switch (dims) {
        case 1: dispatch_exact<T, 1>(Y, N, dd_mode, P, exag, neg, pos, zpart, work, s); break;
        case 2: dispatch_exact<T, 2>(Y, N, dd_mode, P, exag, neg, pos, zpart, work, s); break;
        case 3: dispatch_exact<T, 3>(Y, N, dd_mode, P, exag, neg, pos, zpart, work, s); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
    return exact_zblocks(N);
}

void sum_into_device(const double* v, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_fixed, dim3(1), dim3(256), 0, s, v, n, out);
    TKA_LAUNCH_CHECK();
}

double sum_device(const double* v, int n, MemBudget* mb, hipStream_t s)
{
    DevBuf<double> out(1, mb);
    hipLaunchKernelGGL(sum_fixed, dim3(1), dim3(256), 0, s, v, n, out.get());
    TKA_LAUNCH_CHECK();
    double h = 0.0;
    read_to_host(&h, out.get(), sizeof h, s);
    return h;
}

template <class T>
double kl_edges_reference_device(int dims, const T* Y, const int* row_ptr, const int* col, const T* val, double exag,
                                 double Z, int N, MemBudget* mb, hipStream_t s)
{
    const int nb = tsne_step_blocks(N);
    DevBuf<double> part(nb, mb);
#define TKA_KL(DM) \
    hipLaunchKernelGGL((kl_edges<T, DM, 0>), dim3(nb), dim3(256), 0, s, Y, row_ptr, col, val, exag, Z, N, part.get())
    switch (dims) {
        case 1: TKA_KL(1); break;
        case 2: TKA_KL(2); break;
        case 3: TKA_KL(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_KL
    TKA_LAUNCH_CHECK();
    return sum_device(part.get(), nb, mb, s);
}

template <class T>
double exact_z_device(int dims, const T* Y, int N, MemBudget* mb, hipStream_t s)
{
    const int nb = exact_zblocks(N);
    const ExactGrid g = exact_grid(N);
    const dim3 grid(g.row_blocks, g.slices);
    DevBuf<double> part(nb, mb);
    switch (dims) {
        case 1: hipLaunchKernelGGL((pair_z<T, 1, false>), grid, dim3(256), 0, s, Y, N, g.cols, part.get()); break;
        case 2: hipLaunchKernelGGL((pair_z<T, 2, false>), grid, dim3(256), 0, s, Y, N, g.cols, part.get()); break;
        case 3: hipLaunchKernelGGL((pair_z<T, 3, false>), grid, dim3(256), 0, s, Y, N, g.cols, part.get()); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
    TKA_LAUNCH_CHECK();
    return sum_device(part.get(), nb, mb, s);
}

template <class T>
double kl_edges_exact_device(int dims, const T* Y, const int* row_ptr, const int* col, const T* val, double exag,
                             double Z, int N, MemBudget* mb, hipStream_t s)
{
    const int nb = tsne_step_blocks(N);
    DevBuf<double> part(nb, mb);
#define TKA_KL(DM) \
    hipLaunchKernelGGL((kl_edges<T, DM, 1>), dim3(nb), dim3(256), 0, s, Y, row_ptr, col, val, exag, Z, N, part.get())
    switch (dims) {
        case 1: TKA_KL(1); break;
        case 2: TKA_KL(2); break;
        case 3: TKA_KL(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_KL
    TKA_LAUNCH_CHECK();
    return sum_device(part.get(), nb, mb, s);
}

template <class T>
double kl_dense_device(int dims, const T* Y, const T* P, double exag, int N, int dd_mode, MemBudget* mb,
                       hipStream_t s)
{
    const bool compat = dd_mode == TKA_DD_TAPKEE_COMPAT;
    const int nbz = exact_zblocks(N);
    const ExactGrid g = exact_grid(N);
    const dim3 grid(g.row_blocks, g.slices);
    DevBuf<double> zp(nbz, mb);
#define TKA_PZ(DM)                                                                                              \
    if (compat)                                                                                                 \
        hipLaunchKernelGGL((pair_z<T, DM, true>), grid, dim3(256), 0, s, Y, N, g.cols, zp.get());               \
    else                                                                                                        \
        hipLaunchKernelGGL((pair_z<T, DM, false>), grid, dim3(256), 0, s, Y, N, g.cols, zp.get());
    switch (dims) {
        case 1: TKA_PZ(1); break;
        case 2: TKA_PZ(2); break;
        case 3: TKA_PZ(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_PZ
    TKA_LAUNCH_CHECK();
    const double sumQ = DBL_MIN + sum_device(zp.get(), nbz, mb, s);
    const int nb = (int)std::min<size_t>(4096, ceil_div((size_t)N * N, 256));
    DevBuf<double> part(nb, mb);
#define TKA_KD(DM)                                                                                                 \
    if (compat)                                                                                                    \
        hipLaunchKernelGGL((kl_dense<T, DM, true>), dim3(nb), dim3(256), 0, s, Y, P, exag, N, sumQ, part.get());  \
    else                                                                                                           \
        hipLaunchKernelGGL((kl_dense<T, DM, false>), dim3(nb), dim3(256), 0, s, Y, P, exag, N, sumQ, part.get());
    switch (dims) {
        case 1: TKA_KD(1); break;
        case 2: TKA_KD(2); break;
        case 3: TKA_KD(3); break;
        default: fail(TKA_E_WRONG_PARAMETER, "no_dims must be 1, 2 or 3");
    }
#undef TKA_KD
    TKA_LAUNCH_CHECK();
    return sum_device(part.get(), nb, mb, s);
}

#define TKA_INST(T)                                                                                                  \
    template void center_device<T>(int, T*, int, const double*, int, hipStream_t);                                \
    template void center_fresh_device<T>(int, T*, int, MemBudget*, hipStream_t);                                  \
    template void renumber_points_device<T>(int, int, const uint32_t*, const uint32_t*, T* const*, int, T*,         \
                                            const int*, const int*, const T*, int*, int*, T*, int*, int*, int*,      \
                                            uint32_t*, hipStream_t);                                                 \
    template void unpermute_rows_device<T>(int, int, const T*, const int*, T*, hipStream_t);                        \
    template int tsne_step_device<T>(int, const T*, T*, T*, T*, const T*, const T*, const int*, const int*, const T*, \
                                     const double*, int, int, StepParams, double*, T*, hipStream_t);                 \
    template void attraction_device<T>(int, const T*, const int*, const int*, const T*, double, int, T*,           \
                                       hipStream_t);                                                                 \
    template int tsne_update_device<T>(int, const T*, T*, T*, T*, const T*, const T*, const double*, int, int,       \
                                       StepParams, double*, T*, hipStream_t);                                        \
    template void gradient_combine_device<T>(int, const T*, const T*, const T*, const int*, const int*, const T*,    \
                                             double, const double*, int, int, T*, hipStream_t);                      \
    template int exact_forces_device<T>(int, const T*, int, int, const T*, double, T*, T*, double*, T*,             \
                                        hipStream_t);                                                                \
    template double kl_edges_reference_device<T>(int, const T*, const int*, const int*, const T*, double, double,    \
                                                 int, MemBudget*, hipStream_t);                                       \
    template double exact_z_device<T>(int, const T*, int, MemBudget*, hipStream_t);                                  \
    template double kl_edges_exact_device<T>(int, const T*, const int*, const int*, const T*, double, double, int,   \
                                             MemBudget*, hipStream_t);                                                \
    template double kl_dense_device<T>(int, const T*, const T*, double, int, int, MemBudget*, hipStream_t);
TKA_INST(float)
TKA_INST(double)
#undef TKA_INST

}  // namespace tka
