// Brute-force kNN for the t-SNE input similarities (replaces tapkee's VP-tree
// search, vptree.hpp:126-293, used at tsne.hpp:628-646).
//
// Pipeline per call:
//   1. prep:   low-precision copies of the normalised rows (FP32, or BF16
//              hi/lo splits laid out for MFMA) plus FP64 row norms.
//   2. tiles:  each wave owns 32 query rows; the block streams all candidate
//              rows through LDS in chunks of 128 and computes 32x32 distance
//              tiles (MFMA v_mfma_f32_32x32x16_bf16, or FP32 FMAs laid out the
//              same way).  Values below the row's running threshold are
//              appended to a per-row buffer of kCap entries in HBM; when a
//              buffer could overflow the wave sorts it (register bitonic) and
//              keeps the k' smallest, lowering the threshold.
//   3. rerank: one wave per row recomputes the kept candidates' squared
//              distances in FP64 (same summation as the reference's
//              euclidean_distance, vptree.hpp:95-101), sorts by
//              (distance, index) and checks an a-priori error bound: the K-th
//              exact distance must lie below (threshold - bound), which proves
//              no dropped candidate can belong to the top K.
//   4. rows failing the bound are re-run with FP32 tiles and k' = 384.
#include "knn.hpp"

#include <vector>

namespace tka {
namespace {

constexpr int kCap = 512;      // candidate slots per row
constexpr int kCB = 128;       // candidates per LDS chunk
constexpr int kQB = 128;       // query rows per block (4 waves x 32)
constexpr int kMaxKprime = kCap - kCB;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------- helpers
TKA_DEV uint32_t f2ord(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
TKA_DEV float ord2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

TKA_DEV uint64_t shfl_xor_u64(uint64_t v, int m)
{
    const uint32_t lo = __shfl_xor((uint32_t)v, m, 64);
    const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of 64*R keys held as v[r] at position r*64 + lane.
template <int R>
TKA_DEV void wave_bitonic_sort(uint64_t (&v)[R])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int jr = j >> 6;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if ((r & jr) == 0) {
                        const bool up = (((r * 64 + lane) & k) == 0);
                        const uint64_t a = v[r], b = v[r | jr];
                        const bool sw = up ? (a > b) : (a < b);
                        v[r] = sw ? b : a;
                        v[r | jr] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int p = r * 64 + lane;
                    const bool up = (p & k) == 0;
                    const bool lower = (p & j) == 0;
                    const uint64_t o = shfl_xor_u64(v[r], j);
                    const uint64_t mn = o < v[r] ? o : v[r];
                    const uint64_t mx = o < v[r] ? v[r] : o;
                    v[r] = (lower == up) ? mn : mx;
                }
            }
        }
    }
}

// Lexicographic (distance bits, index) keys for the FP64 re-rank.
struct DKey {
    uint64_t d;
    uint32_t i;
};
TKA_DEV bool dkey_less(const DKey& a, const DKey& b) { return a.d < b.d || (a.d == b.d && a.i < b.i); }

template <int R>
TKA_DEV void wave_bitonic_sort(DKey (&v)[R])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int jr = j >> 6;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if ((r & jr) == 0) {
                        const bool up = (((r * 64 + lane) & k) == 0);
                        const DKey a = v[r], b = v[r | jr];
                        const bool sw = up ? dkey_less(b, a) : dkey_less(a, b);
                        v[r] = sw ? b : a;
                        v[r | jr] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int p = r * 64 + lane;
                    const bool up = (p & k) == 0;
                    const bool lower = (p & j) == 0;
                    DKey o;
                    o.d = shfl_xor_u64(v[r].d, j);
                    o.i = __shfl_xor(v[r].i, j, 64);
                    const bool oless = dkey_less(o, v[r]);
                    const DKey mn = oless ? o : v[r];
                    const DKey mx = oless ? v[r] : o;
                    v[r] = (lower == up) ? mn : mx;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- normalise
// Column means with the reference's summation order: row 0, 1, 2, ... into a
// zero-initialised FP64 accumulator (tsne.hpp:496-506).  One block per 64
// columns; all waves prefetch the next 64-row chunk into registers while wave
// 0 adds the current chunk from LDS.
__global__ __launch_bounds__(256) void col_mean_seq(const double* __restrict__ X, int N, int D,
                                                    double* __restrict__ mean)
{
    constexpr int R = 64;
    __shared__ double tile[R][65];
    const int c0 = blockIdx.x * 64;
    const int tid = threadIdx.x;
    double regs[R * 64 / 256];
    auto load = [&](int r0) {
#pragma unroll
        for (int e = 0; e < R * 64 / 256; e++) {
            const int idx = e * 256 + tid;
            const int r = idx >> 6, c = idx & 63;
            const int row = r0 + r, col = c0 + c;
            regs[e] = (row < N && col < D) ? X[(size_t)row * D + col] : 0.0;
        }
    };
    double acc = 0.0;
    load(0);
    for (int r0 = 0; r0 < N; r0 += R) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < R * 64 / 256; e++) {
            const int idx = e * 256 + tid;
            tile[idx >> 6][idx & 63] = regs[e];
        }
        __syncthreads();
        if (r0 + R < N) load(r0 + R);
        if (tid < 64) {
            const int rows = (N - r0) < R ? (N - r0) : R;
            for (int r = 0; r < rows; r++) acc = dadd_sep(acc, tile[r][tid]);
        }
    }
    if (tid < 64 && c0 + tid < D) mean[c0 + tid] = acc / (double)N;
}

__global__ __launch_bounds__(256) void center_max(double* __restrict__ X, size_t total, int D,
                                                  const double* __restrict__ mean, double* __restrict__ bmax)
{
    __shared__ double sh[4];
    double m = -HUGE_VAL;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const double v = X[i] - mean[i % D];
        X[i] = v;
        m = v > m ? v : m;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int i = 1; i < 4; i++) r = sh[i] > r ? sh[i] : r;
        bmax[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void reduce_max(const double* __restrict__ bmax, int nb, double* __restrict__ out)
{
    __shared__ double sh[4];
    double m = -HUGE_VAL;
    for (int i = threadIdx.x; i < nb; i += 256) m = bmax[i] > m ? bmax[i] : m;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int i = 1; i < 4; i++) r = sh[i] > r ? sh[i] : r;
        *out = r;
    }
}

__global__ __launch_bounds__(256) void divide_all(double* __restrict__ X, size_t total, const double* __restrict__ mx)
{
    const double m = *mx;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) X[i] /= m;
}

// ---------------------------------------------------------------- prep
// Per row: FP64 squared norm (for the certificate) and its square root.
__global__ __launch_bounds__(256) void row_norms(const double* __restrict__ X, int N, int D, double* __restrict__ n2,
                                                 double* __restrict__ rmax_part)
{
    __shared__ double sh[4];
    const int n = blockIdx.x * 256 + threadIdx.x;
    double r = 0.0;
    if (n < N) {
        double s = 0.0;
        for (int d = 0; d < D; d++) {
            const double x = X[(size_t)n * D + d];
            s = dadd_sep(s, dmul_sep(x, x));
        }
        n2[n] = s;
        r = sqrt(s);
    }
    r = wave_max(r);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int i = 1; i < 4; i++) m = sh[i] > m ? sh[i] : m;
        rmax_part[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(256) void prep_fp32(const double* __restrict__ X, size_t total,
                                                 float* __restrict__ Xf)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256)
        Xf[i] = (float)X[i];
}

// BF16 operand rows of width Kp (multiple of 16):
//   candidates A = [c_hi | c_lo | c_hi | n_hi n_lo | 0...]
//   queries    B = [-2 q_hi | -2 q_hi | -2 q_lo | 1 1 | 0...]
// so that A.B = ||c||^2 - 2 q.c with the q_lo*c_lo term dropped (~2^-16
// relative).  Scaling by -2 is exact in BF16.
__global__ __launch_bounds__(256) void prep_bf16x3(const double* __restrict__ X, const double* __restrict__ n2,
                                                   int N, int D, int Kp, __bf16* __restrict__ A,
                                                   __bf16* __restrict__ B)
{
    const size_t total = (size_t)N * Kp;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / Kp), k = (int)(e % Kp);
        float a = 0.f, b = 0.f;
        if (k < 3 * D) {
            const int seg = k / D, d = k % D;
            const double x = X[(size_t)n * D + d];
            const __bf16 hi = (__bf16)x;
            const __bf16 lo = (__bf16)(x - (double)(float)hi);
            if (seg == 0) {
                a = (float)hi;
                b = -2.f * (float)hi;
            } else if (seg == 1) {
                a = (float)lo;
                b = -2.f * (float)hi;
            } else {
                a = (float)hi;
                b = -2.f * (float)lo;
            }
        } else if (k == 3 * D || k == 3 * D + 1) {
            const double nn = n2[n];
            const __bf16 hi = (__bf16)nn;
            const __bf16 lo = (__bf16)(nn - (double)(float)hi);
            a = (k == 3 * D) ? (float)hi : (float)lo;
            b = 1.f;
        }
        A[e] = (__bf16)a;
        B[e] = (__bf16)b;
    }
}

// ---------------------------------------------------------------- tiles
struct TileArgs {
    int N;
    int kprime;
    const int* qlist;   // optional list of query rows (nq entries)
    int nq;
    uint64_t* buf;      // row-slot x kCap
    int* cnt_out;       // row-slot
    float* tau_out;     // row-slot
};

// Row-slot of the wave's query: blocks cover query slots [0, nq).
TKA_DEV int query_row(const TileArgs& a, int slot) { return slot < a.nq ? (a.qlist ? a.qlist[slot] : slot) : -1; }

// Appends the 16 values a lane holds for its query (tile t of the chunk).
// Lane l holds C[m][n] with n = l%32 (query) and
// m = 8*(i/4) + 4*(l/32) + i%4 (candidate within the 32-wide tile).
TKA_DEV void filter_tile(const float (&v)[16], int t, int cb, int q, int slot_local, float tau, const TileArgs& a,
                         int* s_cnt, uint64_t* qbuf)
{
    const int lane = threadIdx.x & 63;
    // Once the threshold has settled almost no value passes: one min per lane
    // decides whether the per-value tests are needed at all.
    float mn = v[0];
#pragma unroll
    for (int i = 1; i < 16; i++) mn = fminf(mn, v[i]);
    if (!(mn < tau)) return;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int m = 32 * t + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int j = cb + m;
        if (q >= 0 && j < a.N && j != q && v[i] < tau) {
            const int s = atomicAdd(&s_cnt[slot_local], 1);
            qbuf[s] = ((uint64_t)f2ord(v[i]) << 32) | (uint32_t)j;
        }
    }
}

// After a chunk: every row of this wave whose buffer could overflow on the
// next chunk is sorted and cut to kprime entries.
TKA_DEV void compact_if_needed(int w, const TileArgs& a, int slot0, int* s_cnt, float* s_tau)
{
    const int lane = threadIdx.x & 63;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int my = 32 * w + (lane & 31);
    const bool need = lane < 32 && s_cnt[my] > kCap - kCB;
    uint64_t mask = __ballot(need);
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int sl = 32 * w + l;
        const int c = s_cnt[sl];
        uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
        uint64_t v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int p = r * 64 + lane;
            v[r] = p < c ? qbuf[p] : ~0ull;
        }
        wave_bitonic_sort<8>(v);
        const int kp = a.kprime;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int p = r * 64 + lane;
            if (p < kp) qbuf[p] = v[r];
            if (p == kp - 1) s_tau[sl] = ord2f((uint32_t)(v[r] >> 32));
        }
        if (lane == 0) s_cnt[sl] = kp;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// FP32 VALU tiles.  Xf: N x D FP32, nrm: ||c||^2 in FP32.
__global__ __launch_bounds__(256) void knn_tiles_fp32(const float* __restrict__ Xf, const float* __restrict__ nrm,
                                                      int D, TileArgs a)
{
    constexpr int KC = 32;
    __shared__ float sC[KC][kCB + 4];
    __shared__ float sQ[KC][kQB + 4];
    __shared__ float sN[kCB];
    __shared__ int s_cnt[kQB];
    __shared__ float s_tau[kQB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int slot0 = blockIdx.x * kQB;
    const int sl = 32 * w + (lane & 31);
    const int q = query_row(a, slot0 + sl);
    if (tid < kQB) {
        s_cnt[tid] = 0;
        s_tau[tid] = HUGE_VALF;
    }
    uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
    for (int cb = 0; cb < a.N; cb += kCB) {
        float acc[4][16];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[t][i] = 0.f;
        for (int kc = 0; kc < D; kc += KC) {
            const int kn = (D - kc) < KC ? (D - kc) : KC;
            __syncthreads();
            for (int e = tid; e < kCB * KC; e += 256) {
                const int m = e / KC, k = e % KC;
                const int row = cb + m;
                sC[k][m] = (row < a.N && k < kn) ? Xf[(size_t)row * D + kc + k] : 0.f;
                const int qs = m;  // also stage queries (kQB == kCB)
                const int qr = query_row(a, slot0 + qs);
                sQ[k][qs] = (qr >= 0 && k < kn) ? Xf[(size_t)qr * D + kc + k] : 0.f;
            }
            if (kc == 0 && tid < kCB) sN[tid] = (cb + tid < a.N) ? nrm[cb + tid] : 0.f;
            __syncthreads();
            for (int k = 0; k < kn; k++) {
                const float qv = sQ[k][sl];
#pragma unroll
                for (int t = 0; t < 4; t++) {
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        const float4 c4 = *reinterpret_cast<const float4*>(&sC[k][32 * t + 8 * g + 4 * half]);
                        acc[t][4 * g + 0] = fmaf(qv, c4.x, acc[t][4 * g + 0]);
                        acc[t][4 * g + 1] = fmaf(qv, c4.y, acc[t][4 * g + 1]);
                        acc[t][4 * g + 2] = fmaf(qv, c4.z, acc[t][4 * g + 2]);
                        acc[t][4 * g + 3] = fmaf(qv, c4.w, acc[t][4 * g + 3]);
                    }
                }
            }
        }
        const float tau = s_tau[sl];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = 32 * t + 8 * (i >> 2) + 4 * half + (i & 3);
                v[i] = fmaf(-2.f, acc[t][i], sN[m]);
            }
            filter_tile(v, t, cb, q, sl, tau, a, s_cnt, qbuf);
        }
        compact_if_needed(w, a, slot0, s_cnt, s_tau);
    }
    if (lane < 32 && slot0 + sl < a.nq) {
        a.cnt_out[slot0 + sl] = s_cnt[sl];
        a.tau_out[slot0 + sl] = s_tau[sl];
    }
}

// BF16 MFMA tiles.  A/B: N x Kp operand rows (see prep_bf16x3).
__global__ __launch_bounds__(256) void knn_tiles_bf16(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                                      int Kp, TileArgs a)
{
    constexpr int KC = 160;                 // bf16 per LDS chunk (multiple of 16)
    constexpr int KV = KC / 8;              // bf16x8 vectors per row in a chunk
    __shared__ bf16x8 sA[kCB][KV + 1];      // +1 vector of padding per row
    __shared__ int s_cnt[kQB];
    __shared__ float s_tau[kQB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int slot0 = blockIdx.x * kQB;
    const int sl = 32 * w + (lane & 31);
    const int q = query_row(a, slot0 + sl);
    if (tid < kQB) {
        s_cnt[tid] = 0;
        s_tau[tid] = HUGE_VALF;
    }
    uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
    const bf16x8 zero8 = {};
    for (int cb = 0; cb < a.N; cb += kCB) {
        f32x16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = f32x16{};
        for (int kc = 0; kc < Kp; kc += KC) {
            const int kn = (Kp - kc) < KC ? (Kp - kc) : KC;  // multiple of 16
            const int nv = kn / 8;
            __syncthreads();
            for (int e = tid; e < kCB * KV; e += 256) {
                const int m = e / KV, v = e % KV;
                const int row = cb + m;
                sA[m][v] = (row < a.N && v < nv)
                               ? *reinterpret_cast<const bf16x8*>(A + (size_t)row * Kp + kc + 8 * v)
                               : zero8;
            }
            __syncthreads();
            for (int ks = 0; ks < kn / 16; ks++) {
                const bf16x8 b = (q >= 0) ? *reinterpret_cast<const bf16x8*>(B + (size_t)q * Kp + kc + 16 * ks + 8 * half)
                                          : zero8;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const bf16x8 av = sA[32 * t + (lane & 31)][2 * ks + half];
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b, acc[t], 0, 0, 0);
                }
            }
        }
        const float tau = s_tau[sl];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) v[i] = acc[t][i];
            filter_tile(v, t, cb, q, sl, tau, a, s_cnt, qbuf);
        }
        compact_if_needed(w, a, slot0, s_cnt, s_tau);
    }
    if (lane < 32 && slot0 + sl < a.nq) {
        a.cnt_out[slot0 + sl] = s_cnt[sl];
        a.tau_out[slot0 + sl] = s_tau[sl];
    }
}

// FP8 (e4m3) MFMA tiles with per-row absmax scaling: A8 rows hold
// fp8(x / s_x), s_x = absmax(x) / 224, so q.c ~= s_q s_c sum(q8 c8).
// The squared norms come from FP32 copies of the FP64 norms.
__global__ __launch_bounds__(256) void knn_tiles_fp8(const uint8_t* __restrict__ A8, const float* __restrict__ scale,
                                                     const float* __restrict__ nrm, int Kp, TileArgs a)
{
    constexpr int KC = 256;          // bytes (fp8 values) per LDS chunk
    constexpr int KV = KC / 8;       // 8-byte vectors per row in a chunk
    __shared__ uint64_t sA[kCB][KV + 1];
    __shared__ float sN[kCB], sS[kCB];
    __shared__ int s_cnt[kQB];
    __shared__ float s_tau[kQB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int slot0 = blockIdx.x * kQB;
    const int sl = 32 * w + (lane & 31);
    const int q = query_row(a, slot0 + sl);
    const float sq = q >= 0 ? scale[q] : 0.f;
    if (tid < kQB) {
        s_cnt[tid] = 0;
        s_tau[tid] = HUGE_VALF;
    }
    uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
    for (int cb = 0; cb < a.N; cb += kCB) {
        f32x16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = f32x16{};
        for (int kc = 0; kc < Kp; kc += KC) {
            const int kn = (Kp - kc) < KC ? (Kp - kc) : KC;  // multiple of 16
            const int nv = kn / 8;
            __syncthreads();
            for (int e = tid; e < kCB * KV; e += 256) {
                const int m = e / KV, v = e % KV;
                const int row = cb + m;
                sA[m][v] = (row < a.N && v < nv) ? *reinterpret_cast<const uint64_t*>(A8 + (size_t)row * Kp + kc + 8 * v)
                                                 : 0ull;
            }
            if (kc == 0 && tid < kCB) {
                const int row = cb + tid;
                sN[tid] = row < a.N ? nrm[row] : 0.f;
                sS[tid] = row < a.N ? scale[row] : 0.f;
            }
            __syncthreads();
            for (int ks = 0; ks < kn / 16; ks++) {
                const long b = (q >= 0) ? *reinterpret_cast<const long*>(A8 + (size_t)q * Kp + kc + 16 * ks + 8 * half)
                                        : 0l;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const long av = (long)sA[32 * t + (lane & 31)][2 * ks + half];
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(av, b, acc[t], 0, 0, 0);
                }
            }
        }
        const float tau = s_tau[sl];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = 32 * t + 8 * (i >> 2) + 4 * half + (i & 3);
                v[i] = fmaf(-2.f * sq * sS[m], acc[t][i], sN[m]);
            }
            filter_tile(v, t, cb, q, sl, tau, a, s_cnt, qbuf);
        }
        compact_if_needed(w, a, slot0, s_cnt, s_tau);
    }
    if (lane < 32 && slot0 + sl < a.nq) {
        a.cnt_out[slot0 + sl] = s_cnt[sl];
        a.tau_out[slot0 + sl] = s_tau[sl];
    }
}

__global__ __launch_bounds__(256) void prep_fp8(const double* __restrict__ X, const double* __restrict__ n2, int N,
                                                int D, int Kp, uint8_t* __restrict__ A8, float* __restrict__ scale,
                                                float* __restrict__ nrm)
{
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const double* x = X + (size_t)n * D;
    double amax = 0.0;
    for (int d = 0; d < D; d++) amax = fmax(amax, fabs(x[d]));
    const double s = amax > 0.0 ? amax / 224.0 : 1.0;
    scale[n] = (float)s;
    nrm[n] = (float)n2[n];
    uint32_t* row = reinterpret_cast<uint32_t*>(A8 + (size_t)n * Kp);
    for (int d0 = 0; d0 < Kp; d0 += 4) {
        float f[4];
        for (int k = 0; k < 4; k++) f[k] = (d0 + k < D) ? (float)(x[d0 + k] / s) : 0.f;
        int wv = 0;
        wv = __builtin_amdgcn_cvt_pk_fp8_f32(f[0], f[1], wv, false);
        wv = __builtin_amdgcn_cvt_pk_fp8_f32(f[2], f[3], wv, true);
        row[d0 / 4] = (uint32_t)wv;
    }
}

// ---------------------------------------------------------------- rerank
struct RerankArgs {
    const double* X;      // N x D normalised FP64
    const double* n2;     // ||x||^2 FP64
    int N, D, K;
    const int* qlist;     // optional: slot -> row
    int nq;
    const uint64_t* buf;
    const int* cnt;
    const float* tau;
    double eps_a;         // bound = eps_a * ||q|| * Rmax + eps_b * Rmax^2
    double eps_b;
    double rmax;
    int* idx;             // N x K (indexed by row)
    double* dist;
    int* fail;            // per slot: 1 if the bound failed
};

// Exact FP64 squared distance, summed in index order without contraction
// (the reference's euclidean_distance, vptree.hpp:95-101).
TKA_DEV double exact_sqdist(const double* __restrict__ xq, const double* __restrict__ xj, int D)
{
    double dd = 0.0;
    for (int d = 0; d < D; d++) {
        const double t = dadd_sep(xq[d], -xj[d]);
        dd = dadd_sep(dd, dmul_sep(t, t));
    }
    return dd;
}

// Sorts the first R registers' worth of keys (R in {1, 2, 4, 8}).
TKA_DEV void sort_prefix(DKey (&v)[8], int R)
{
    if (R <= 1) {
        DKey u[1] = {v[0]};
        wave_bitonic_sort<1>(u);
        v[0] = u[0];
    } else if (R <= 2) {
        DKey u[2] = {v[0], v[1]};
        wave_bitonic_sort<2>(u);
        v[0] = u[0];
        v[1] = u[1];
    } else if (R <= 4) {
        DKey u[4] = {v[0], v[1], v[2], v[3]};
        wave_bitonic_sort<4>(u);
        for (int r = 0; r < 4; r++) v[r] = u[r];
    } else {
        wave_bitonic_sort<8>(v);
    }
}

// One wave per row.  The kept candidates are sorted by their approximate
// values v_j (|v_j - e_j| <= eps, e_j = ||c_j||^2 - 2 q.c_j); with v_(K) the
// K-th smallest, any candidate with v_j > v_(K) + 2 eps has e_j > e_K and
// cannot be in the top K, so only the prefix up to that threshold gets exact
// FP64 distances.  The row is certified when its K-th exact distance lies
// below (final threshold + ||q||^2 - eps): no candidate dropped during the
// stream can then belong to the top K.
__global__ __launch_bounds__(256) void knn_rerank(RerankArgs a)
{
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (slot >= a.nq) return;
    const int q = a.qlist ? a.qlist[slot] : slot;
    const int c = a.cnt[slot] < kCap ? a.cnt[slot] : kCap;
    const uint64_t* qbuf = a.buf + (size_t)slot * kCap;
    const double* xq = a.X + (size_t)q * a.D;
    const double qn2 = a.n2[q];
    const double eps = a.eps_a * sqrt(qn2) * a.rmax + a.eps_b * a.rmax * a.rmax + 1e-12 * (qn2 + a.rmax * a.rmax);
    const int K = a.K;

    // approximate order
    uint64_t u[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        u[r] = p < c ? qbuf[p] : ~0ull;
    }
    wave_bitonic_sort<8>(u);
    uint32_t vk_bits = 0;
#pragma unroll
    for (int r = 0; r < 8; r++)
        if (r * 64 + lane == K - 1) vk_bits = (uint32_t)(u[r] >> 32);
    vk_bits = __shfl(vk_bits, (K - 1) & 63, 64);
    const double thr = (double)ord2f(vk_bits) + 2.0 * eps;
    int m = 0;  // candidates within the threshold (a prefix of the sorted list)
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        const bool in = p < c && (double)ord2f((uint32_t)(u[r] >> 32)) <= thr;
        m += __popcll(__ballot(in));
    }
    if (c < K) m = c;

    DKey v[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        if (p < m) {
            const uint32_t j = (uint32_t)u[r];
            v[r].d = (uint64_t)__double_as_longlong(exact_sqdist(xq, a.X + (size_t)j * a.D, a.D));
            v[r].i = j;
        } else {
            v[r].d = ~0ull;
            v[r].i = 0xffffffffu;
        }
    }
    const int R = m <= 64 ? 1 : m <= 128 ? 2 : m <= 256 ? 4 : 8;
    sort_prefix(v, R);
    double dk = 0.0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        if (p < K) {
            a.idx[(size_t)q * K + p] = (int)v[r].i;
            a.dist[(size_t)q * K + p] = __longlong_as_double((long long)v[r].d);
        }
        if (p == K - 1) dk = __longlong_as_double((long long)v[r].d);
    }
    dk = __shfl(dk, (K - 1) & 63, 64);
    if (lane == 0) {
        const float tau = a.tau[slot];
        int bad = 0;
        if (c < K) {
            bad = 1;  // cannot happen with K <= N-1 unless values were non-finite
        } else if (isfinite(tau)) {
            bad = !(dk < (double)tau + qn2 - eps);
        }
        a.fail[slot] = bad;
    }
}

__global__ void collect_failures(const int* __restrict__ fail, int n, int* __restrict__ list, int* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && fail[i]) list[atomicAdd(count, 1)] = i;
}

}  // namespace

// ============================================================== host side
void normalize_device(double* X, int N, int D, MemBudget* mb, hipStream_t s)
{
    DevBuf<double> mean(D, mb), bmax(1024, mb), mx(1, mb);
    hipLaunchKernelGGL(col_mean_seq, dim3(ceil_div(D, 64)), dim3(256), 0, s, X, N, D, mean.get());
    TKA_LAUNCH_CHECK();
    const size_t total = (size_t)N * D;
    const int nb = (int)std::min<size_t>(1024, ceil_div(total, 256));
    hipLaunchKernelGGL(center_max, dim3(nb), dim3(256), 0, s, X, total, D, mean.get(), bmax.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_max, dim3(1), dim3(256), 0, s, bmax.get(), nb, mx.get());
    TKA_LAUNCH_CHECK();
    double h = 0.0;
    TKA_HIP(hipMemcpyAsync(&h, mx.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
    require(h > 0.0 && std::isfinite(h), TKA_E_WRONG_PARAMETER,
            "input has no spread: the largest centred coefficient is " + std::to_string(h) +
                " (the reference would divide by it and return NaN)");
    hipLaunchKernelGGL(divide_all, dim3(nb), dim3(256), 0, s, X, total, mx.get());
    TKA_LAUNCH_CHECK();
}

namespace {

struct TileRun {
    DevBuf<uint64_t> buf;
    DevBuf<int> cnt;
    DevBuf<float> tau;
    DevBuf<int> fail;
};

void run_tiles(int dtype, const void* op_a, const void* op_b, const float* nrm, int N, int D, int Kp, int kprime,
               const int* qlist, int nq, TileRun& tr, MemBudget* mb, hipStream_t s)
{
    const int nslots = (int)ceil_div(nq, kQB) * kQB;
    tr.buf.alloc((size_t)nslots * kCap, mb);
    tr.cnt.alloc(nslots, mb);
    tr.tau.alloc(nslots, mb);
    tr.fail.alloc(nslots, mb);
    TileArgs ta{N, kprime, qlist, nq, tr.buf.get(), tr.cnt.get(), tr.tau.get()};
    const dim3 grid(ceil_div(nq, kQB));
    if (dtype == TKA_DIST_BF16) {
        hipLaunchKernelGGL(knn_tiles_bf16, grid, dim3(256), 0, s, (const __bf16*)op_a, (const __bf16*)op_b, Kp, ta);
    } else if (dtype == TKA_DIST_FP8) {
        hipLaunchKernelGGL(knn_tiles_fp8, grid, dim3(256), 0, s, (const uint8_t*)op_a, (const float*)op_b, nrm, Kp,
                           ta);
    } else {
        hipLaunchKernelGGL(knn_tiles_fp32, grid, dim3(256), 0, s, (const float*)op_a, nrm, D, ta);
    }
    TKA_LAUNCH_CHECK();
}

}  // namespace

void knn_device(const double* Xn, int N, int D, int K, int dtype, int kprime, int* idx, double* dist, KnnStats* st,
                MemBudget* mb, hipStream_t s)
{
    require(N >= 2 && K >= 1 && K <= N - 1, TKA_E_WRONG_PARAMETER, "kNN needs 1 <= K <= N-1");
    require(K <= kMaxKprime, TKA_E_WRONG_PARAMETER,
            "K = " + std::to_string(K) + " exceeds the engine's limit of " + std::to_string(kMaxKprime) +
                " neighbours (perplexity <= 128)");
    require(dtype == TKA_DIST_FP32 || dtype == TKA_DIST_BF16 || dtype == TKA_DIST_FP8, TKA_E_UNSUPPORTED,
            "distance dtype " + std::to_string(dtype) + " is not available");
    if (kprime <= 0) kprime = 2 * K;
    kprime = std::max(kprime, K);
    kprime = std::min(kprime, kMaxKprime);
    kprime = std::min(kprime, N - 1);

    // FP64 norms and the largest row norm (for the error bound).
    const int nbr = (int)ceil_div(N, 256);
    DevBuf<double> n2(N, mb), rpart(nbr, mb), rmaxd(1, mb);
    hipLaunchKernelGGL(row_norms, dim3(nbr), dim3(256), 0, s, Xn, N, D, n2.get(), rpart.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_max, dim3(1), dim3(256), 0, s, rpart.get(), nbr, rmaxd.get());
    TKA_LAUNCH_CHECK();
    double rmax = 0.0;
    TKA_HIP(hipMemcpyAsync(&rmax, rmaxd.get(), sizeof(double), hipMemcpyDeviceToHost, s));

    // Operands.
    DevBuf<float> Xf, nrm;
    DevBuf<__bf16> A, B;
    DevBuf<uint8_t> A8;
    DevBuf<float> scale8, nrm8;
    int Kp = 0;
    auto make_fp32 = [&]() {
        if (Xf.get()) return;
        Xf.alloc((size_t)N * D, mb);
        nrm.alloc(N, mb);
        const size_t total = (size_t)N * D;
        hipLaunchKernelGGL(prep_fp32, dim3(std::min<size_t>(4096, ceil_div(total, 256))), dim3(256), 0, s, Xn, total,
                           Xf.get());
        TKA_LAUNCH_CHECK();
        hipLaunchKernelGGL(prep_fp32, dim3(ceil_div(N, 256)), dim3(256), 0, s, n2.get(), (size_t)N, nrm.get());
        TKA_LAUNCH_CHECK();
    };
    const double u32 = std::ldexp(1.0, -24);
    double eps_a = 0.0, eps_b = 0.0;
    auto fp32_bound = [&]() {
        eps_a = 2.0 * (2.0 * (D + 3) * u32);
        eps_b = 2.0 * (2.0 * u32);
    };
    TileRun tr;
    if (dtype == TKA_DIST_BF16) {
        Kp = ((3 * D + 2) + 15) / 16 * 16;
        A.alloc((size_t)N * Kp, mb);
        B.alloc((size_t)N * Kp, mb);
        hipLaunchKernelGGL(prep_bf16x3, dim3(std::min<size_t>(8192, ceil_div((size_t)N * Kp, 256))), dim3(256), 0, s,
                           Xn, n2.get(), N, D, Kp, A.get(), B.get());
        TKA_LAUNCH_CHECK();
        run_tiles(dtype, A.get(), B.get(), nullptr, N, D, Kp, kprime, nullptr, N, tr, mb, s);
        const double ub2 = std::ldexp(1.0, -16);
        const double gam = Kp * u32;
        eps_a = 2.0 * (6.2 * ub2 + 2.1 * gam);
        eps_b = 2.0 * (1.1 * ub2 + 1.1 * gam);
    } else if (dtype == TKA_DIST_FP8) {
        Kp = (D + 15) / 16 * 16;
        A8.alloc((size_t)N * Kp, mb);
        scale8.alloc(N, mb);
        nrm8.alloc(N, mb);
        hipLaunchKernelGGL(prep_fp8, dim3(ceil_div(N, 256)), dim3(256), 0, s, Xn, n2.get(), N, D, Kp, A8.get(),
                           scale8.get(), nrm8.get());
        TKA_LAUNCH_CHECK();
        run_tiles(dtype, A8.get(), scale8.get(), nrm8.get(), N, D, Kp, kprime, nullptr, N, tr, mb, s);
        // e4m3 keeps 4 significant bits (unit roundoff 2^-4): the bound is
        // loose at any useful k'; FP8 rows are not re-solved (approximate mode)
        const double u8 = std::ldexp(1.0, -4);
        eps_a = 2.0 * (2.0 * (2 * u8 + u8 * u8) + 2.0 * D * u32);
        eps_b = 2.0 * (2.0 * u32);
    } else {
        make_fp32();
        run_tiles(dtype, Xf.get(), nullptr, nrm.get(), N, D, 0, kprime, nullptr, N, tr, mb, s);
        fp32_bound();
    }
    TKA_HIP(hipStreamSynchronize(s));  // rmax is on the host now
    RerankArgs ra{Xn, n2.get(), N, D, K, nullptr, N, tr.buf.get(), tr.cnt.get(), tr.tau.get(),
                  eps_a, eps_b, rmax, idx, dist, tr.fail.get()};
    hipLaunchKernelGGL(knn_rerank, dim3(ceil_div(N, 4)), dim3(256), 0, s, ra);
    TKA_LAUNCH_CHECK();

    // Rows whose bound failed: FP32 tiles with the widest candidate list.
    DevBuf<int> list(N, mb), nfail(1, mb);
    TKA_HIP(hipMemsetAsync(nfail.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(collect_failures, dim3(ceil_div(N, 256)), dim3(256), 0, s, tr.fail.get(), N, list.get(),
                       nfail.get());
    TKA_LAUNCH_CHECK();
    int hf = 0;
    TKA_HIP(hipMemcpyAsync(&hf, nfail.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
    if (st) st->certificate_failures = hf;
    if (st) st->unresolved = hf;
    if (hf > 0 && dtype != TKA_DIST_FP8) {
        // Deterministic order of the re-solved rows does not matter: each row
        // is solved independently and written to its own output slot.
        make_fp32();
        fp32_bound();
        TileRun tr2;
        const int kp2 = std::min(kMaxKprime, N - 1);
        run_tiles(TKA_DIST_FP32, Xf.get(), nullptr, nrm.get(), N, D, 0, kp2, list.get(), hf, tr2, mb, s);
        RerankArgs rb{Xn, n2.get(), N, D, K, list.get(), hf, tr2.buf.get(), tr2.cnt.get(), tr2.tau.get(),
                      eps_a, eps_b, rmax, idx, dist, tr2.fail.get()};
        hipLaunchKernelGGL(knn_rerank, dim3(ceil_div(hf, 4)), dim3(256), 0, s, rb);
        TKA_LAUNCH_CHECK();
        TKA_HIP(hipMemsetAsync(nfail.get(), 0, sizeof(int), s));
        hipLaunchKernelGGL(collect_failures, dim3(ceil_div(hf, 256)), dim3(256), 0, s, tr2.fail.get(), hf, list.get(),
                           nfail.get());
        TKA_LAUNCH_CHECK();
        int hf2 = 0;
        TKA_HIP(hipMemcpyAsync(&hf2, nfail.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        TKA_HIP(hipStreamSynchronize(s));
        if (st) st->unresolved = hf2;
    }
}

}  // namespace tka
