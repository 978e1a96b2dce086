// Brute-force kNN for the t-SNE input similarities (replaces tapkee's VP-tree
// search, vptree.hpp:126-293, used at tsne.hpp:628-646).
//
// Pipeline per call:
//   1. prep:   rows put in a locality order (Morton order of a projection on
//              their estimated top-3 principal subspace), then low-precision
//              operands laid out for MFMA (FP32; BF16 centred on each query
//              block's mean, or hi/lo splits for rows too wide for that; FP8),
//              plus FP64 row norms.
//   2. tiles:  each wave owns 32 query rows; the block streams all candidate
//              rows through LDS in chunks of 128 and computes 32x32 distance
//              tiles on the matrix cores (register-resident query operands,
//              or both operands streamed in k-stages for wide rows).  Values
//              below the row's running threshold are appended to a per-row
//              buffer of kCap entries in HBM; when a buffer could overflow the
//              wave keeps the k' smallest (a selection by ballot bisection),
//              lowering the threshold.
//   3. rerank: one wave per row recomputes the kept candidates' squared
//              distances in FP64 (same summation as the reference's
//              euclidean_distance, vptree.hpp:95-101), sorts by
//              (distance, index) and checks an a-priori error bound: the K-th
//              exact distance must lie below the smallest exact distance a
//              value at the threshold can stand for, which proves no dropped
//              candidate can belong to the top K.
//   4. rows failing the bound get every distance in FP64 and an exact
//      selection among those within the K-th distance already found
//      (knn_exact_rows), as do all rows when K exceeds the buffers.
#include "knn.hpp"
#include "radix_sort.hpp"

#include <climits>
#include <cmath>
#include <type_traits>
#include <vector>

namespace tka {
namespace {

constexpr int kCap = 512;      // candidate slots per row
constexpr int kCB = 128;       // candidates per LDS chunk
constexpr int kSeedChunks = 4; // chunks a block streams first, stored whole (seed_store)
static_assert(kSeedChunks * kCB <= kCap, "the seed chunks must fit the row buffer");
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

// Ascending wave bitonic sort of 64*R (distance bits, index) pairs held as
// d[r], ix[r] at position r*64 + lane (the first R of 8 registers; the pair
// order is lexicographic).  Two plain register arrays rather than an array
// of structs: the compiler keeps these in registers.
TKA_DEV bool pair_less(uint64_t ad, uint32_t ai, uint64_t bd, uint32_t bi) { return ad < bd || (ad == bd && ai < bi); }

template <int R>
TKA_DEV void wave_sort_pairs(uint64_t (&d)[8], uint32_t (&ix)[8])
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
                        const int r2 = r | jr;
                        const bool up = (((r * 64 + lane) & k) == 0);
                        const bool sw = up ? pair_less(d[r2], ix[r2], d[r], ix[r]) : pair_less(d[r], ix[r], d[r2], ix[r2]);
                        const uint64_t d0 = d[r], d1 = d[r2];
                        const uint32_t i0 = ix[r], i1 = ix[r2];
                        d[r] = sw ? d1 : d0;
                        d[r2] = sw ? d0 : d1;
                        ix[r] = sw ? i1 : i0;
                        ix[r2] = sw ? i0 : i1;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int p = r * 64 + lane;
                    const bool up = (p & k) == 0;
                    const bool lower = (p & j) == 0;
                    const uint64_t od = shfl_xor_u64(d[r], j);
                    const uint32_t oi = __shfl_xor(ix[r], j, 64);
                    const bool oless = pair_less(od, oi, d[r], ix[r]);
                    const bool take = (lower == up) ? oless : !oless;  // keep the min in the lower slot when ascending
                    d[r] = take ? od : d[r];
                    ix[r] = take ? oi : ix[r];
                }
            }
        }
    }
}

// ---------------------------------------------------------------- normalise
// Column means with the reference's summation order: row 0, 1, 2, ... into a
// zero-initialised FP64 accumulator, no contraction (tsne.hpp:496-506).  Each
// column's sum is one chain of N dependent adds, so its time is N times the
// latency of one FP64 add.  One wave takes one column: its lanes load
// kColGroups groups of 64 consecutive rows one batch ahead (strided loads of
// that column; the columns' waves share the cache lines) and store the batch
// to LDS in row order; every lane then runs the same chain over the batch
// from broadcast LDS reads (no cross-lane moves inside the chain).
constexpr int kColGroups = 8;
constexpr int kColBatch = 64 * kColGroups;

__global__ __launch_bounds__(64) void col_mean_seq(const double* __restrict__ X, int N, int D,
                                                   double* __restrict__ mean)
{
    __shared__ __attribute__((aligned(16))) double s_rows[2][kColBatch];
    const int d = blockIdx.x, lane = threadIdx.x;
    double buf[kColGroups];
#pragma unroll
    for (int g = 0; g < kColGroups; g++) {
        const int row = g * 64 + lane;
        buf[g] = row < N ? X[(size_t)row * D + d] : 0.0;
    }
    double acc = 0.0;
    int b = 0;
    for (int r0 = 0; r0 < N; r0 += kColBatch, b ^= 1) {
        double* sb = s_rows[b];
#pragma unroll
        for (int g = 0; g < kColGroups; g++) sb[g * 64 + lane] = buf[g];
#pragma unroll
        for (int g = 0; g < kColGroups; g++) {  // the next batch, in flight under this one's chain
            const int next = r0 + kColBatch + g * 64 + lane;
            buf[g] = next < N ? X[(size_t)next * D + d] : 0.0;
        }
        __syncthreads();  // one wave: orders the stores before the broadcast reads
        const int cnt = N - r0 < kColBatch ? N - r0 : kColBatch;
        if (cnt == kColBatch) {
            // software-pipelined: the broadcast reads of the next 32 values
            // are issued before the chain over the current 32
            const double2* s2 = reinterpret_cast<const double2*>(sb);
            constexpr int kP = 16;
            double2 cur[kP], nxt[kP];
#pragma unroll
            for (int k = 0; k < kP; k++) cur[k] = s2[k];
#pragma unroll
            for (int t0 = 0; t0 < kColBatch / 2; t0 += kP) {
                if (t0 + kP < kColBatch / 2) {
#pragma unroll
                    for (int k = 0; k < kP; k++) nxt[k] = s2[t0 + kP + k];
                }
#pragma unroll
                for (int k = 0; k < kP; k++) {
                    acc = dadd_sep(acc, cur[k].x);
                    acc = dadd_sep(acc, cur[k].y);
                }
#pragma unroll
                for (int k = 0; k < kP; k++) cur[k] = nxt[k];
            }
        } else {
            for (int t = 0; t < cnt; t++) acc = dadd_sep(acc, sb[t]);
        }
    }
    if (lane == 0) mean[d] = acc / (double)N;
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

// ---------------------------------------------------------------- tiles
struct TileArgs {
    int N;
    int kprime;
    int rotate;         // 1: block b streams candidate chunks b, b+1, ... (mod), starting with its own rows
    const int* qlist;   // optional list of query rows (nq entries)
    int nq;
    uint64_t* buf;      // row-slot x kCap
    int* cnt_out;       // row-slot
    float* tau_out;     // row-slot
};

// Row-slot of the wave's query: blocks cover query slots [0, nq).
TKA_DEV int query_row(const TileArgs& a, int slot) { return slot < a.nq ? (a.qlist ? a.qlist[slot] : slot) : -1; }

TKA_DEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Threshold selection over the ordered keys (high words) of up to 64*R
// entries u[r] at position r*64 + lane (empty entries are ~0): the smallest
// T, to 8*rounds leading bits (the remaining low bits of T set), with at
// least kth entries whose key is <= T; all ones if there are fewer than kth
// entries.  With rounds = 4, T is exactly the kth smallest key.  Bisection
// from the top bit: each step counts the keys at or below a probe with one
// compare and ballot per register.  (Candidate distances share their
// leading bits, so a digit histogram in LDS would serialise its atomics on
// one or two bins.)
template <int R>
TKA_DEV uint32_t wave_select(const uint64_t (&u)[R], int kth, int rounds)
{
    uint64_t valid[R];
    int total = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        valid[r] = __ballot(u[r] != ~0ull);
        total += __popcll(valid[r]);
    }
    if (total < kth) return 0xffffffffu;
    const int low = 32 - 8 * rounds;
    uint32_t res = 0;
    for (int b = 31; b >= low; b--) {
        const uint32_t probe = res | ((1u << b) - 1u);  // bit b clear, every lower bit set
        int c = 0;
#pragma unroll
        for (int r = 0; r < R; r++) c += __popcll(__ballot((uint32_t)(u[r] >> 32) <= probe) & valid[r]);
        if (c < kth) res |= 1u << b;
    }
    return low == 0 ? res : (res | ((1u << low) - 1u));
}

// v[i] for a per-lane index i, by a tree of bit-mask merges (register arrays
// indexed by a divergent value, or selects the compiler folds into such an
// index, go to scratch memory).
TKA_DEV uint32_t merge_bit(uint32_t lo, uint32_t hi, int i, int b)
{
    const uint32_t m = 0u - (uint32_t)((i >> b) & 1);
    return (lo & ~m) | (hi & m);
}
TKA_DEV float pick16(const float (&v)[16], int i)
{
    uint32_t a[8], b[4], c[2];
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = merge_bit(__float_as_uint(v[2 * k]), __float_as_uint(v[2 * k + 1]), i, 0);
#pragma unroll
    for (int k = 0; k < 4; k++) b[k] = merge_bit(a[2 * k], a[2 * k + 1], i, 1);
#pragma unroll
    for (int k = 0; k < 2; k++) c[k] = merge_bit(b[2 * k], b[2 * k + 1], i, 2);
    return __uint_as_float(merge_bit(c[0], c[1], i, 3));
}

// Filter of the 16 values a lane holds for its query (tile t of the chunk):
// lane l holds C[m][n] with n = l%32 (query) and m = 8*(i/4) + 4*(l/32) + i%4
// (candidate within the 32-wide tile).  Values below the row's threshold are
// appended to the row buffer: one min and one compare per lane when nothing
// passes; otherwise one LDS atomic reserves the lane's slots and a loop over
// the lane's passing values stores them (a uniform loop over all 16 values
// with predicated stores measured slower).
TKA_DEV void filter_append(const float (&v)[16], int t, int cb, int q, int slot_local, float tau, const TileArgs& a,
                           int* s_cnt, uint64_t* qbuf);

TKA_DEV void filter_tile(const float (&v)[16], int t, int cb, int q, int slot_local, float tau, const TileArgs& a,
                         int* s_cnt, uint64_t* qbuf)
{
    float mn = v[0];
#pragma unroll
    for (int i = 1; i < 16; i++) mn = fminf(mn, v[i]);
    if (!(mn < tau) || q < 0) return;
    filter_append(v, t, cb, q, slot_local, tau, a, s_cnt, qbuf);
}

// The appends of filter_tile, for a tile known to hold a value below tau.
TKA_DEV void filter_append(const float (&v)[16], int t, int cb, int q, int slot_local, float tau, const TileArgs& a,
                           int* s_cnt, uint64_t* qbuf)
{
    const int half = (threadIdx.x & 63) >> 5;
    const int base = cb + 32 * t + 4 * half;  // candidate of value i: base + 8*(i/4) + i%4
    const bool edge = base + 27 >= a.N || (q >= base && q <= base + 27);
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int j = base + 8 * (i >> 2) + (i & 3);
        const bool ok = v[i] < tau && !(edge && (j >= a.N || j == q));
        m |= (uint32_t)ok << i;
    }
    if (m == 0) return;
    int s = atomicAdd(&s_cnt[slot_local], __popc(m));
    while (m) {
        const int i = __ffs(m) - 1;
        m &= m - 1;
        const int j = base + 8 * (i >> 2) + (i & 3);
        qbuf[s++] = ((uint64_t)f2ord(pick16(v, i)) << 32) | (uint32_t)j;
    }
}

// Seed chunks (the first ones a block streams, while the row threshold is
// still unset): every value is kept, so the lane's 16 values of tile t are
// stored at their fixed places in the row buffer -- chunk c of the seed
// phase fills slots c*kCB .. c*kCB + kCB-1 in candidate order -- with 32-byte
// vector stores instead of appends.  The query itself and rows past N are
// stored as empty entries (~0), which the compaction and the re-rank skip.
TKA_DEV void seed_store(const float (&v)[16], int t, int cb, int q, int c, const TileArgs& a, uint64_t* qbuf)
{
    const int half = (threadIdx.x & 63) >> 5;
    const int off = 32 * t + 4 * half;  // candidate of value i: cb + off + 8*(i/4) + i%4
#pragma unroll
    for (int g = 0; g < 4; g++) {
        uint64_t e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = cb + off + 8 * g + k;
            e[k] = (j < a.N && j != q) ? (((uint64_t)f2ord(v[4 * g + k]) << 32) | (uint32_t)j) : ~0ull;
        }
        ulonglong2* dst = reinterpret_cast<ulonglong2*>(qbuf + c * kCB + off + 8 * g);
        dst[0] = make_ulonglong2(e[0], e[1]);
        dst[1] = make_ulonglong2(e[2], e[3]);
    }
}

// After a chunk: every row of this wave whose buffer could overflow on the
// next chunk keeps only its entries up to a selected threshold (>= kprime of
// them, 16-bit resolution), which becomes the row's new threshold.  If that
// leaves too many, the threshold T is the exact kprime-th key: entries below
// it are kept, and entries equal to it only up to the capacity, the new
// threshold being T itself (a dropped entry then equals the threshold, which
// the re-rank's strict certificate treats as possibly closer: such a row is
// re-solved if it matters).
// Wave w owns the ROWS row-slots ROWS*w .. ROWS*w + ROWS-1 of the block
// (ROWS = 32 or 64: one or two 32-query tiles).
template <int ROWS = 32>
TKA_DEV void compact_if_needed(int w, const TileArgs& a, int slot0, int* s_cnt, float* s_tau)
{
    const int lane = threadIdx.x & 63;
    wave_lds_sync();
    const int my = ROWS * w + (lane % ROWS);
    const bool need = lane < ROWS && s_cnt[my] > kCap - kCB;
    uint64_t mask = __ballot(need);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int sl = ROWS * w + l;
        const int c = s_cnt[sl];
        uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
        uint64_t v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int p = r * 64 + lane;
            v[r] = p < c ? qbuf[p] : ~0ull;
        }
        const int kp = a.kprime;
        uint32_t T = wave_select<8>(v, kp, 2);
        int kept = 0;
#pragma unroll
        for (int r = 0; r < 8; r++) kept += __popcll(__ballot(v[r] != ~0ull && (uint32_t)(v[r] >> 32) <= T));
        bool exact = false;
        if (kept > kCap - kCB) {
            T = wave_select<8>(v, kp, 4);
            exact = true;
        }
        // keep key < T, and key == T (all of them, or with an exact T up to
        // the capacity left), in buffer order
        int pos = 0;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t key = (uint32_t)(v[r] >> 32);
            const bool keep = v[r] != ~0ull && key < T;
            const uint64_t km = __ballot(keep);
            if (keep) qbuf[pos + __popcll(km & lt)] = v[r];
            pos += __popcll(km);
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t key = (uint32_t)(v[r] >> 32);
            const bool tie = v[r] != ~0ull && key == T;
            const uint64_t tm = __ballot(tie);
            const int p = pos + __popcll(tm & lt);
            if (tie && p < kCap - kCB) qbuf[p] = v[r];
            pos += __popcll(tm);
        }
        if (pos > kCap - kCB) pos = kCap - kCB;
        if (lane == 0) {
            s_cnt[sl] = pos;
            s_tau[sl] = exact ? ord2f(T) : ((T >= 0xff7fffffu) ? HUGE_VALF : ord2f(T + 1u));
        }
        wave_lds_sync();
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

// ---------------------------------------------------------------- register tiles
// Operand rows in half-major order for the MFMA layouts: a row is 2*NV
// vectors plus one pad vector, and vector h*NV + g holds the k values that
// lane half h feeds to step group g.  The pad carries the FP32 norm ||c||^2
// and (FP8) the row scale.  Rows are allocated for a whole number of kCB-row
// chunks (zero rows past N).  A wave keeps its query operand in registers
// for the whole candidate stream; candidate chunks are copied global -> LDS
// by the direct LDS loads of CDNA (no VGPR staging), double-buffered, the
// copy of chunk c+1 running while chunk c is multiplied and filtered.  The
// odd row stride (2*NV + 1 vectors) spreads consecutive rows over the LDS
// banks.
constexpr int kNVLimit = 20;       // widest row (vectors per half) kept in registers

// The register tiles' double-buffered chunks plus the kernel's own LDS (row
// counters and thresholds) must fit the 160 KiB of a CU.
inline bool reg_tiles_fit(int NV, size_t vec_bytes)
{
    if (NV > kNVLimit) return false;
    const size_t dyn = 2 * (size_t)kCB * (2 * NV + 1) * vec_bytes;
    return dyn + 256 * 8 <= 160 * 1024;
}

struct OpBF16 {  // v_mfma_f32_32x32x16_bf16: one k-step of 16 per vector pair
    typedef bf16x8 V;
    TKA_DEV static f32x16 mma(V a, V b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};
struct OpFP8 {  // v_mfma_f32_32x32x16_fp8_fp8
    typedef long V;
    TKA_DEV static f32x16 mma(V a, V b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, b, c, 0, 0, 0); }
};
struct OpF32 {  // four k-steps of v_mfma_f32_32x32x2_f32 per vector
    typedef float4 V;
    TKA_DEV static f32x16 mma(V a, V b, f32x16 c)
    {
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.x, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.y, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b.z, c, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b.w, c, 0, 0, 0);
    }
};

// Order in which a block streams the nch candidate chunks.  With rows in
// locality order, chunk s holds the block's own queries: the block first
// takes the seeds s-1, s, s+1, s+2 (its nearest rows, so the first
// compaction already sets tight thresholds), then every other chunk in
// ascending order.  That second part is the same for all blocks, so the
// blocks running together on an XCD read the same chunks at about the same
// time and the XCD's L2 serves them; streams that each started at their own
// chunk would all miss.  s < 0: plain ascending order.
struct ChunkOrder {
    int nch, nseed, a;  // seeds: a, a+1, ..., a+nseed-1 (mod nch)
    TKA_DEV ChunkOrder(int s, int n) : nch(n), nseed(s < 0 ? 0 : min(kSeedChunks, n)), a(s < 0 ? 0 : (s - 1 + n) % n) {}
    // c-th chunk of the stream, 0 <= c < nch
    TKA_DEV int at(int c) const
    {
        if (c < nseed) return (a + c) % nch;
        const int k = c - nseed;            // k-th chunk that is not a seed, ascending
        const int wrap = a + nseed - nch;   // seeds that wrapped to 0, 1, ...
        if (wrap > 0) return wrap + k;      // non-seeds are wrap .. a-1
        return k < a ? k : k + nseed;
    }
};

// The first chunk of a block's stream: the chunk holding its first query row
// (rows are in locality order, so the rows near a query come early).
TKA_DEV int seed_chunk(const TileArgs& a, int slot0, int nch)
{
    if (!a.rotate) return -1;
    const int r = query_row(a, slot0);
    return (r >= 0 ? r / kCB : 0) % nch;
}

typedef __attribute__((address_space(3))) void lds_void;

// Copies `bytes` (a multiple of 1024) from global src to LDS dst, 16 bytes
// per lane, each wave instruction filling one contiguous KiB; the block's
// `waves` waves share the copy.
TKA_DEV void dma_to_lds(const unsigned char* src, unsigned char* dst, int bytes, int waves)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int e = w; e < bytes / 1024; e += waves)
        __builtin_amdgcn_global_load_lds(src + (size_t)e * 1024 + 16 * lane, (lds_void*)(dst + e * 1024), 16, 0, 0);
}

// acc[u][t] += the products of k-steps [GQ, GQ + NVS) of the chunk image sA
// (row stride RSS, NVS vectors per half) with the wave's query operand.  Step
// g+1's four candidate fragments are read from LDS before step g's MFMAs
// issue, so the reads complete under them (the scheduling barriers keep the
// compiler from pairing each read with its MFMA behind a full wait).
template <class Op, int NVS, int RSS, int GQ, int QT, int NV>
TKA_DEV void tile_mma(const typename Op::V* sA, const typename Op::V (&bq)[QT][NV], f32x16 (&acc)[QT][4])
{
    typedef typename Op::V V;
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const V* rowp = sA + (lane & 31) * RSS + half * NVS;
    V af[4], an[4];
#pragma unroll
    for (int t = 0; t < 4; t++) af[t] = rowp[32 * t * RSS];
#pragma unroll
    for (int g = 0; g < NVS; g++) {
        if (g + 1 < NVS) {
#pragma unroll
            for (int t = 0; t < 4; t++) an[t] = rowp[32 * t * RSS + g + 1];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < QT; u++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[u][t] = Op::mma(af[t], bq[u][GQ + g], acc[u][t]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; t++) af[t] = an[t];
    }
}

// The same for the two candidate tiles T0, T0 + 1 of the chunk and one
// query tile: half the accumulators, so a kernel can hold four waves per SIMD.
template <class Op, int NVS, int RSS, int NV, int T0>
TKA_DEV void tile_mma2(const typename Op::V* sA, const typename Op::V (&bq)[1][NV], f32x16 (&acc)[2])
{
    typedef typename Op::V V;
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const V* rowp = sA + (lane & 31) * RSS + half * NVS + 32 * T0 * RSS;
    V af[2], an[2];
#pragma unroll
    for (int t = 0; t < 2; t++) af[t] = rowp[32 * t * RSS];
#pragma unroll
    for (int g = 0; g < NVS; g++) {
        if (g + 1 < NVS) {
#pragma unroll
            for (int t = 0; t < 2; t++) an[t] = rowp[32 * t * RSS + g + 1];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 2; t++) acc[t] = Op::mma(af[t], bq[0][g], acc[t]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 2; t++) af[t] = an[t];
    }
}

// kMode 0: the value is the accumulator (BF16: ||c||^2 folded into the operands)
//       1: ||c||^2 - 2 acc (FP32)
//       2: ||c||^2 - 2 s_q s_c acc (FP8 with per-row scales)
// QT: 32-query tiles per wave, WAVES: waves per block (block = 32 WAVES QT
// query rows).  With QT = 2 every candidate fragment read from LDS feeds two
// MFMAs; every chunk copied into LDS serves all the block's queries.  Two
// waves per SIMD (WAVES = 8, or two blocks per CU) let one wave's filter run
// under another's MFMAs.
template <class Op, int kMode, int NV, int QT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void knn_tiles_reg(const typename Op::V* __restrict__ A,
                                                            const typename Op::V* __restrict__ B, TileArgs a)
{
    typedef typename Op::V V;
    constexpr int QB = 32 * WAVES * QT;  // query rows per block
    // two chunks of kCB rows x (2*NV + 1) vectors
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ int s_cnt[QB];
    __shared__ float s_tau[QB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int slot0 = blockIdx.x * QB;
    constexpr int RS = 2 * NV + 1;
    const int chunk_bytes = kCB * RS * (int)sizeof(V);  // one LDS buffer
    for (int i = tid; i < QB; i += 64 * WAVES) {
        s_cnt[i] = 0;
        s_tau[i] = HUGE_VALF;
    }
    int sl[QT], q[QT];
    V bq[QT][NV];
    float sq[QT];
    uint64_t* qbuf[QT];
#pragma unroll
    for (int u = 0; u < QT; u++) {
        sl[u] = 32 * QT * w + 32 * u + (lane & 31);
        q[u] = query_row(a, slot0 + sl[u]);
        const V* qr = B + (size_t)(q[u] >= 0 ? q[u] : 0) * RS;
#pragma unroll
        for (int g = 0; g < NV; g++) bq[u][g] = (q[u] >= 0) ? qr[half * NV + g] : V{};
        sq[u] = (kMode == 2 && q[u] >= 0) ? reinterpret_cast<const float*>(qr + 2 * NV)[1] : 0.f;
        qbuf[u] = a.buf + (size_t)(slot0 + sl[u]) * kCap;
    }
    const unsigned char* src = reinterpret_cast<const unsigned char*>(A);
    const int nch = (a.N + kCB - 1) / kCB;
    const ChunkOrder order(seed_chunk(a, slot0, nch), nch);
    dma_to_lds(src + (size_t)order.at(0) * kCB * RS * sizeof(V), dyn_lds, chunk_bytes, WAVES);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
        const int cb = order.at(c) * kCB;
        f32x16 acc[QT][4];
#pragma unroll
        for (int u = 0; u < QT; u++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[u][t] = f32x16{};
        unsigned char* cur = dyn_lds + (c & 1) * chunk_bytes;
        if (c + 1 < nch)
            dma_to_lds(src + (size_t)order.at(c + 1) * kCB * RS * sizeof(V), dyn_lds + ((c + 1) & 1) * chunk_bytes,
                       chunk_bytes, WAVES);
        tile_mma<Op, NV, RS, 0, QT, NV>(reinterpret_cast<const V*>(cur), bq, acc);
        const V* sA = reinterpret_cast<const V*>(cur);  // the chunk with its pad vectors, for the filter
        constexpr int sRS = RS;
        const bool seeding = c < order.nseed;
#pragma unroll
        for (int u = 0; u < QT; u++) {
            const float tau = s_tau[sl[u]];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                float v[16];
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const int m = 32 * t + 8 * (i >> 2) + 4 * half + (i & 3);
                    if (kMode == 0) {
                        v[i] = acc[u][t][i];
                    } else {
                        const float* pad = reinterpret_cast<const float*>(sA + m * sRS + sRS - 1);
                        v[i] = kMode == 1 ? fmaf(-2.f, acc[u][t][i], pad[0])
                                          : fmaf(-2.f * sq[u] * pad[1], acc[u][t][i], pad[0]);
                    }
                }
                if (seeding) {
                    if (q[u] >= 0) seed_store(v, t, cb, q[u], c, a, qbuf[u]);
                } else {
                    filter_tile(v, t, cb, q[u], sl[u], tau, a, s_cnt, qbuf[u]);
                }
            }
        }
        if (c == order.nseed - 1 && lane < 32) {
#pragma unroll
            for (int u = 0; u < QT; u++)  // the seed chunks' slots (rows past N stored nothing)
                if (q[u] >= 0) s_cnt[sl[u]] = order.nseed * kCB;
        }
        if (c >= order.nseed - 1) compact_if_needed<32 * QT>(w, a, slot0, s_cnt, s_tau);
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copy of the next chunk (stage) has landed
        __syncthreads();                // every wave's has, and the buffer just used is free
    }
#pragma unroll
    for (int u = 0; u < QT; u++) {
        if (lane < 32 && slot0 + sl[u] < a.nq) {
            a.cnt_out[slot0 + sl[u]] = s_cnt[sl[u]];
            a.tau_out[slot0 + sl[u]] = s_tau[sl[u]];
        }
    }
}

// ---------------------------------------------------------------- BF16, block-centred
// One BF16 product per pair on values centred at the query block's mean m
// (a block's 256 queries are consecutive rows of the locality order, so they
// lie close together): candidate y~ = bf16(y - m), query q~ = bf16(q - m),
// value v = ||y~||^2 - 2 q~.y~ = ||q~ - y~||^2 - ||q~||^2 (the image row holds
// the D values, zeros, then ||y~||^2 as a hi/lo pair in its last two slots,
// against query values of 1).  The rounding error
// is relative to ||q - m|| and ||y - m|| instead of the largest row norm, so
// one product certifies where values on the global scale need the hi/lo
// split's three (LocalBound in the re-rank).  Candidate rows stream from HBM
// as FP32 into registers one chunk ahead; each chunk is converted to its BF16
// image (values, then the norm as a hi/lo pair against query values of 1) in
// one of two LDS buffers while the other feeds the MFMAs.  Filtering, seeds
// and compaction are knn_tiles_reg's.
constexpr int kLocNVMax = 10;            // widest image row: 16 x 10 values >= D + 2 (D <= 158)
constexpr int kLocWaves = 8;
constexpr int kLocQB = 32 * kLocWaves;   // query rows per block (and per centre)

inline int local_nv(int D) { return (D + 2 + 15) / 16; }  // D values + the norm pair
inline int local_dp(int D) { return 16 * local_nv(D); }  // FP32 row stride: the image width, zero past D

// Up to D = 62 (NV <= 4) the kernel fits 128 VGPRs: four waves per SIMD, so
// two blocks share a CU and 391 blocks at N=100k run in one round.
template <int NV>
__global__ __launch_bounds__(64 * kLocWaves) __attribute__((amdgpu_waves_per_eu(NV <= 4 ? 4 : 1)))
void knn_tiles_local(const float* __restrict__ Xf, int D, const float* __restrict__ centres, TileArgs a)
{
    constexpr int KP = 16 * NV;                     // values per image row = FP32 row stride
    constexpr int RS = 2 * NV + 1;                  // image row stride in vectors (odd: LDS banks)
    constexpr int VPT = (2 * NV + 3) / 4;           // image vectors a thread converts per chunk
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    bf16x8* img = reinterpret_cast<bf16x8*>(dyn_lds);  // two kCB x RS images
    __shared__ int s_cnt[kLocQB];
    __shared__ float s_tau[kLocQB];
    __shared__ __align__(16) float s_cen[16 * kLocNVMax + 32];  // the block's centre, zero past D
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int slot0 = blockIdx.x * kLocQB;
    for (int i = tid; i < kLocQB; i += 64 * kLocWaves) {
        s_cnt[i] = 0;
        s_tau[i] = HUGE_VALF;
    }
    for (int i = tid; i < 16 * kLocNVMax + 32; i += 64 * kLocWaves)
        s_cen[i] = i < KP ? centres[(size_t)blockIdx.x * KP + i] : 0.f;
    __syncthreads();
    // the query operand: -2 bf16(q - m) for the values, 1 against the norm pair
    const int sl = 32 * w + (lane & 31);
    const int q = query_row(a, slot0 + sl);
    bf16x8 bq[1][NV];
    {
        const float* xr = Xf + (size_t)(q >= 0 ? q : 0) * KP;
#pragma unroll
        for (int g = 0; g < NV; g++) {
#pragma unroll
            for (int el = 0; el < 8; el++) {
                const int p = (half * NV + g) * 8 + el;
                float v = -2.f * (float)(__bf16)(xr[p] - s_cen[p]);  // 0 past D (both are)
                v = p >= KP - 2 ? 1.f : v;                            // against the norm pair
                bq[0][g][el] = (__bf16)(q >= 0 ? v : 0.f);
            }
        }
    }
    uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
    // conversion roles: row r of a chunk, image vectors pp, pp + 4, ...; a
    // vector index past the row (2 NV not a multiple of 4) is clamped to the
    // last vector and its store skipped
    const int r = tid >> 2, pp = tid & 3;
    int vis[VPT];
#pragma unroll
    for (int k = 0; k < VPT; k++) vis[k] = min(pp + 4 * k, 2 * NV - 1);
    float4 pre[VPT][2];
    auto fetch = [&](int chunk) {
        const float4* xr = reinterpret_cast<const float4*>(Xf + ((size_t)chunk * kCB + r) * KP);
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            pre[k][0] = xr[2 * vis[k]];
            pre[k][1] = xr[2 * vis[k] + 1];
        }
    };
    auto convert = [&](bf16x8* dst) {
        bf16x8 out[VPT];
        float part = 0.f;
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const float4 c0 = reinterpret_cast<const float4*>(s_cen)[2 * vis[k]];
            const float4 c1 = reinterpret_cast<const float4*>(s_cen)[2 * vis[k] + 1];
            const float d[8] = {pre[k][0].x - c0.x, pre[k][0].y - c0.y, pre[k][0].z - c0.z, pre[k][0].w - c0.w,
                                pre[k][1].x - c1.x, pre[k][1].y - c1.y, pre[k][1].z - c1.z, pre[k][1].w - c1.w};
            const bool dup = pp + 4 * k >= 2 * NV;  // clamped vector: counted once
#pragma unroll
            for (int el = 0; el < 8; el++) {
                const __bf16 b = (__bf16)d[el];
                const float f = (float)b;
                part = dup ? part : fmaf(f, f, part);
                out[k][el] = b;
            }
        }
        part += __shfl_xor(part, 1, 64);  // the row's four parts, (p0 + p1) + (p2 + p3) on every lane
        part += __shfl_xor(part, 2, 64);
        const __bf16 nh = (__bf16)part;
        const __bf16 nl = (__bf16)(part - (float)nh);
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const bool last = vis[k] == 2 * NV - 1;  // the norm pair is the row's last two values
            out[k][6] = last ? nh : out[k][6];
            out[k][7] = last ? nl : out[k][7];
            if (pp + 4 * k < 2 * NV) dst[r * RS + vis[k]] = out[k];
        }
    };
    const int nch = (a.N + kCB - 1) / kCB;
    const ChunkOrder order(seed_chunk(a, slot0, nch), nch);
    fetch(order.at(0));
    convert(img);
    if (nch > 1) fetch(order.at(1));
    __syncthreads();
    for (int c = 0; c < nch; c++) {
        const int cb = order.at(c) * kCB;
        const bf16x8* cur = img + (size_t)(c & 1) * kCB * RS;
        const bool seeding = c < order.nseed;
        const float tau = s_tau[sl];
        // the chunk's four candidate tiles in two halves of two (half the
        // accumulators: four waves per SIMD)
        auto filter2 = [&](const f32x16 (&acc)[2], int t0) {
#pragma unroll
            for (int t = 0; t < 2; t++) {
                float v[16];
#pragma unroll
                for (int i = 0; i < 16; i++) v[i] = acc[t][i];
                if (seeding) {
                    if (q >= 0) seed_store(v, t0 + t, cb, q, c, a, qbuf);
                } else {
                    filter_tile(v, t0 + t, cb, q, sl, tau, a, s_cnt, qbuf);
                }
            }
        };
        {
            f32x16 acc[2] = {f32x16{}, f32x16{}};
            tile_mma2<OpBF16, NV, RS, NV, 0>(cur, bq, acc);
            if (c + 1 < nch) {  // the next chunk's image (its buffer was last read before the previous barrier)
                convert(img + (size_t)((c + 1) & 1) * kCB * RS);
                if (c + 2 < nch) fetch(order.at(c + 2));
            }
            filter2(acc, 0);
        }
        {
            f32x16 acc[2] = {f32x16{}, f32x16{}};
            tile_mma2<OpBF16, NV, RS, NV, 2>(cur, bq, acc);
            filter2(acc, 2);
        }
        if (c == order.nseed - 1 && lane < 32 && q >= 0) s_cnt[sl] = order.nseed * kCB;
        if (c >= order.nseed - 1) compact_if_needed<32>(w, a, slot0, s_cnt, s_tau);
        __syncthreads();
    }
    if (lane < 32 && slot0 + sl < a.nq) {
        a.cnt_out[slot0 + sl] = s_cnt[sl];
        a.tau_out[slot0 + sl] = s_tau[sl];
    }
}

// FP32 copy of the rows (Npad x Dp, zero padding) and every query block's
// mean (FP64 sums, stored as FP32): the centres of knn_tiles_local.
__global__ __launch_bounds__(256) void prep_local_rows(const double* __restrict__ X, int N, int Npad, int D, int Dp,
                                                       float* __restrict__ Xf)
{
    const size_t total = (size_t)Npad * Dp;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / Dp), d = (int)(e % Dp);
        Xf[e] = (n < N && d < D) ? (float)X[(size_t)n * D + d] : 0.f;
    }
}

__global__ __launch_bounds__(256) void prep_local_centres(const double* __restrict__ X, int N, int D, int Dp,
                                                          float* __restrict__ centres)
{
    __shared__ double sh[4][64];
    const int r0 = blockIdx.x * kLocQB, r1 = min(N, r0 + kLocQB);
    const int dl = threadIdx.x & 63, ph = threadIdx.x >> 6;  // column, row phase (every fourth row)
    for (int d0 = 0; d0 < Dp; d0 += 64) {
        const int d = d0 + dl;
        double sum = 0.0;
        if (d < D)
            for (int r = r0 + ph; r < r1; r += 4) sum += X[(size_t)r * D + d];
        sh[ph][dl] = sum;
        __syncthreads();
        if (ph == 0 && d < Dp) {
            const double t = ((sh[0][dl] + sh[1][dl]) + sh[2][dl]) + sh[3][dl];
            centres[(size_t)blockIdx.x * Dp + d] = d < D && r1 > r0 ? (float)(t / (r1 - r0)) : 0.f;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- wide rows
// Rows too wide to keep a query operand in registers (BF16 beyond the centred
// tiles' D = 158, on the hi/lo splits; FP32 beyond 20 vectors per half row,
// D above 77): both operands are
// streamed through LDS in k-stages of KS vectors per half row.  The operands
// are stage-major (hm_to_stages): stage st of row r is the 2 KS vectors
// (h, st KS + g), h = 0, 1, g < KS, at ((st Npad) + r) 2 KS, so a stage of a
// chunk is one contiguous block.  A stage image holds those vectors for the
// kCB candidate rows of a chunk and then for the block's kWideQB query rows,
// with a pad slot per row (RSS = 2 KS + 1 vectors, an odd stride for the LDS
// banks), copied by direct LDS loads.  Every wave issues the same number of
// them per stage (the last one is repeated where the image does not divide
// evenly).
constexpr int kWideQB = 128;    // query rows per block: four 32-row tiles
constexpr int kWideWaves = 8;   // two waves per query tile, two candidate tiles each

template <int KS>
struct WideStage {
    static constexpr int RSS = 2 * KS + 1;
    static constexpr int IMG = (kCB + kWideQB) * RSS;            // vectors per stage image
    static constexpr int TL = IMG / 64;                           // direct loads per image
    static constexpr int L = (TL + kWideWaves - 1) / kWideWaves;  // per wave
    static_assert(IMG % 64 == 0 && kCB * RSS % 64 == 0, "whole direct loads, rows of one kind per load");
};

// Stage i of the stream (chunk order.at(i / nst), stage i % nst) into image
// buffer i % 2: the lane's direct loads at their precomputed offsets.
template <int KS>
TKA_DEV void wide_copy(const uint4* __restrict__ A, const uint4* __restrict__ B, int Npad, int nst,
                       const ChunkOrder& order, int i, const int* off, const bool* isq, unsigned char* lds, int w)
{
    typedef WideStage<KS> S;
    const size_t so = (size_t)(i % nst) * Npad * (2 * KS);
    const uint4* ca = A + so + (size_t)order.at(i / nst) * kCB * (2 * KS);
    const uint4* qb = B + so;
    unsigned char* dst = lds + (size_t)(i & 1) * S::IMG * 16;
#pragma unroll
    for (int j = 0; j < S::L; j++) {
        const int e = min(w + j * kWideWaves, S::TL - 1);
        __builtin_amdgcn_global_load_lds((isq[j] ? qb : ca) + off[j], (lds_void*)(dst + (size_t)e * 1024), 16, 0, 0);
    }
}

// All rows against the block's query rows with both operands streamed
// through two stage images (the next one is copied while this one is
// multiplied).  Wave w multiplies query tile w % 4 by candidate tiles
// 2 (w / 4) and 2 (w / 4) + 1 of the chunk, so two waves on each SIMD
// overlap one's copies and sums with the other's MFMAs.  Each stage is
// accumulated from zero and added to the chunk's running sums (no value
// carries more than 16 KS (BF16) or 8 KS (FP32) MFMA roundings plus one
// addition per stage).  After a chunk's last stage its values are filtered
// exactly as in knn_tiles_reg (same seeds, order, appends -- by LDS atomics,
// so the two waves of a query tile share its rows -- and compaction), so the
// candidate buffers have the same meaning.
template <class Op, int KS>
__global__ __launch_bounds__(64 * kWideWaves) void knn_tiles_wide(const typename Op::V* __restrict__ A,
                                                                  const typename Op::V* __restrict__ B, int NV,
                                                                  int Npad, TileArgs a)
{
    typedef typename Op::V V;
    typedef WideStage<KS> S;
    static_assert(sizeof(V) == 16, "one vector per lane per direct load");
    constexpr int RSS = S::RSS;
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const V* lds = reinterpret_cast<const V*>(dyn_lds);  // two stage images
    __shared__ int s_cnt[kWideQB];
    __shared__ float s_tau[kWideQB];
    __shared__ int s_qrow[kWideQB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int qt = w & 3, ch = w >> 2;  // query tile, candidate half
    const int slot0 = blockIdx.x * kWideQB;
    for (int i = tid; i < kWideQB; i += 64 * kWideWaves) {
        s_cnt[i] = 0;
        s_tau[i] = HUGE_VALF;
        s_qrow[i] = max(query_row(a, slot0 + i), 0);  // slots past the queries copy row 0
    }
    const int sl = 32 * qt + (lane & 31);
    const int q = query_row(a, slot0 + sl);
    uint64_t* qbuf = a.buf + (size_t)(slot0 + sl) * kCap;
    const int nch = (a.N + kCB - 1) / kCB;
    const int nst = NV / KS;
    const int T = nch * nst;  // stages of the whole stream
    const ChunkOrder order(seed_chunk(a, slot0, nch), nch);
    __syncthreads();  // s_qrow
    // this lane's direct loads: source offset within a stage (candidate rows
    // relative to the chunk) and whether it is a query row
    int off[S::L];
    bool isq[S::L];
#pragma unroll
    for (int j = 0; j < S::L; j++) {
        const int e = min(w + j * kWideWaves, S::TL - 1);
        const int v = e * 64 + lane, r = v / RSS, k = v - r * RSS;
        const int kk = k < 2 * KS ? k : 0;  // the pad slot: any vector of the row (never read)
        isq[j] = r >= kCB;
        off[j] = (isq[j] ? s_qrow[r - kCB] : r) * (2 * KS) + kk;
    }
    const uint4* A4 = reinterpret_cast<const uint4*>(A);
    const uint4* B4 = reinterpret_cast<const uint4*>(B);
    unsigned char* ring = dyn_lds;
    wide_copy<KS>(A4, B4, Npad, nst, order, 0, off, isq, ring, w);
    int i = 0;  // stage of the stream
    for (int c = 0; c < nch; c++) {
        const int cb = order.at(c) * kCB;
        f32x16 tot[2];
#pragma unroll
        for (int t = 0; t < 2; t++) tot[t] = f32x16{};
        for (int st = 0; st < nst; st++, i++) {
            __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of stage i have landed
            __syncthreads();                // every wave's have; the other image is free
            if (i + 1 < T) wide_copy<KS>(A4, B4, Npad, nst, order, i + 1, off, isq, ring, w);
            const V* sC = lds + (i & 1) * S::IMG;
            const V* crow = sC + (64 * ch + (lane & 31)) * RSS + half * KS;
            const V* qrow = sC + kCB * RSS + sl * RSS + half * KS;
            // every fragment of the stage first, then the MFMAs back to back
            V fa[KS][2], fb[KS];
#pragma unroll
            for (int g = 0; g < KS; g++) {
                fb[g] = qrow[g];
#pragma unroll
                for (int t = 0; t < 2; t++) fa[g][t] = crow[32 * t * RSS + g];
            }
            f32x16 acc[2];
#pragma unroll
            for (int g = 0; g < KS; g++)
#pragma unroll
                for (int t = 0; t < 2; t++) acc[t] = Op::mma(fa[g][t], fb[g], g == 0 ? f32x16{} : acc[t]);
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int e = 0; e < 16; e++) tot[t][e] += acc[t][e];
            if (st == nst - 1) {  // the chunk's values (norms folded into the operands)
                const bool seeding = c < order.nseed;
                const float tau = s_tau[sl];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    float v[16];
#pragma unroll
                    for (int e = 0; e < 16; e++) v[e] = tot[t][e];
                    if (seeding) {
                        if (q >= 0) seed_store(v, 2 * ch + t, cb, q, c, a, qbuf);
                    } else {
                        filter_tile(v, 2 * ch + t, cb, q, sl, tau, a, s_cnt, qbuf);
                    }
                }
                __syncthreads();  // both candidate halves of the chunk are in the rows' buffers
                if (ch == 0) {
                    if (c == order.nseed - 1 && lane < 32 && q >= 0) s_cnt[sl] = order.nseed * kCB;
                    if (c >= order.nseed - 1) compact_if_needed<32>(qt, a, slot0, s_cnt, s_tau);
                }
            }
        }
    }
    __syncthreads();
    if (ch == 0 && lane < 32 && slot0 + sl < a.nq) {
        a.cnt_out[slot0 + sl] = s_cnt[sl];
        a.tau_out[slot0 + sl] = s_tau[sl];
    }
}

// Padded half-major rows (RS = 2 NV + 1 vectors, Npad rows) to the
// stage-major layout of knn_tiles_wide (stages of KS vectors per half).
template <int KS>
__global__ __launch_bounds__(256) void hm_to_stages(const uint4* __restrict__ hm, int Npad, int NV,
                                                    uint4* __restrict__ W)
{
    const int RS = 2 * NV + 1;
    const size_t total = (size_t)(NV / KS) * Npad * 2 * KS;
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < total; o += (size_t)gridDim.x * 256) {
        const int k = (int)(o % (2 * KS));
        const size_t sr = o / (2 * KS);
        const int r = (int)(sr % Npad), st = (int)(sr / Npad);
        const int h = k / KS, g = st * KS + (k - h * KS);
        W[o] = hm[(size_t)r * RS + h * NV + g];
    }
}

constexpr int kWideKS = 8;  // vectors per half row per k-stage of knn_tiles_wide

// Operand width (vectors per half row) of the wide tiles: a whole number of stages.
inline int wide_nv(int NV) { return (NV + kWideKS - 1) / kWideKS * kWideKS; }

// BF16 operands over the k-values
//   candidates A = [c_hi | c_lo | c_hi | n_hi n_lo | 0...]
//   queries    B = [-2 q_hi | -2 q_hi | -2 q_lo | 1 1 | 0...]
// so that A.B = ||c||^2 - 2 q.c with the q_lo*c_lo term dropped (~2^-16
// relative); scaling by -2 is exact in BF16.  Padded half-major rows:
// vector (h, g) of a row holds k = 16 g + 8 h + 0..7; the pad vector and
// the rows past N are zero.  A and B share the layout.
__global__ __launch_bounds__(256) void prep_bf16x3_hm(const double* __restrict__ X, const double* __restrict__ n2,
                                                      int N, int Npad, int D, int NV, __bf16* __restrict__ A,
                                                      __bf16* __restrict__ B)
{
    const int RSe = 8 * (2 * NV + 1);  // bf16 per padded row
    const size_t total = (size_t)Npad * RSe;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / RSe), p = (int)(e % RSe);
        float av = 0.f, bv = 0.f;
        if (n < N && p < 16 * NV) {
            const int vi = p >> 3, el = p & 7, h = vi / NV, g = vi - h * NV;
            const int k = 16 * g + 8 * h + el;
            if (k < 3 * D) {
                const int seg = k / D, d = k % D;
                const double x = X[(size_t)n * D + d];
                const __bf16 hi = (__bf16)x;
                const __bf16 lo = (__bf16)(x - (double)(float)hi);
                if (seg == 0) {
                    av = (float)hi;
                    bv = -2.f * (float)hi;
                } else if (seg == 1) {
                    av = (float)lo;
                    bv = -2.f * (float)hi;
                } else {
                    av = (float)hi;
                    bv = -2.f * (float)lo;
                }
            } else if (k == 3 * D || k == 3 * D + 1) {
                const double nn = n2[n];
                const __bf16 hi = (__bf16)nn;
                const __bf16 lo = (__bf16)(nn - (double)(float)hi);
                av = (k == 3 * D) ? (float)hi : (float)lo;
                bv = 1.f;
            }
        }
        A[e] = (__bf16)av;
        B[e] = (__bf16)bv;
    }
}

// Padded half-major FP32 operands with the norm folded in, so the MFMA
// chain yields ||c||^2 - 2 q.c: candidates A = [c | ||c||^2], queries
// B = [-2q | 1] (x -2 is exact), D+1 k-values.  S = 4 NV k-steps of 2 per
// half row; position h*S + s holds k = 2 s + h; the pad vector and the rows
// past N are zero.
__global__ __launch_bounds__(256) void prep_fp32n_hm(const double* __restrict__ X, const double* __restrict__ n2,
                                                     int N, int Npad, int D, int NV, float* __restrict__ A,
                                                     float* __restrict__ B)
{
    const int S = 4 * NV, RSe = 4 * (2 * NV + 1);
    const size_t total = (size_t)Npad * RSe;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / RSe), p = (int)(e % RSe);
        float av = 0.f, bv = 0.f;
        if (n < N && p < 2 * S) {
            const int h = p / S, k = 2 * (p - h * S) + h;
            if (k < D) {
                const float x = (float)X[(size_t)n * D + k];
                av = x;
                bv = -2.f * x;
            } else if (k == D) {
                av = (float)n2[n];
                bv = 1.f;
            }
        }
        A[e] = av;
        B[e] = bv;
    }
}

// Padded half-major FP8 operand from the k-ordered rows written by prep_fp8
// (vector (h, g) = 8 bytes holding k = 16 g + 8 h + 0..7); the pad vector
// holds the floats {||x||^2, scale}.
__global__ __launch_bounds__(256) void fp8_to_hm(const uint8_t* __restrict__ src, const float* __restrict__ nrm,
                                                 const float* __restrict__ scale, int N, int Npad, int NVsrc, int NV,
                                                 uint64_t* __restrict__ dst)
{
    const int RS = 2 * NV + 1;
    const size_t total = (size_t)Npad * RS;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int n = (int)(e / RS), vi = (int)(e % RS);
        uint64_t v = 0;
        if (n < N) {
            if (vi < 2 * NV) {
                const int h = vi / NV, g = vi - h * NV;  // source rows: NVsrc vectors per half
                if (g < NVsrc) v = *reinterpret_cast<const uint64_t*>(src + (size_t)n * 16 * NVsrc + 16 * g + 8 * h);
            } else {
                v = ((uint64_t)__float_as_uint(scale[n]) << 32) | __float_as_uint(nrm[n]);
            }
        }
        dst[e] = v;
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
    const double* rmax;   // device word: the largest row norm
    int* idx;             // N x K (indexed by row)
    double* dist;
    int* fail;            // per slot: 1 if the bound failed
    double* ubound;       // per slot: the K-th exact distance found (>= the true one), +inf if fewer than K
    const uint32_t* perm; // optional: X rows are the caller's rows perm[0..N) (outputs and ties use those)
    // block-centred BF16 tiles (knn_tiles_local): per-row bound instead of eps_a, eps_b
    const float* centres; // null: global bound
    int cdim;             // centre stride (Dp)
    int kp;               // values per image row (16 NV)
};

// Bounds on the exact squared distance d of a candidate from its value v in
// knn_tiles_local, for a query q of centre m.  q~, y~ are the BF16 values
// (FP32 copy, FP32 subtraction of m, round to nearest BF16: per value
// |y~ - (y - m)| <= u |y - m| + 2^-24 |y|), a = ||q - m||, Q = ||q~||^2 and
// e = ||q~ - (q - m)|| (both exact: the re-rank repeats the query's
// rounding), s = ||q~ - y~||, w = 2^-24 Rmax:
//   |sqrt(d) - s| <= u ||y - m|| + w + e,  ||y - m|| <= sqrt(d) + a;
//   v = s^2 - Q + E,  |E| <= g (||y~||^2 + 2 sqrt(Q) ||y~||) (FP32
//   accumulation of the KP products and of the norm, hi/lo split of the
//   norm), ||y~|| <= s + sqrt(Q).
// Solving these for s, then for sqrt(d), gives lower(v) <= d for every
// candidate of value >= v, and upper(v) >= d for every candidate of value <= v.
struct LocalBound {
    double u, g, t, Q, qn;  // t: the additive part u a + w + e
    TKA_DEV LocalBound(double a, double Q_, double e, double rmax, int kp, int D) : Q(Q_)
    {
        u = 0x1.01p-9;
        g = (2.0 * kp + D + 4) * 0x1.0p-24 + 0x1.0p-17;
        t = u * a + 0x1.0p-24 * rmax + e;
        qn = sqrt(Q);
    }
    TKA_DEV double lower(double v) const
    {
        const double R = v + Q - 3.0 * g * Q;
        if (!(R > 0.0)) return 0.0;
        const double sl = (-4.0 * g * qn + sqrt(16.0 * g * g * Q + 4.0 * (1.0 + g) * R)) / (2.0 * (1.0 + g));
        const double r = (sl - t) / (1.0 + u);
        return r > 0.0 ? r * r * (1.0 - 1e-12) : 0.0;
    }
    TKA_DEV double upper(double v) const
    {
        double R = v + Q + 3.0 * g * Q;
        R = R > 0.0 ? R : 0.0;
        const double sh = (4.0 * g * qn + sqrt(16.0 * g * g * Q + 4.0 * (1.0 - g) * R)) / (2.0 * (1.0 - g));
        const double r = (sh + t) / (1.0 - u);
        return r * r * (1.0 + 1e-12);
    }
};


// Exact FP64 squared distance, summed in index order without contraction
// (the reference's euclidean_distance, vptree.hpp:95-101).
TKA_DEV double exact_sqdist(const double* __restrict__ xq, const double* __restrict__ xj, int D)
{
    double dd = 0.0;
    int d = 0;
    if ((((uintptr_t)xq | (uintptr_t)xj) & 15) == 0) {  // 16-byte aligned rows: pairs of values per load,
                                                         // sums still in index order
        for (; d + 8 <= D; d += 8) {
            double2 a[4], b[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                a[k] = *reinterpret_cast<const double2*>(xq + d + 2 * k);
                b[k] = *reinterpret_cast<const double2*>(xj + d + 2 * k);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double t0 = dadd_sep(a[k].x, -b[k].x);
                dd = dadd_sep(dd, dmul_sep(t0, t0));
                const double t1 = dadd_sep(a[k].y, -b[k].y);
                dd = dadd_sep(dd, dmul_sep(t1, t1));
            }
        }
    }
    for (; d + 8 <= D; d += 8) {  // loads of a group issued together; sums stay in index order
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a[k] = xq[d + k];
            b[k] = xj[d + k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const double t = dadd_sep(a[k], -b[k]);
            dd = dadd_sep(dd, dmul_sep(t, t));
        }
    }
    for (; d < D; d++) {
        const double t = dadd_sep(xq[d], -xj[d]);
        dd = dadd_sep(dd, dmul_sep(t, t));
    }
    return dd;
}

// Sorts the first 64*R entries (R in {1, 2, 4, 8}: the smallest that holds them).
TKA_DEV void sort_prefix(uint64_t (&d)[8], uint32_t (&ix)[8], int R)
{
    if (R <= 1)
        wave_sort_pairs<1>(d, ix);
    else if (R <= 2)
        wave_sort_pairs<2>(d, ix);
    else if (R <= 4)
        wave_sort_pairs<4>(d, ix);
    else
        wave_sort_pairs<8>(d, ix);
}

// ---------------------------------------------------------------- large K
// K beyond the candidate buffers (kMaxKprime) up to kLargeKMax: every
// distance in FP64 with the re-rank's arithmetic, then an exact selection
// per row, so no certificate is needed.
constexpr int kLargeKMax = 8192;
constexpr int kPairTile = 64;   // query and candidate rows per workgroup of pair_sqdist_rows
constexpr int kPairDims = 32;   // dimensions per LDS stage

// out[r N + c] = squared distance of query row q(r) to row c, for the B
// query rows q(r) = qlist[q0 + r] (or q0 + r without a list), summed in
// dimension order without contraction (exact_sqdist's arithmetic); the row
// itself is +inf.  64 x 64 pairs per workgroup, 4 x 4 per thread,
// coordinates staged through LDS 32 dimensions at a time.
__global__ __launch_bounds__(256) void pair_sqdist_rows(const double* __restrict__ X, int N, int D,
                                                        const int* __restrict__ qlist, int q0, int B,
                                                        double* __restrict__ out)
{
    __shared__ double sq[kPairTile][kPairDims + 1], sc[kPairTile][kPairDims + 1];
    const int qb = blockIdx.y * kPairTile, cb = blockIdx.x * kPairTile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int d0 = 0; d0 < D; d0 += kPairDims) {
        const int dn = min(kPairDims, D - d0);
        for (int e = threadIdx.x; e < kPairTile * kPairDims; e += 256) {
            const int r = e / kPairDims, k = e % kPairDims;
            const bool kin = k < dn;
            const int qr = qb + r < B ? (qlist ? qlist[q0 + qb + r] : q0 + qb + r) : 0;
            sq[r][k] = (kin && qb + r < B) ? X[(size_t)qr * D + d0 + k] : 0.0;
            sc[r][k] = (kin && cb + r < N) ? X[(size_t)(cb + r) * D + d0 + k] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < dn; k++) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                a[i] = sq[4 * ty + i][k];
                b[i] = sc[4 * tx + i][k];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const double t = dadd_sep(a[i], -b[j]);
                    acc[i][j] = dadd_sep(acc[i][j], dmul_sep(t, t));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int q = qb + 4 * ty + i;
        if (q >= B) continue;
        const int qr = qlist ? qlist[q0 + q] : q0 + q;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = cb + 4 * tx + j;
            if (c < N) out[(size_t)q * N + c] = (qr == c) ? HUGE_VAL : acc[i][j];
        }
    }
}

// Block-wide exclusive prefix of one count per thread (256 threads), in
// thread order; the block total in *total.
TKA_DEV int block_excl_scan(int v, int* sh, int* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    __syncthreads();
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; k++) base += sh[k];
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    return base + x - v;
}

// The K smallest (distance, index) of each of B rows of `dist` (N values,
// non-negative or +inf), sorted: a radix select of the K-th distance's bits
// (six rounds of 11, 11, 11, 11, 11 and 9 bits), the rows below it and the
// equal ones of smallest index gathered into LDS, then a bitonic sort of
// (bits, index).  One workgroup per row; K <= kLargeKMax.  Ties are broken
// by row index (X in the caller's order).
__global__ __launch_bounds__(256) void select_k_rows(const double* __restrict__ dist, int N, int B, int K,
                                                     const int* __restrict__ qlist, int q0,
                                                     const double* __restrict__ ubound, int* __restrict__ idx,
                                                     double* __restrict__ dout)
{
    __shared__ uint64_t skey[kLargeKMax];
    __shared__ int sidx[kLargeKMax];
    __shared__ int hist[2048];
    __shared__ int sh[4];
    __shared__ int s_cnt, s_digit, s_below;
    const int r = blockIdx.x;
    if (r >= B) return;
    const uint64_t* row = reinterpret_cast<const uint64_t*>(dist + (size_t)r * N);
    // With an upper bound U on the row's K-th distance (the K-th exact
    // distance of candidates already seen), the K smallest are among the
    // distances <= U: gathered in one pass and sorted, when they fit.
    int nsort = K;
    bool bounded = false;
    if (ubound != nullptr && isfinite(ubound[q0 + r])) {
        const uint64_t U = (uint64_t)__double_as_longlong(ubound[q0 + r]);
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        for (int c = threadIdx.x; c < N; c += 256) {
            const uint64_t key = row[c];
            if (key <= U) {
                const int pos = atomicAdd(&s_cnt, 1);
                if (pos < kLargeKMax) {
                    skey[pos] = key;
                    sidx[pos] = c;
                }
            }
        }
        __syncthreads();
        bounded = s_cnt >= K && s_cnt <= kLargeKMax;
        nsort = bounded ? s_cnt : K;
        __syncthreads();
    }
    if (!bounded) {
    uint64_t prefix = 0ull, pmask = 0ull;
    int k = K;  // rank (1-based) of the wanted key among the keys that match the prefix
    for (int round = 0; round < 6; round++) {
        const int w = round < 5 ? 11 : 9;
        const int sft = 64 - 11 * (round + 1) < 0 ? 0 : 64 - 11 * (round + 1);
        const int nb = 1 << w;
        for (int b = threadIdx.x; b < nb; b += 256) hist[b] = 0;
        __syncthreads();
        for (int c = threadIdx.x; c < N; c += 256) {
            const uint64_t key = row[c];
            if ((key & pmask) == prefix) atomicAdd(&hist[(int)((key >> sft) & (uint64_t)(nb - 1))], 1);
        }
        __syncthreads();
        // the bin holding the k-th key: per-thread partial sums of 8 bins, a block scan, then the thread whose range
        // crosses k
        const int per = nb / 256;
        int local = 0;
        for (int b = 0; b < per; b++) local += hist[threadIdx.x * per + b];
        int tot;
        const int before = block_excl_scan(local, sh, &tot);
        if (before < k && k <= before + local) {
            int acc = before;
            for (int b = 0; b < per; b++) {
                const int h = hist[threadIdx.x * per + b];
                if (acc + h >= k) {
                    s_digit = threadIdx.x * per + b;
                    s_below = acc;
                    break;
                }
                acc += h;
            }
        }
        __syncthreads();
        prefix |= (uint64_t)s_digit << sft;
        pmask |= (uint64_t)(nb - 1) << sft;
        k -= s_below;
        __syncthreads();
    }
    // prefix is the K-th key; K - k keys lie below it; take the k equal keys of smallest index
    const uint64_t T = prefix;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    int eq_taken = 0;
    for (int c0 = 0; c0 < N; c0 += 256) {
        const int c = c0 + threadIdx.x;
        const uint64_t key = c < N ? row[c] : ~0ull;
        if (key < T) {
            const int pos = atomicAdd(&s_cnt, 1);
            skey[pos] = key;
            sidx[pos] = c;
        }
        const int eq = (key == T) ? 1 : 0;
        int tot;
        const int rank = block_excl_scan(eq, sh, &tot);
        if (eq && eq_taken + rank < k) {
            skey[K - k + eq_taken + rank] = key;
            sidx[K - k + eq_taken + rank] = c;
        }
        eq_taken += tot;
    }
    }  // radix select
    __syncthreads();
    // bitonic sort of (key, index) over the next power of two >= nsort, padded with +inf
    int P2 = 1;
    while (P2 < nsort) P2 <<= 1;
    for (int i = nsort + threadIdx.x; i < P2; i += 256) {
        skey[i] = ~0ull;
        sidx[i] = INT_MAX;
    }
    __syncthreads();
    for (int size = 2; size <= P2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < P2 / 2; t += 256) {
                const int i = 2 * stride * (t / stride) + (t % stride), j = i + stride;
                const bool up = (i & size) == 0;
                const uint64_t ki = skey[i], kj = skey[j];
                const int ii = sidx[i], ij = sidx[j];
                const bool gt = ki > kj || (ki == kj && ii > ij);
                if (gt == up) {
                    skey[i] = kj;
                    skey[j] = ki;
                    sidx[i] = ij;
                    sidx[j] = ii;
                }
            }
            __syncthreads();
        }
    }
    const int qr = qlist ? qlist[q0 + r] : q0 + r;
    for (int j = threadIdx.x; j < K; j += 256) {
        idx[(size_t)qr * K + j] = sidx[j];
        dout[(size_t)qr * K + j] = __longlong_as_double((long long)skey[j]);
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
    const int qo = a.perm ? (int)a.perm[q] : q;  // the caller's row
    const int c = a.cnt[slot] < kCap ? a.cnt[slot] : kCap;
    const uint64_t* qbuf = a.buf + (size_t)slot * kCap;
    const double* xq = a.X + (size_t)q * a.D;
    const double qn2 = a.n2[q];
    const double rmax = *a.rmax;
    const double eps = a.eps_a * sqrt(qn2) * rmax + a.eps_b * rmax * rmax + 1e-12 * (qn2 + rmax * rmax);
    const int K = a.K;
    const bool local = a.centres != nullptr;
    double ac = 0.0, Q = 0.0, eq = 0.0;  // ||q - m||, ||q~||^2, ||q~ - (q - m)|| for the block-centred tiles
    if (local) {
        const float* m = a.centres + (size_t)(slot / kLocQB) * a.cdim;
        for (int d = lane; d < a.D; d += 64) {
            const double t = xq[d] - (double)m[d];
            ac += t * t;
            const double b = (double)(float)(__bf16)((float)xq[d] - m[d]);  // the tiles' rounding
            Q += b * b;
            eq += (b - t) * (b - t);
        }
        ac = sqrt(wave_sum(ac)) * (1.0 + 1e-12);
        Q = wave_sum(Q);
        eq = sqrt(wave_sum(eq)) * (1.0 + 1e-12) + 1e-15 * ac;  // t carries one FP64 rounding per value
    }
    const LocalBound lb(ac, Q, eq, rmax, a.kp, a.D);

    __shared__ uint64_t s_list[4][kCap];
    const int w = threadIdx.x >> 6;
    uint64_t u[8];
    // Every slot below cnt holds a stored candidate (seed chunks store all
    // their slots, appends store every slot they reserve, compaction rewrites
    // the kept prefix), so a row index >= N cannot come from the tiles; it is
    // still never gathered: such an entry is dropped and the row is reported
    // as failed instead of reading X out of bounds.
    bool foreign = false;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        u[r] = p < c ? qbuf[p] : ~0ull;
        if (u[r] != ~0ull && (uint32_t)u[r] >= (uint32_t)a.N) {
            foreign = true;
            u[r] = ~0ull;
        }
    }
    foreign = __ballot(foreign) != 0ull;
    // v_(K) exactly (radix select), then the candidates within thr, packed
    const uint32_t vk_bits = wave_select<8>(u, K, 4);
    const double thr = c >= K ? (local ? lb.upper((double)ord2f(vk_bits)) : (double)ord2f(vk_bits) + 2.0 * eps)
                              : HUGE_VAL;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int m = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const double vr = (double)ord2f((uint32_t)(u[r] >> 32));
        const bool in = u[r] != ~0ull && (local ? lb.lower(vr) : vr) <= thr;
        const uint64_t im = __ballot(in);
        if (in) s_list[w][m + __popcll(im & lt)] = u[r];
        m += __popcll(im);
    }
    wave_lds_sync();

    uint64_t vd[8];
    uint32_t vi[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        if (p < m) {
            const uint32_t j = (uint32_t)s_list[w][p];
            vd[r] = (uint64_t)__double_as_longlong(exact_sqdist(xq, a.X + (size_t)j * a.D, a.D));
            vi[r] = a.perm ? a.perm[j] : j;
        } else {
            vd[r] = ~0ull;
            vi[r] = 0xffffffffu;
        }
    }
    const int R = m <= 64 ? 1 : m <= 128 ? 2 : m <= 256 ? 4 : 8;
    sort_prefix(vd, vi, R);
    double dk = 0.0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int p = r * 64 + lane;
        if (p < K) {
            a.idx[(size_t)qo * K + p] = (int)vi[r];
            a.dist[(size_t)qo * K + p] = __longlong_as_double((long long)vd[r]);
        }
        if (p == K - 1) dk = __longlong_as_double((long long)vd[r]);
    }
    dk = __shfl(dk, (K - 1) & 63, 64);
    if (lane == 0) {
        const float tau = a.tau[slot];
        int bad = 0;
        if (c < K || foreign) {
            bad = 1;  // cannot happen with K <= N-1 unless values were non-finite
        } else if (isfinite(tau)) {
            bad = local ? !(dk < lb.lower((double)tau)) : !(dk < (double)tau + qn2 - eps);
        }
        a.fail[slot] = bad;
        a.ubound[slot] = c >= K && !foreign ? dk : HUGE_VAL;
    }
}

__global__ void collect_failures(const int* __restrict__ fail, const double* __restrict__ ubound, int n,
                                 const uint32_t* __restrict__ perm, int* __restrict__ list, double* __restrict__ ulist,
                                 int* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && fail[i]) {
        const int p = atomicAdd(count, 1);
        list[p] = perm ? (int)perm[i] : i;
        ulist[p] = ubound[i];
    }
}

// ---------------------------------------------------------------- locality order
// Rows are streamed in an order where nearby rows sit in nearby chunks, so a
// block's first chunks (its own rows, rotate = 1) already hold most of its
// neighbours and the thresholds fall early: few values pass the filter
// afterwards.  Order: Morton code of the rows' projection onto an estimate of
// their top-3 principal subspace (two rounds of subspace iteration from a
// fixed random projection, R <- orth(X^T X R)).  On the 50-D swiss roll the
// four seed chunks around a row then hold 76% of its 90 nearest neighbours,
// against 41% for the random projection.  The kNN result does not depend on
// the order (ties are broken by the caller's row index in the re-rank).
__global__ __launch_bounds__(256) void proj3(const double* __restrict__ X, int N, int D, const float* __restrict__ R,
                                             float* __restrict__ P, float* __restrict__ part)
{
    __shared__ float sh[6][4];
    const int n = blockIdx.x * 256 + threadIdx.x;
    float v[6] = {HUGE_VALF, HUGE_VALF, HUGE_VALF, -HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
    if (n < N) {
        float p[3] = {0.f, 0.f, 0.f};
        for (int d = 0; d < D; d++) {
            const float x = (float)X[(size_t)n * D + d];
#pragma unroll
            for (int k = 0; k < 3; k++) p[k] = fmaf(x, R[3 * d + k], p[k]);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            P[3 * (size_t)n + k] = p[k];
            v[k] = v[k + 3] = p[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        for (int o = 32; o > 0; o >>= 1) {
            v[k] = fminf(v[k], __shfl_xor(v[k], o, 64));
            v[k + 3] = fmaxf(v[k + 3], __shfl_xor(v[k + 3], o, 64));
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        float r = sh[threadIdx.x][0];
        for (int w = 1; w < 4; w++) r = threadIdx.x < 3 ? fminf(r, sh[threadIdx.x][w]) : fmaxf(r, sh[threadIdx.x][w]);
        part[6 * blockIdx.x + threadIdx.x] = r;
    }
}

// Per-workgroup partials of X^T P (D x 3) over kGramRows rows: thread t
// takes column t % 64 (+ 64 j) of every fourth row (coalesced row reads), the
// four row phases are added in LDS (fixed order: deterministic).
constexpr int kGramRows = 256;
__global__ __launch_bounds__(256) void gram3(const double* __restrict__ X, int N, int D, const float* __restrict__ P,
                                             float* __restrict__ part)
{
    __shared__ float sh[4][64][3];
    const int r0 = blockIdx.x * kGramRows, r1 = min(N, r0 + kGramRows);
    const int dl = threadIdx.x & 63, ph = threadIdx.x >> 6;
    for (int d0 = 0; d0 < D; d0 += 64) {
        const int d = d0 + dl;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f;
        if (d < D) {
            for (int r = r0 + ph; r < r1; r += 4) {
                const float x = (float)X[(size_t)r * D + d];
                g0 = fmaf(x, P[3 * (size_t)r], g0);
                g1 = fmaf(x, P[3 * (size_t)r + 1], g1);
                g2 = fmaf(x, P[3 * (size_t)r + 2], g2);
            }
        }
        sh[ph][dl][0] = g0;
        sh[ph][dl][1] = g1;
        sh[ph][dl][2] = g2;
        __syncthreads();
        if (ph == 0 && d < D) {
            float* o = part + ((size_t)blockIdx.x * D + d) * 3;
#pragma unroll
            for (int k = 0; k < 3; k++) o[k] = ((sh[0][dl][k] + sh[1][dl][k]) + sh[2][dl][k]) + sh[3][dl][k];
        }
        __syncthreads();
    }
}

// Bounding box of the projections (min x, y, z, max x, y, z) from proj3's
// per-workgroup partials.  One workgroup.
__global__ __launch_bounds__(256) void reduce_lim(const float* __restrict__ part, int nb, float* __restrict__ lim)
{
    __shared__ float sh[6][4];
    float v[6] = {HUGE_VALF, HUGE_VALF, HUGE_VALF, -HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
    for (int b = threadIdx.x; b < nb; b += 256) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], part[6 * b + k]);
            v[k + 3] = fmaxf(v[k + 3], part[6 * b + k + 3]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        for (int o = 32; o > 0; o >>= 1) {
            v[k] = fminf(v[k], __shfl_xor(v[k], o, 64));
            v[k + 3] = fmaxf(v[k + 3], __shfl_xor(v[k + 3], o, 64));
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        float r = sh[threadIdx.x][0];
        for (int w = 1; w < 4; w++) r = threadIdx.x < 3 ? fminf(r, sh[threadIdx.x][w]) : fmaxf(r, sh[threadIdx.x][w]);
        lim[threadIdx.x] = r;
    }
}

// The fixed starting projection: 3 x D values uniform in (-1, 1) from a
// 64-bit LCG (any fixed full-rank projection works).
__global__ void proj_init(float* __restrict__ R, int n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (int e = 0; e < n; e++) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        R[e] = (float)((double)(st >> 11) * 0x1.0p-53 * 2.0 - 1.0);
    }
}

// G = X^T P (D x 3) from gram3's ngr partials; R <- orth(G): Gram-Schmidt on
// the three columns in FP64 (one workgroup); a column that vanishes (data of
// rank < 3) keeps the previous direction.
__global__ __launch_bounds__(256) void gram_sum(const float* __restrict__ gpart, int ngr, int D, double* __restrict__ G)
{
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);  // a wave per entry
    if (e >= 3 * D) return;
    double g = 0.0;
    for (int b = lane; b < ngr; b += 64) g += gpart[(size_t)b * D * 3 + e];
    g = wave_sum(g);
    if (lane == 0) G[e] = g;
}

__global__ __launch_bounds__(256) void orth3(int D, double* __restrict__ G, float* __restrict__ R)
{
    __shared__ double sh[4];
    for (int k = 0; k < 3; k++) {
        for (int j = 0; j < k; j++) {
            double dot = 0.0;
            for (int d = threadIdx.x; d < D; d += 256) dot += G[3 * d + k] * G[3 * d + j];
            dot = block_sum(dot, sh);
            for (int d = threadIdx.x; d < D; d += 256) G[3 * d + k] -= dot * G[3 * d + j];
            __syncthreads();
        }
        double nrm = 0.0;
        for (int d = threadIdx.x; d < D; d += 256) nrm += G[3 * d + k] * G[3 * d + k];
        nrm = sqrt(block_sum(nrm, sh));
        for (int d = threadIdx.x; d < D; d += 256) {
            if (nrm > 1e-30) {
                G[3 * d + k] /= nrm;
                R[3 * d + k] = (float)G[3 * d + k];
            } else {
                G[3 * d + k] = 0.0;  // later columns ignore it
            }
        }
        __syncthreads();
    }
}

TKA_DEV uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton keys (10 bits per projected axis) over the projections'
// bounding box.
__global__ __launch_bounds__(256) void morton3(const float* __restrict__ P, int N, const float* __restrict__ lim,
                                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    uint32_t key = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float span = lim[k + 3] - lim[k];
        float t = span > 0.f ? (P[3 * (size_t)n + k] - lim[k]) / span : 0.f;
        t = fminf(fmaxf(t, 0.f), 1.f);
        key |= spread10((uint32_t)(t * 1023.f)) << k;
    }
    keys[n] = key;
    vals[n] = (uint32_t)n;
}

__global__ void copy_word(const uint32_t* __restrict__ src, int* __restrict__ dst) { *dst = (int)*src; }

__global__ __launch_bounds__(256) void gather_rows(const double* __restrict__ X, const double* __restrict__ n2,
                                                   const uint32_t* __restrict__ perm, int N, int D,
                                                   double* __restrict__ Xp, double* __restrict__ n2p)
{
    const size_t total = (size_t)N * D;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t i = e / D, d = e % D;
        const uint32_t p = perm[i];
        const uint32_t src = p < (uint32_t)N ? p : 0u;  // in bounds whatever the sort left
        Xp[e] = X[(size_t)src * D + d];
        if (d == 0) n2p[i] = n2[src];
    }
}

}  // namespace

// ============================================================== host side
void normalize_device(double* X, int N, int D, MemBudget* mb, hipStream_t s)
{
    DevBuf<double> mean(D, mb), bmax(1024, mb), mx(1, mb);
    hipLaunchKernelGGL(col_mean_seq, dim3(D), dim3(64), 0, s, X, N, D, mean.get());
    TKA_LAUNCH_CHECK();
    const size_t total = (size_t)N * D;
    const int nb = (int)std::min<size_t>(1024, ceil_div(total, 256));
    hipLaunchKernelGGL(center_max, dim3(nb), dim3(256), 0, s, X, total, D, mean.get(), bmax.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_max, dim3(1), dim3(256), 0, s, bmax.get(), nb, mx.get());
    TKA_LAUNCH_CHECK();
    double h = 0.0;
    read_to_host(&h, mx.get(), sizeof h, s);
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

// FP8 rows too wide for the register tiles: LDS-staged tiles.
void run_tiles_fp8(const uint8_t* A8, const float* scale, const float* nrm, int N, int Kp, int kprime, TileRun& tr,
                   MemBudget* mb, hipStream_t s)
{
    const int nslots = (int)ceil_div(N, kQB) * kQB;
    tr.buf.alloc((size_t)nslots * kCap, mb);
    tr.cnt.alloc(nslots, mb);
    tr.tau.alloc(nslots, mb);
    tr.fail.alloc(nslots, mb);
    TileArgs ta{N, kprime, 0, nullptr, N, tr.buf.get(), tr.cnt.get(), tr.tau.get()};
    hipLaunchKernelGGL(knn_tiles_fp8, dim3(ceil_div(N, kQB)), dim3(256), 0, s, A8, scale, nrm, Kp, ta);
    TKA_LAUNCH_CHECK();
}

// All rows against all rows with the query operand in registers.  NV
// (vectors per half row, 1..kNVLimit) is a template constant, so the k-loop
// is fully unrolled and the LDS reads of step g+1 are scheduled under the
// MFMAs of step g.  QT: query tiles per wave (see knn_tiles_reg).
// A (candidates), B (queries): padded half-major rows for ceil(N / kCB) * kCB rows.
template <class Op, int kMode, int NV, int QT, int WAVES>
void launch_tiles_reg(const void* A, const void* B, dim3 grid, size_t lds, const TileArgs& ta, hipStream_t s)
{
    typedef typename Op::V V;
    hipLaunchKernelGGL((knn_tiles_reg<Op, kMode, NV, QT, WAVES>), grid, dim3(64 * WAVES), lds, s, (const V*)A,
                       (const V*)B, ta);
}

template <class Op, int kMode, int QT, int WAVES>
void run_tiles_reg(const void* A, const void* B, int NV, int N, int kprime, const int* qlist, int nq, TileRun& tr,
                   MemBudget* mb, hipStream_t s)
{
    typedef typename Op::V V;
    require(NV >= 1 && NV <= kNVLimit, TKA_E_INTERNAL, "register tiles: bad operand width");
    constexpr int QB = 32 * WAVES * QT;
    const int nslots = (int)ceil_div(nq, QB) * QB;
    tr.buf.alloc((size_t)nslots * kCap, mb);
    tr.cnt.alloc(nslots, mb);
    tr.tau.alloc(nslots, mb);
    tr.fail.alloc(nslots, mb);
    TileArgs ta{N, kprime, 1, qlist, nq, tr.buf.get(), tr.cnt.get(), tr.tau.get()};
    const size_t lds = 2 * (size_t)kCB * (2 * NV + 1) * sizeof(V);
    const dim3 grid(ceil_div(nq, QB));
#define TKA_W(W) \
    case W: launch_tiles_reg<Op, kMode, W, QT, WAVES>(A, B, grid, lds, ta, s); break
    switch (NV) {
        TKA_W(1); TKA_W(2); TKA_W(3); TKA_W(4); TKA_W(5); TKA_W(6); TKA_W(7); TKA_W(8); TKA_W(9); TKA_W(10);
        TKA_W(11); TKA_W(12); TKA_W(13); TKA_W(14); TKA_W(15); TKA_W(16); TKA_W(17); TKA_W(18); TKA_W(19);
        TKA_W(20);
    }
#undef TKA_W
    TKA_LAUNCH_CHECK();
}

// All rows against all rows on the block-centred BF16 tiles (knn_tiles_local):
// Xs N x D FP64 rows in locality order.  Fills the centres (one per block of
// kLocQB query rows, stride Dp) for the re-rank's bound.
void run_tiles_local(const double* Xs, int N, int D, int kprime, TileRun& tr, DevBuf<float>& centres, MemBudget* mb,
                     hipStream_t s)
{
    const int NV = local_nv(D), Dp = local_dp(D);
    require(NV >= 1 && NV <= kLocNVMax, TKA_E_INTERNAL, "block-centred tiles: bad operand width");
    const int Npad = (int)ceil_div(N, kCB) * kCB;
    const int nb = (int)ceil_div(N, kLocQB);
    DevBuf<float> Xf((size_t)Npad * Dp, mb);
    centres.alloc((size_t)nb * Dp, mb);
    hipLaunchKernelGGL(prep_local_rows, dim3(std::min<size_t>(8192, ceil_div((size_t)Npad * Dp, 256))), dim3(256), 0,
                       s, Xs, N, Npad, D, Dp, Xf.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(prep_local_centres, dim3(nb), dim3(256), 0, s, Xs, N, D, Dp, centres.get());
    TKA_LAUNCH_CHECK();
    const int nslots = nb * kLocQB;
    tr.buf.alloc((size_t)nslots * kCap, mb);
    tr.cnt.alloc(nslots, mb);
    tr.tau.alloc(nslots, mb);
    tr.fail.alloc(nslots, mb);
    TileArgs ta{N, kprime, 1, nullptr, N, tr.buf.get(), tr.cnt.get(), tr.tau.get()};
    const size_t lds = 2 * (size_t)kCB * (2 * NV + 1) * sizeof(bf16x8);
#define TKA_L(W)                                                                                                  \
    case W:                                                                                                       \
        hipLaunchKernelGGL(knn_tiles_local<W>, dim3(nb), dim3(64 * kLocWaves), lds, s, Xf.get(), D, centres.get(), \
                           ta);                                                                                   \
        break
    switch (NV) {
        TKA_L(1); TKA_L(2); TKA_L(3); TKA_L(4); TKA_L(5); TKA_L(6); TKA_L(7); TKA_L(8); TKA_L(9); TKA_L(10);
    }
#undef TKA_L
    TKA_LAUNCH_CHECK();
}

// Query rows (qlist[0, nq), or all rows) against all rows with both operands
// streamed in k-stages (knn_tiles_wide).  A, B: padded half-major rows of NV
// vectors per half (NV a multiple of kWideKS) for ceil(N / kCB) * kCB rows.
template <class Op>
void launch_tiles_wide(const void* A, const void* B, int NV, int Npad, int nq, const TileArgs& ta, hipStream_t s)
{
    typedef typename Op::V V;
    const size_t lds = 2 * (size_t)WideStage<kWideKS>::IMG * sizeof(V);
    hipLaunchKernelGGL((knn_tiles_wide<Op, kWideKS>), dim3(ceil_div(nq, kWideQB)), dim3(64 * kWideWaves), lds, s,
                       (const V*)A, (const V*)B, NV, Npad, ta);
}

// A, B: padded half-major rows (NV vectors per half, a multiple of kWideKS)
// for Npad = ceil(N / kCB) * kCB rows; copied to the stage-major layout here.
template <class Op>
void run_tiles_wide(const void* A, const void* B, int NV, int N, int kprime, const int* qlist, int nq, TileRun& tr,
                    MemBudget* mb, hipStream_t s)
{
    require(NV >= kWideKS && NV % kWideKS == 0, TKA_E_INTERNAL, "wide tiles: bad operand width");
    const int Npad = (int)ceil_div(N, kCB) * kCB;
    const size_t nvec = (size_t)Npad * 2 * NV;
    DevBuf<uint4> Aw(nvec, mb), Bw(nvec, mb);
    const unsigned g = (unsigned)std::min<size_t>(8192, ceil_div(nvec, 256));
    hipLaunchKernelGGL(hm_to_stages<kWideKS>, dim3(g), dim3(256), 0, s, (const uint4*)A, Npad, NV, Aw.get());
    TKA_LAUNCH_CHECK();
    hipLaunchKernelGGL(hm_to_stages<kWideKS>, dim3(g), dim3(256), 0, s, (const uint4*)B, Npad, NV, Bw.get());
    TKA_LAUNCH_CHECK();
    const int nslots = (int)ceil_div(nq, kWideQB) * kWideQB;
    tr.buf.alloc((size_t)nslots * kCap, mb);
    tr.cnt.alloc(nslots, mb);
    tr.tau.alloc(nslots, mb);
    tr.fail.alloc(nslots, mb);
    TileArgs ta{N, kprime, 1, qlist, nq, tr.buf.get(), tr.cnt.get(), tr.tau.get()};
    launch_tiles_wide<Op>(Aw.get(), Bw.get(), NV, Npad, nq, ta, s);
    TKA_LAUNCH_CHECK();
}

}  // namespace

// Every distance in FP64, selected exactly per row (X in the caller's
// order, so ties go by the caller's index): for K beyond the candidate
// buffers (all rows), and to re-solve the rows whose certificate failed
// (qlist, nq rows, with ubound[i] >= their K-th distance).  Query rows in
// batches of at most 512 MB of distances.
static void knn_exact_rows(const double* Xn, int N, int D, int K, const int* qlist, int nq, int* idx, double* dist,
                           MemBudget* mb, hipStream_t s, const double* ubound = nullptr)
{
    require(K <= kLargeKMax, TKA_E_WRONG_PARAMETER,
            "K = " + std::to_string(K) + " exceeds the engine's limit of " + std::to_string(kLargeKMax) +
                " neighbours (perplexity < " + std::to_string(kLargeKMax / 3 + 1) + ")");
    const size_t budget = (size_t)512 << 20;
    int B = (int)std::min<size_t>((size_t)nq, std::max<size_t>(kPairTile, budget / ((size_t)N * sizeof(double))));
    if (B < nq) B = B / kPairTile * kPairTile;
    DevBuf<double> buf((size_t)B * N, mb);
    for (int q0 = 0; q0 < nq; q0 += B) {
        const int b = std::min(B, nq - q0);
        hipLaunchKernelGGL(pair_sqdist_rows, dim3(ceil_div(N, kPairTile), ceil_div(b, kPairTile)), dim3(256), 0, s,
                           Xn, N, D, qlist, q0, b, buf.get());
        TKA_LAUNCH_CHECK();
        hipLaunchKernelGGL(select_k_rows, dim3(b), dim3(256), 0, s, (const double*)buf.get(), N, b, K, qlist, q0,
                           ubound, idx, dist);
        TKA_LAUNCH_CHECK();
    }
}

void knn_device(const double* Xn, int N, int D, int K, int dtype, int kprime, int* idx, double* dist, KnnStats* st,
                MemBudget* mb, hipStream_t s, uint32_t* order_out)
{
    require(N >= 2 && K >= 1 && K <= N - 1, TKA_E_WRONG_PARAMETER, "kNN needs 1 <= K <= N-1");
    require(dtype == TKA_DIST_FP32 || dtype == TKA_DIST_BF16 || dtype == TKA_DIST_FP8, TKA_E_UNSUPPORTED,
            "distance dtype " + std::to_string(dtype) + " is not available");
    if (K > kMaxKprime) {  // all distances in FP64 (whatever the tile dtype): exact without a certificate
        knn_exact_rows(Xn, N, D, K, nullptr, N, idx, dist, mb, s);
        return;
    }
    // default candidates per row: 1.5 K for the FP32 and split-BF16 tiles,
    // 5/3 K for the block-centred BF16 tiles (their bound grows with a row's
    // distance to its block's centre), rows the bound cannot certify being
    // re-solved exactly; 2 K for FP8 (not re-solved)
    const bool centred = dtype == TKA_DIST_BF16 && local_nv(D) <= kLocNVMax;
    if (kprime <= 0)
        kprime = dtype == TKA_DIST_FP8 ? 2 * K
                 : centred            ? std::max(K + (2 * K + 2) / 3, K + 16)
                                      : std::max(K + (K + 1) / 2, K + 16);
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
    // status words read once after the re-rank: [0] rows that failed the
    // bound, [1] the locality sort's fault word
    DevBuf<int> status(2, mb);
    TKA_HIP(hipMemsetAsync(status.get(), 0, 2 * sizeof(int), s));

    // Operands.
    DevBuf<__bf16> A, B;
    DevBuf<uint8_t> A8;
    DevBuf<float> scale8, nrm8;
    int Kp = 0;
    const double u32 = std::ldexp(1.0, -24);
    double eps_a = 0.0, eps_b = 0.0;
    TileRun tr;
    const unsigned gprep = (unsigned)std::min<size_t>(8192, ceil_div((size_t)N * 64, 256));
    const int Npad = (int)ceil_div(N, kCB) * kCB;

    // Locality order for the register tiles: rows permuted by the Morton code
    // of a 3-D projection onto their estimated principal subspace (see
    // proj3).  Xs/ns are the rows the tiles and the re-rank read; perm maps
    // them back to the caller's rows.
    DevBuf<uint32_t> perm;
    DevBuf<double> Xp, n2p;
    const double* Xs = Xn;
    const double* ns = nullptr;
    auto locality = [&]() {
        DevBuf<float> Rd(3 * (size_t)D, mb), P((size_t)N * 3, mb), part((size_t)6 * nbr, mb);
        const int ngr = (int)ceil_div(N, kGramRows);
        DevBuf<float> gpart((size_t)ngr * D * 3, mb);
        DevBuf<double> G(3 * (size_t)D, mb);
        hipLaunchKernelGGL(proj_init, dim3(1), dim3(64), 0, s, Rd.get(), 3 * D);
        TKA_LAUNCH_CHECK();
        for (int it = 0; it < 2; it++) {  // R <- orth(X^T X R)
            hipLaunchKernelGGL(proj3, dim3(nbr), dim3(256), 0, s, Xn, N, D, Rd.get(), P.get(), part.get());
            TKA_LAUNCH_CHECK();
            hipLaunchKernelGGL(gram3, dim3(ngr), dim3(256), 0, s, Xn, N, D, P.get(), gpart.get());
            TKA_LAUNCH_CHECK();
            hipLaunchKernelGGL(gram_sum, dim3(ceil_div(3 * D, 4)), dim3(256), 0, s, gpart.get(), ngr, D, G.get());
            TKA_LAUNCH_CHECK();
            hipLaunchKernelGGL(orth3, dim3(1), dim3(256), 0, s, D, G.get(), Rd.get());
            TKA_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(proj3, dim3(nbr), dim3(256), 0, s, Xn, N, D, Rd.get(), P.get(), part.get());
        TKA_LAUNCH_CHECK();
        DevBuf<uint32_t> k0(N, mb), k1(N, mb), v0(N, mb), v1(N, mb), scr(radix_sort_scratch_words(N), mb);
        TKA_HIP(hipMemsetAsync(scr.get(), 0, scr.size() * sizeof(uint32_t), s));
        DevBuf<float> lim(6, mb);
        hipLaunchKernelGGL(reduce_lim, dim3(1), dim3(256), 0, s, part.get(), nbr, lim.get());
        TKA_LAUNCH_CHECK();
        hipLaunchKernelGGL(morton3, dim3(nbr), dim3(256), 0, s, P.get(), N, lim.get(), k0.get(), v0.get());
        TKA_LAUNCH_CHECK();
        const bool alt = radix_sort_pairs<uint32_t, uint32_t>(k0.get(), v0.get(), k1.get(), v1.get(), N, 30, scr.get(), s);
        // the order gathers rows of X: a sort that gave up is reported at the
        // failure count's read (gather_rows keeps every read in bounds)
        hipLaunchKernelGGL(copy_word, dim3(1), dim3(1), 0, s, scr.get() + 1, status.get() + 1);
        TKA_LAUNCH_CHECK();
        perm = std::move(alt ? v1 : v0);
        Xp.alloc((size_t)N * D, mb);
        n2p.alloc(N, mb);
        hipLaunchKernelGGL(gather_rows, dim3(gprep), dim3(256), 0, s, Xn, n2.get(), perm.get(), N, D, Xp.get(),
                           n2p.get());
        TKA_LAUNCH_CHECK();
        Xs = Xp.get();
        ns = n2p.get();
    };
    // FP32 MFMA tiles over the locality-ordered rows for the query rows
    // ql[0, nq) (all rows when ql is null), candidate lists of kp; sets the
    // error bound of the values they produce.  The norm is one more term of
    // the FP32 sum: its rounding is bounded like the products' (D+1 terms
    // plus the two conversions).
    auto fp32_tiles = [&](const int* ql, int nq, int kp, TileRun& t) {
        const int NV = ((D + 2) / 2 + 3) / 4;  // float4 vectors of k-steps (2 k per step) per half row
        const bool reg = reg_tiles_fit(NV, sizeof(float4));
        const int NVo = reg ? NV : wide_nv(NV);
        DevBuf<float> Ah((size_t)Npad * 4 * (2 * NVo + 1), mb), Bh((size_t)Npad * 4 * (2 * NVo + 1), mb);
        hipLaunchKernelGGL(prep_fp32n_hm, dim3(gprep), dim3(256), 0, s, Xs, ns ? ns : n2.get(), N, Npad, D, NVo,
                           Ah.get(), Bh.get());
        TKA_LAUNCH_CHECK();
        if (reg) {
            run_tiles_reg<OpF32, 0, 1, 8>(Ah.get(), Bh.get(), NV, N, kp, ql, nq, t, mb, s);
            eps_a = 2.0 * (2.0 * (D + 3) * u32);
        } else {
            // stages of 8 kWideKS products summed from zero, one addition per stage
            run_tiles_wide<OpF32>(Ah.get(), Bh.get(), NVo, N, kp, ql, nq, t, mb, s);
            eps_a = 2.0 * (2.0 * (8 * kWideKS + NVo / kWideKS + 2) * u32);
        }
        eps_b = eps_a;
    };
    DevBuf<float> centres;  // block-centred BF16 tiles
    if (centred) {
        locality();
        Kp = 16 * local_nv(D);
        run_tiles_local(Xs, N, D, kprime, tr, centres, mb, s);
    } else if (dtype == TKA_DIST_BF16) {
        // rows too wide for the centred tiles' registers: the hi/lo splits on
        // the global scale, both operands streamed in k-stages (NV padded
        // with zero vectors to whole stages)
        Kp = ((3 * D + 2) + 15) / 16 * 16;
        const int NVw = wide_nv(Kp / 16);
        locality();
        const size_t rows = (size_t)Npad * 8 * (2 * NVw + 1);
        A.alloc(rows, mb);
        B.alloc(rows, mb);
        hipLaunchKernelGGL(prep_bf16x3_hm, dim3(gprep), dim3(256), 0, s, Xs, ns, N, Npad, D, NVw, A.get(), B.get());
        TKA_LAUNCH_CHECK();
        run_tiles_wide<OpBF16>(A.get(), B.get(), NVw, N, kprime, nullptr, N, tr, mb, s);
        // each stage's 16 kWideKS products are summed from zero, then added to
        // the running sum: at most that many roundings plus one per stage
        const int nsum = 16 * kWideKS + NVw / kWideKS;
        const double ub2 = std::ldexp(1.0, -16);
        const double gam = nsum * u32;
        eps_a = 2.0 * (6.2 * ub2 + 2.1 * gam);
        eps_b = 2.0 * (1.1 * ub2 + 1.1 * gam);
    } else if (dtype == TKA_DIST_FP8) {
        Kp = (D + 15) / 16 * 16;
        const int NV = Kp / 16;
        const bool reg = reg_tiles_fit(NV, sizeof(uint64_t));
        if (reg) locality();
        A8.alloc((size_t)N * Kp, mb);
        scale8.alloc(N, mb);
        nrm8.alloc(N, mb);
        hipLaunchKernelGGL(prep_fp8, dim3(ceil_div(N, 256)), dim3(256), 0, s, Xs, ns ? ns : n2.get(), N, D, Kp,
                           A8.get(), scale8.get(), nrm8.get());
        TKA_LAUNCH_CHECK();
        if (reg) {
            DevBuf<uint64_t> A8h((size_t)Npad * (2 * NV + 1), mb);
            hipLaunchKernelGGL(fp8_to_hm, dim3(gprep), dim3(256), 0, s, A8.get(), nrm8.get(), scale8.get(), N, Npad,
                               Kp / 16, NV, A8h.get());
            TKA_LAUNCH_CHECK();
            run_tiles_reg<OpFP8, 2, 1, 4>(A8h.get(), A8h.get(), NV, N, kprime, nullptr, N, tr, mb, s);
        } else {
            run_tiles_fp8(A8.get(), scale8.get(), nrm8.get(), N, Kp, kprime, tr, mb, s);
        }
        // e4m3 keeps 4 significant bits (unit roundoff 2^-4): the bound is
        // loose at any useful k'; FP8 rows are not re-solved (approximate mode)
        const double u8 = std::ldexp(1.0, -4);
        eps_a = 2.0 * (2.0 * (2 * u8 + u8 * u8) + 2.0 * D * u32);
        eps_b = 2.0 * (2.0 * u32);
    } else {
        locality();
        fp32_tiles(nullptr, N, kprime, tr);
    }
    if (order_out && perm.get()) {
        TKA_HIP(hipMemcpyAsync(order_out, perm.get(), (size_t)N * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        if (st) st->order_written = true;
    }
    DevBuf<double> ubound(N, mb), ulist(N, mb);
    RerankArgs ra{Xs, ns ? ns : n2.get(), N, D, K, nullptr, N, tr.buf.get(), tr.cnt.get(), tr.tau.get(),
                  eps_a, eps_b, rmaxd.get(), idx, dist, tr.fail.get(), ubound.get(), perm.get(),
                  centres.get(), local_dp(D), Kp};
    hipLaunchKernelGGL(knn_rerank, dim3(ceil_div(N, 4)), dim3(256), 0, s, ra);
    TKA_LAUNCH_CHECK();

    // Rows whose bound failed: every distance of those rows in FP64 (exact,
    // no certificate needed), with X in the caller's order so that ties go
    // by the caller's index.  Their cost is the rows' share of the N^2 D
    // work, where a tile pass costs a whole stream of the candidates.
    DevBuf<int> list(N, mb);
    hipLaunchKernelGGL(collect_failures, dim3(ceil_div(N, 256)), dim3(256), 0, s, tr.fail.get(), ubound.get(), N,
                       perm.get(), list.get(), ulist.get(), status.get());
    TKA_LAUNCH_CHECK();
    int hs[2] = {0, 0};
    read_to_host(hs, status.get(), sizeof hs, s);
    require(hs[1] == 0, TKA_E_INTERNAL, "radix sort look-back gave up");
    const int hf = hs[0];
    if (st) st->certificate_failures = hf;
    if (st) st->unresolved = dtype == TKA_DIST_FP8 ? hf : 0;
    if (hf > 0 && dtype != TKA_DIST_FP8) knn_exact_rows(Xn, N, D, K, list.get(), hf, idx, dist, mb, s, ulist.get());
}

}  // namespace tka
