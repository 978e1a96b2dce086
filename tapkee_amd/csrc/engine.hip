// Host engine and C-ABI entry points (include/tapkee_amd.h).
//
// tka_tsne_run reproduces the schedule of tsne::TSNE::run
// (tsne.hpp:61-205): normalise X, build P (kNN + perplexity + symmetrise +
// normalise at theta > 0, dense at theta == 0), exaggerate by 12, draw Y0
// from std::rand, then 1000 iterations of gradient + gains/momentum update +
// zeroMean, switching off the exaggeration and raising the momentum after
// iteration 250, with a KL evaluation every 50 iterations and at the last.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "affinity.hpp"
#include "bh.hpp"
#include "bh3.hpp"
#include "common.hpp"
#include "forces.hpp"
#include "knn.hpp"
#include "radix_sort.hpp"

namespace tka {

namespace {
struct CachedBlock {
    void* p;
    hipEvent_t ready;  // recorded at the free: the last work that may use p
};
std::mutex g_cache_mu;
std::map<std::pair<int, size_t>, std::vector<CachedBlock>> g_cache;  // (device, block bytes) -> free blocks
size_t g_cache_bytes = 0;
constexpr size_t kCacheLimit = (size_t)32 << 30;  // cached bytes beyond this are returned with hipFree
// Events of reused blocks, kept for the next frees on the same device
// (creating an event per free costs more host time than recording one).
std::map<int, std::vector<hipEvent_t>> g_free_events;

// Size classes: powers of two up to 1 MiB, whole 2 MiB pages above.
size_t size_class(size_t b)
{
    if (b <= ((size_t)1 << 20)) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t page = (size_t)2 << 20;
    return (b + page - 1) / page * page;
}

// Returns every cached block of the device (after its last use) to the runtime.
void trim_cache(int dev)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto it = g_cache.begin(); it != g_cache.end();) {
        if (it->first.first != dev) {
            ++it;
            continue;
        }
        for (auto& b : it->second) {
            (void)hipEventSynchronize(b.ready);
            (void)hipEventDestroy(b.ready);
            (void)hipFree(b.p);
            g_cache_bytes -= it->first.second;
        }
        it = g_cache.erase(it);
    }
}
}  // namespace

std::string& last_error()
{
    thread_local std::string msg;
    return msg;
}

namespace {
std::mutex g_log_mu;
tka_log_fn g_log_fn = nullptr;
void* g_log_user = nullptr;
uint32_t g_log_levels = TKA_LOG_DEFAULT_LEVELS;
const char* const kLevelNames[] = {"info", "warning", "debug", "error", "benchmark"};
}  // namespace

void log_message(tka_log_level level, const std::string& msg, bool force)
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    if (!force && !(g_log_levels & (1u << level))) return;
    if (g_log_fn)
        g_log_fn((int32_t)level, msg.c_str(), g_log_user);
    else
        fprintf(stderr, "[%s] %s\n", kLevelNames[level], msg.c_str());
}

void log_took(const char* operation, double seconds, bool force)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.9g", seconds);
    log_message(TKA_LOG_BENCHMARK, std::string(operation) + " took " + buf + " seconds.", force);
}

int device_cus()
{
    static int cus[64] = {};  // per device, filled on first use
    int dev = 0;
    TKA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    int v = __atomic_load_n(&cus[dev], __ATOMIC_RELAXED);
    if (v == 0) {
        TKA_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        if (v <= 0) v = 1;
        __atomic_store_n(&cus[dev], v, __ATOMIC_RELAXED);
    }
    return v;
}

void* dev_alloc(size_t bytes, size_t* block_bytes, int* device)
{
    const size_t bb = size_class(bytes);
    int dev = 0;
    TKA_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache.find({dev, bb});
        if (it != g_cache.end()) {
            auto& v = it->second;
            for (size_t i = 0; i < v.size(); i++) {
                if (hipEventQuery(v[i].ready) != hipSuccess) continue;  // still in use by queued work
                void* p = v[i].p;
                g_free_events[dev].push_back(v[i].ready);
                v.erase(v.begin() + (long)i);
                g_cache_bytes -= bb;
                *block_bytes = bb;
                *device = dev;
                return p;
            }
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bb);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        trim_cache(dev);
        e = hipMalloc(&p, bb);
    }
    TKA_HIP(e);
    *block_bytes = bb;
    *device = dev;
    return p;
}

void read_to_host(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    static thread_local void* pinned = nullptr;  // held until the process ends
    require(bytes <= 256, TKA_E_INTERNAL, "read_to_host: at most 256 bytes");
    if (!pinned) TKA_HIP(hipHostMalloc(&pinned, 256, hipHostMallocDefault));
    TKA_HIP(hipMemcpyAsync(pinned, src, bytes, hipMemcpyDeviceToHost, s));
    TKA_HIP(hipStreamSynchronize(s));
    std::memcpy(dst, pinned, bytes);
}

void dev_free(void* p, size_t bb, int dev, hipStream_t stream) noexcept
{
    hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto& pool = g_free_events[dev];
        if (!pool.empty()) {
            ev = pool.back();
            pool.pop_back();
        }
    }
    if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
    if (ev && hipEventRecord(ev, stream) == hipSuccess) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        if (g_cache_bytes + bb <= kCacheLimit) {
            g_cache[{dev, bb}].push_back(CachedBlock{p, ev});
            g_cache_bytes += bb;
            return;
        }
    }
    if (ev) {
        (void)hipEventSynchronize(ev);
        (void)hipEventDestroy(ev);
    }
    (void)hipFree(p);
}

namespace {


double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Stream {
    hipStream_t s = nullptr;
    bool owned = false;
    explicit Stream(void* ext = nullptr)
    {
        if (ext) {
            s = (hipStream_t)ext;
        } else {
            TKA_HIP(hipStreamCreate(&s));
            owned = true;
        }
    }
    ~Stream()
    {
        if (owned && s) (void)hipStreamDestroy(s);
    }
    void sync() { TKA_HIP(hipStreamSynchronize(s)); }
};

template <class T>
void upload(DevBuf<T>& d, const T* h, size_t n, hipStream_t s, MemBudget* mb = nullptr)
{
    d.alloc(n, mb);
    if (n) TKA_HIP(hipMemcpyAsync(d.get(), h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

template <class T>
void download(T* h, const T* d, size_t n, hipStream_t s)
{
    if (n) TKA_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

__global__ void cast_kernel_dd(const double* __restrict__ a, double* __restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}
__global__ void cast_kernel_df(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = (float)a[i];
}
__global__ void cast_kernel_fd(const float* __restrict__ a, double* __restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = (double)a[i];
}

template <class A, class B>
void cast_device(const A* a, B* b, size_t n, hipStream_t s)
{
    const unsigned g = (unsigned)std::min<size_t>(8192, ceil_div(n ? n : 1, 256));
    if constexpr (std::is_same<A, double>::value && std::is_same<B, float>::value)
        hipLaunchKernelGGL(cast_kernel_df, dim3(g), dim3(256), 0, s, a, b, n);
    else if constexpr (std::is_same<A, float>::value && std::is_same<B, double>::value)
        hipLaunchKernelGGL(cast_kernel_fd, dim3(g), dim3(256), 0, s, a, b, n);
    else
        hipLaunchKernelGGL(cast_kernel_dd, dim3(g), dim3(256), 0, s, (const double*)a, (double*)b, n);
    TKA_LAUNCH_CHECK();
}

__global__ void fill_kernel(double* __restrict__ a, size_t n, double v)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) a[i] = v;
}
__global__ void fill_kernel_f(float* __restrict__ a, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) a[i] = v;
}
template <class T>
void fill_device(T* a, size_t n, double v, hipStream_t s)
{
    const unsigned g = (unsigned)std::min<size_t>(8192, ceil_div(n ? n : 1, 256));
    if constexpr (std::is_same<T, float>::value)
        hipLaunchKernelGGL(fill_kernel_f, dim3(g), dim3(256), 0, s, a, n, (float)v);
    else
        hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, a, n, v);
    TKA_LAUNCH_CHECK();
}

void check_finite(const double* X, size_t n, const char* what)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(X[i])) fail(TKA_E_WRONG_PARAMETER, std::string(what) + " contains a non-finite value");
}

// std::rand replay of the reference's Y0 (tsne.hpp:137-138 with
// defines/random.hpp:38-54), after the VP-tree build's draws
// (vptree.hpp:226: one draw per node with more than one item).
long long vptree_draws(long long n)
{
    long long calls = 0;
    std::vector<long long> st{n};
    while (!st.empty()) {
        const long long m = st.back();
        st.pop_back();
        if (m <= 1) continue;
        calls++;
        st.push_back(m / 2 - 1);
        st.push_back(m - m / 2);
    }
    return calls;
}

double gaussian_random()
{
    double x, y, r;
    do {
        x = 2 * (std::rand() / ((double)RAND_MAX + 1)) - 1;
        y = 2 * (std::rand() / ((double)RAND_MAX + 1)) - 1;
        r = x * x + y * y;
    } while (r >= 1.0 || r == 0.0);
    r = std::sqrt(-2 * std::log(r) / r);
    return x * r;
}

void validate(int N, int D, int no_dims, double perplexity, double theta)
{
    // methods/base.hpp:40-43, methods/tsne.hpp:20-21
    if (N <= 0) fail(TKA_E_NO_DATA, "No data to embed");
    if (!(no_dims >= 1 && no_dims < N))
        fail(TKA_E_WRONG_PARAMETER, "[1, " + std::to_string(N) + ") range check failed for target dimension, its value is " +
                                        std::to_string(no_dims));
    if (!(perplexity >= 0.0 && perplexity <= (N - 1) / 3.0))
        fail(TKA_E_WRONG_PARAMETER, "[0, " + std::to_string((N - 1) / 3.0) +
                                        "] range check failed for SNE perplexity, its value is " +
                                        std::to_string(perplexity));
    if (!(theta >= 0.0)) fail(TKA_E_WRONG_PARAMETER, "Non-negativity check failed for SNE theta");
    if (D <= 0) fail(TKA_E_WRONG_PARAMETER, "feature dimension must be positive");
    // Engine limits (documented deviations): Barnes-Hut runs on a quadtree
    // (2-D, the reference's) or an octree (3-D; the reference's quadtree
    // silently assumes 2-D, quadtree.hpp:71), the exact path supports 1..3
    // dims, and K = (int)(3*perplexity) must be >= 1.
    if (theta > 0.0 && no_dims != 2 && no_dims != 3)
        fail(TKA_E_WRONG_PARAMETER, "Barnes-Hut t-SNE (theta > 0) needs target dimension 2 or 3");
    if (theta > 0.0 && N > kMaxTreePoints)
        fail(TKA_E_UNSUPPORTED, "Barnes-Hut t-SNE: at most 2^26 - 1 points (tree node counts are 26-bit)");
    if (no_dims > 3) fail(TKA_E_WRONG_PARAMETER, "target dimension above 3 is not supported");
}

template <class T>
struct LoopState {
    DevBuf<T> Ya, Yb, uY, gains, neg, pos;
    DevBuf<T> work;  // theta == 0: per-slice row sums of the all-pairs forces
    DevBuf<double> zpart, box;
    DevBuf<int> row, col;
    DevBuf<T> val;   // CSR values (sparse) or dense P (N x N)
    bool dense = false;
};

// Accumulates HIP-event timings of one kernel class on the session stream.
// The events carry no system-scope fence (they only time device work), so
// recording them adds no cache writeback between the kernels they bracket.
struct EventTimer {
    struct Span {
        hipEvent_t a, b;
        int counts;  // launches that end with the span
    };
    std::vector<Span> ev;
    hipEvent_t pending = nullptr;
    double ms = 0.0;
    int64_t launches = 0;
    void begin(hipStream_t s)
    {
        TKA_HIP(hipEventCreateWithFlags(&pending, hipEventDisableSystemFence));
        TKA_HIP(hipEventRecord(pending, s));
    }
    // `launches` launches end here (0: a part of one, which continues at the
    // next begin())
    void end(hipStream_t s, int launches_done = 1)
    {
        hipEvent_t e;
        TKA_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        TKA_HIP(hipEventRecord(e, s));
        ev.push_back({pending, e, launches_done});
        pending = nullptr;
    }
    void pause(hipStream_t s) { end(s, 0); }
    void collect()
    {
        for (auto& p : ev) {
            float t = 0.f;
            TKA_HIP(hipEventElapsedTime(&t, p.a, p.b));
            ms += t;
            launches += p.counts;
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        ev.clear();
    }
    ~EventTimer()
    {
        for (auto& p : ev) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
};

enum { kTimBuild = 0, kTimRepulsion = 1, kTimStep = 2, kTimExact = 3, kTimAttract = 4, kTimCount = 5 };
// Profiled steps run the traversal this many times back to back (it rewrites
// the same forces) between one pair of events, so that the events' own cost
// is spread over the launches.
constexpr int kProfileRepeats = 4;

}  // namespace
}  // namespace tka

// Opaque handle of the C-ABI: one embedding problem with its device state.
struct tka_session {
    tka_tsne_options o;
    tka_tsne_stats st{};
    tka::MemBudget mb;
    tka::Stream stream;
    int N = 0, dims = 0;
    double theta = 0.0;
    bool exact = false;
    int dd_mode = TKA_DD_TEXTBOOK;
    int iter = 0;
    double exag = 12.0, momentum = 0.5;
    bool profile = false;
    tka::EventTimer timers[tka::kTimCount];
    tka::DevBuf<unsigned long long> interactions;
    virtual ~tka_session() = default;
    virtual void step(int n) = 0;
    virtual void embedding(double* Y) = 0;
    virtual double kl_exact() = 0;
    virtual void inject_sort_fault() {}
    hipStream_t s() const { return stream.s; }
};

namespace tka {
namespace {

void log_kl(tka_session& c, int iter, double C)
{
    tka_tsne_stats& st = c.st;
    if (st.kl_count < TKA_MAX_KL_TRACE) {
        st.kl_iter[st.kl_count] = iter;
        st.kl_value[st.kl_count] = C;
        st.kl_count++;
    }
    char buf[96];
    snprintf(buf, sizeof buf, "Iteration %d: error is %.17g", iter, C);  // tsne.hpp:191
    log_message(TKA_LOG_INFO, buf, c.o.verbose != 0);
}

template <class T>
struct Session : tka_session {
    LoopState<T> L;
    BhTree<T> tree;    // 2-D
    BhTree3<T> tree3;  // 3-D
    T* Ycur = nullptr;
    T* Ynext = nullptr;
    int nbox = 0, nz = 0;
    bool tree_ready = false;  // tree, neg and zpart already computed for Ycur
    bool centred = false;     // Ycur is zero-mean (else its mean is in the box partials)

    // Points are periodically renumbered in the Morton order of the latest
    // tree so that CSR neighbours and tree leaves sit close in memory; orig
    // maps the current numbering back to the caller's.
    static constexpr int kRenumberEvery = 50;
    bool renumbered = false;
    DevBuf<int> orig, orig_tmp, inv, row2, col2;
    DevBuf<T> val2, state_tmp;
    DevBuf<uint32_t> rn_scratch;

    // Barnes-Hut loop: one stream.  The attraction rows ride in extra
    // workgroups of the tree build's sort and node launches, which leave most
    // CUs idle; no second stream, so the overlap does not depend on how the
    // runtime maps streams onto hardware queues (two streams that share a
    // queue run one after the other).
    template <int DIMS = 2>
    AttractSlice<T, DIMS> attract_job() const
    {
        AttractSlice<T, DIMS> a;
        a.Y = Ycur;
        a.row_ptr = L.row.get();
        a.col = L.col.get();
        a.val = L.val.get();
        a.exag = (T)exag;
        a.pos = L.pos.get();
        a.row0 = 0;
        a.row1 = N;
        return a;
    }

    // The attraction as a launch of its own (after a KL evaluation, whose
    // tree the step reuses, and in profiling mode, where it is timed alone).
    void attraction_alone()
    {
        hipStream_t s = this->s();
        if (profile) timers[kTimAttract].begin(s);
        attraction_device<T>(dims, Ycur, L.row.get(), L.col.get(), L.val.get(), exag, N, L.pos.get(), s);
        if (profile) timers[kTimAttract].end(s);
    }

    // The sort fault word of the session's tree (quadtree or octree).
    const uint32_t* tree_fault() const { return dims == 3 ? tree3.sort_fault() : tree.sort_fault(); }

    // `fault`: the sort fault word of the tree whose order `perm` is
    void renumber(const uint32_t* perm, const uint32_t* fault)
    {
        hipStream_t s = this->s();
        MemBudget* m = &mb;
        const size_t nd = (size_t)N * dims;
        if (!renumbered) {
            orig_tmp.alloc(N, m);
            inv.alloc(N, m);
            row2.alloc((size_t)N + 1, m);
            col2.alloc(L.col.size(), m);
            val2.alloc(L.val.size(), m);
            state_tmp.alloc(nd, m);
            rn_scratch.alloc(renumber_scan_scratch_words(N), m);
            renumbered = true;
        }
        T* state[3] = {Ycur, L.uY.get(), L.gains.get()};
        renumber_points_device<T>(N, dims, perm, fault, state, 3, state_tmp.get(), L.row.get(), L.col.get(), L.val.get(),
                                  row2.get(), col2.get(), val2.get(), orig.get(), orig_tmp.get(), inv.get(),
                                  rn_scratch.get(), s);
        std::swap(L.row, row2);
        std::swap(L.col, col2);
        std::swap(L.val, val2);
    }

    void init(const std::vector<double>& Y0)
    {
        hipStream_t s = this->s();
        MemBudget* m = &mb;
        const size_t nd = (size_t)N * dims;
        std::vector<T> h(nd);
        for (size_t i = 0; i < nd; i++) h[i] = (T)Y0[i];
        upload(L.Ya, h.data(), nd, s, m);
        TKA_HIP(hipStreamSynchronize(s));
        L.Yb.alloc(nd, m);
        L.uY.alloc(nd, m);
        L.gains.alloc(nd, m);
        L.neg.alloc(nd, m);
        fill_device(L.uY.get(), nd, 0.0, s);
        fill_device(L.gains.get(), nd, 1.0, s);
        const int nstep = tsne_step_blocks(N);
        const int nbox0 = bbox_blocks(N);
        L.box.alloc((size_t)kBoxStride * std::max(nstep, nbox0), m);
        nz = exact ? exact_zblocks(N) : bh_zblocks(N);
        L.zpart.alloc(nz, m);
        if (L.dense || !exact) L.pos.alloc(nd, m);
        if (exact) L.work.alloc(exact_work_elems(N, dims, L.dense), m);
        // tsne.hpp:75-80: uY = 0, gains = 1; Y0 is not centred before iteration 0
        bbox_partials_device<T>(L.Ya.get(), N, L.box.get(), s);
        if (!exact) {
            if (dims == 3)
                tree3.alloc(N, m);
            else
                tree.alloc(N, m);
        }
        nbox = nbox0;
        Ycur = L.Ya.get();
        Ynext = L.Yb.get();
        exag = o.exaggeration;
        momentum = o.momentum;
        // [0]: interactions counted over profiled traversals; [1 + b]: traversal workgroup b's, last launch
        interactions.alloc(1 + (size_t)bh_zblocks(N), m);
        TKA_HIP(hipMemsetAsync(interactions.get(), 0, sizeof(unsigned long long), s));
        std::vector<int> id(N);
        for (int i = 0; i < N; i++) id[i] = i;
        upload(orig, id.data(), (size_t)N, s, m);
        TKA_HIP(hipStreamSynchronize(s));
    }

    void step(int n) override
    {
        hipStream_t s = this->s();
        MemBudget* m = &mb;
        for (int k = 0; k < n; k++, iter++) {
            if (exact) {
                if (profile) timers[kTimExact].begin(s);
                exact_forces_device<T>(dims, Ycur, N, dd_mode, L.dense ? L.val.get() : nullptr, exag, L.neg.get(),
                                       L.dense ? L.pos.get() : nullptr, L.zpart.get(), L.work.get(), s);
                if (profile) timers[kTimExact].end(s);
            } else if (!tree_ready && dims == 3) {
                // octree: its own box partials; the attraction rows (on the Y
                // the key kernel centres) ride in the node kernels, or, when
                // profiling, run alone after the build
                if (iter > 0 && iter % kRenumberEvery == 0 && tree3.sorted_perm != nullptr)
                    renumber(tree3.sorted_perm, tree3.sort_fault());  // order of the previous iteration's octree
                const AttractSlice<T, 3> job3 = attract_job<3>();
                if (profile) timers[kTimBuild].begin(s);
                bh3_build<T>(Ycur, N, tree3, s, profile ? nullptr : &job3);
                if (profile) timers[kTimBuild].end(s);
                if (profile) attraction_alone();
                if (profile) timers[kTimRepulsion].begin(s);
                int g3 = 0;
                for (int r = 0; r < (profile ? kProfileRepeats : 1); r++)
                    g3 = bh3_repulsion<T>(tree3, theta, L.neg.get(), L.zpart.get(),
                                          profile ? interactions.get() : nullptr, s);
                if (profile) {
                    timers[kTimRepulsion].end(s, kProfileRepeats);
                    bh_add_interactions(interactions.get(), g3, kProfileRepeats, s);
                }
            } else if (!tree_ready) {
                if (iter > 0 && iter % kRenumberEvery == 0 && tree.sorted_perm != nullptr)
                    renumber(tree.sorted_perm, tree.sort_fault());  // order of the previous iteration's tree
                // Y is final after the key kernel: the (HBM-bound) attraction
                // runs inside the sort and node launches, which occupy few
                // CUs; beside the traversal it would slow the traversal's
                // node loads by about its own duration.  Profiling times the
                // build without it and the attraction alone between the key
                // kernel and the rest of the build, so that the traversal
                // still follows the node kernels as in the loop.
                const AttractSlice<T> job = attract_job();
                if (profile) timers[kTimBuild].begin(s);
                bh_build_keys<T>(Ycur, N, L.box.get(), nbox, tree, s);
                if (profile) {
                    timers[kTimBuild].pause(s);
                    attraction_alone();
                    timers[kTimBuild].begin(s);
                }
                bh_build_tree<T>(Ycur, N, tree, s, profile ? nullptr : &job);
                if (profile) timers[kTimBuild].end(s);
                if (profile) timers[kTimRepulsion].begin(s);
                int g2 = 0;
                for (int r = 0; r < (profile ? kProfileRepeats : 1); r++)
                    g2 = bh_repulsion<T>(tree, Ycur, theta, L.neg.get(), L.zpart.get(),
                                         profile ? interactions.get() : nullptr, s);
                if (profile) {
                    timers[kTimRepulsion].end(s, kProfileRepeats);
                    bh_add_interactions(interactions.get(), g2, kProfileRepeats, s);
                }
            } else {
                attraction_alone();  // tree and repulsion came with the last KL evaluation
            }
            tree_ready = false;
            // after renumbering: the CSR arrays were swapped for their renumbered copies
            const int* row = L.dense ? nullptr : L.row.get();
            const int* col = L.dense ? nullptr : L.col.get();
            StepParams sp{momentum, o.eta, exag};
            if (exact) {
                if (profile) timers[kTimStep].begin(s);
                nbox = tsne_step_device<T>(dims, Ycur, Ynext, L.uY.get(), L.gains.get(), L.neg.get(),
                                           L.dense ? L.pos.get() : nullptr, row, col,
                                           L.dense ? nullptr : L.val.get(), L.zpart.get(), nz, N, sp, L.box.get(),
                                           nullptr, s);
                if (profile) timers[kTimStep].end(s);
            } else {
                if (profile) timers[kTimStep].begin(s);
                nbox = tsne_update_device<T>(dims, Ycur, Ynext, L.uY.get(), L.gains.get(), L.neg.get(), L.pos.get(),
                                             L.zpart.get(), nz, N, sp, L.box.get(), nullptr, s);
                if (profile) timers[kTimStep].end(s);
            }
            std::swap(Ycur, Ynext);
            centred = false;
            if (exact) {
                center_device<T>(dims, Ycur, N, L.box.get(), nbox, s);  // tsne.hpp:167
                centred = true;
            }
            if (iter == o.stop_lying_iter) exag = 1.0;                   // tsne.hpp:170-179
            if (iter == o.mom_switch_iter) momentum = o.final_momentum;  // tsne.hpp:180-181
            const bool log_now =
                o.log_every > 0 && iter > 0 && ((iter % o.log_every == 0) || (iter == o.max_iter - 1));
            if (!log_now) continue;
            double C;
            if (exact) {
                if (L.dense) {
                    C = kl_dense_device<T>(dims, Ycur, L.val.get(), exag, N, dd_mode, m, s);
                } else {
                    const double Z = exact_z_device<T>(dims, Ycur, N, m, s);
                    C = kl_edges_exact_device<T>(dims, Ycur, row, col, L.val.get(), exag, Z, N, m, s);
                }
            } else {
                // tsne.hpp:454-486: Barnes-Hut estimate of Z from a tree on the
                // updated, zero-mean Y, then the sum over P's pattern.  The
                // next iteration reuses this tree and these forces.
                if (dims == 3) {
                    bh3_build<T>(Ycur, N, tree3, s);
                    bh3_repulsion<T>(tree3, theta, L.neg.get(), L.zpart.get(), nullptr, s);
                } else {
                    bh_build<T>(Ycur, N, L.box.get(), nbox, tree, s);
                    bh_repulsion<T>(tree, Ycur, theta, L.neg.get(), L.zpart.get(), nullptr, s);
                }
                tree_ready = true;
                centred = true;
                if (o.kl_mode == TKA_KL_EXACT_Z) {
                    const double Z = exact_z_device<T>(dims, Ycur, N, m, s);
                    C = kl_edges_exact_device<T>(dims, Ycur, row, col, L.val.get(), exag, Z, N, m, s);
                } else {
                    const double Z = sum_device(L.zpart.get(), nz, m, s);
                    C = kl_edges_reference_device<T>(dims, Ycur, row, col, L.val.get(), exag, Z, N, m, s);
                }
            }
            log_kl(*this, iter, C);
        }
    }

    // The current solution in FP64, made zero-mean in FP64 (tsne.hpp:167).
    DevBuf<double> centred_fp64()
    {
        hipStream_t s = this->s();
        const size_t nd = (size_t)N * dims;
        DevBuf<double> Yd(nd, &mb);
        cast_device<T, double>(Ycur, Yd.get(), nd, s);
        center_fresh_device<double>(dims, Yd.get(), N, &mb, s);
        return Yd;
    }

    // The Morton sort's tile look-back is bounded; a sort that gave up has
    // left a fault word (never expected: every wait is on a running tile).
    void check_sort()
    {
        if (!exact && tree.N > 0 && bh_sort_failed(tree, s()))
            fail(TKA_E_INTERNAL, "radix sort look-back gave up; the tree of some iteration was wrong");
        if (!exact && tree3.N > 0 && bh3_sort_failed(tree3, s()))
            fail(TKA_E_INTERNAL, "radix sort look-back gave up; the tree of some iteration was wrong");
    }

    void inject_sort_fault() override
    {
        if (!exact && tree.N > 0) bh_inject_sort_fault(tree, s());
        if (!exact && tree3.N > 0) bh3_inject_sort_fault(tree3, s());
    }

    void embedding(double* Y) override
    {
        check_sort();
        DevBuf<double> Yd = centred_fp64();
        if (renumbered) {
            DevBuf<double> out((size_t)N * dims, &mb);
            unpermute_rows_device<double>(N, dims, Yd.get(), orig.get(), out.get(), s());
            Yd = std::move(out);
        }
        download(Y, Yd.get(), (size_t)N * dims, s());
        TKA_HIP(hipStreamSynchronize(s()));
    }

    // KL(P||Q) of the current solution (the embedding() values) with the
    // exact normaliser, FP64, P without exaggeration.
    double kl_exact() override
    {
        hipStream_t s = this->s();
        MemBudget* m = &mb;
        check_sort();
        DevBuf<double> Yd = centred_fp64();
        if (L.dense) {
            DevBuf<double> Pd((size_t)N * N, m);
            cast_device<T, double>(L.val.get(), Pd.get(), (size_t)N * N, s);
            return kl_dense_device<double>(dims, Yd.get(), Pd.get(), 1.0, N, TKA_DD_TEXTBOOK, m, s);
        }
        DevBuf<double> vd(L.val.size(), m);
        cast_device<T, double>(L.val.get(), vd.get(), L.val.size(), s);
        const double Z = exact_z_device<double>(dims, Yd.get(), N, m, s);
        return kl_edges_exact_device<double>(dims, Yd.get(), L.row.get(), L.col.get(), vd.get(), 1.0, Z, N, m, s);
    }
};

// What a session's P is computed from: feature vectors (the reference's
// path), a precomputed neighbour graph (indices and squared distances), or
// a precomputed joint P.
struct SimilarityInput {
    enum Kind { kFeatures, kKnn, kJointP } kind = kFeatures;
    const double* X = nullptr;  // kFeatures: N x D
    int D = 0;
    const int* idx = nullptr;   // kKnn: N x K neighbours (no self), K
    const double* dist = nullptr;
    int K = 0;
    const int* row = nullptr;   // kJointP: CSR
    const int* col = nullptr;
    const double* val = nullptr;
};

// Host checks of a precomputed graph before any device work.
void validate_graph(const SimilarityInput& in, int N)
{
    if (in.kind == SimilarityInput::kKnn) {
        require(in.idx && in.dist, TKA_E_WRONG_PARAMETER, "null neighbour arrays");
        require(in.K >= 1 && in.K <= N - 1, TKA_E_WRONG_PARAMETER, "K must be in [1, N-1]");
        for (int i = 0; i < N; i++)
            for (int k = 0; k < in.K; k++) {
                const size_t e = (size_t)i * in.K + k;
                const int j = in.idx[e];
                require(j >= 0 && j < N && j != i, TKA_E_WRONG_PARAMETER,
                        "neighbour " + std::to_string(k) + " of object " + std::to_string(i) +
                            " is out of range or the object itself");
                require(std::isfinite(in.dist[e]) && in.dist[e] >= 0.0, TKA_E_WRONG_PARAMETER,
                        "neighbour distances must be finite and non-negative");
            }
        // a neighbour listed twice in one row has no meaning in the union pattern E u E^T
        std::vector<int> row(in.K);
        for (int i = 0; i < N; i++) {
            std::copy(in.idx + (size_t)i * in.K, in.idx + (size_t)(i + 1) * in.K, row.begin());
            std::sort(row.begin(), row.end());
            const auto dup = std::adjacent_find(row.begin(), row.end());
            if (dup != row.end())
                fail(TKA_E_WRONG_PARAMETER, "object " + std::to_string(*dup) +
                                                " is listed twice among the neighbours of object " + std::to_string(i));
        }
    } else if (in.kind == SimilarityInput::kJointP) {
        require(in.row && in.col && in.val, TKA_E_WRONG_PARAMETER, "null CSR arrays");
        require(in.row[0] == 0, TKA_E_WRONG_PARAMETER, "row_ptr[0] must be 0");
        double sum = 0.0;
        for (int i = 0; i < N; i++) {
            require(in.row[i + 1] >= in.row[i], TKA_E_WRONG_PARAMETER, "row_ptr must be non-decreasing");
            for (int e = in.row[i]; e < in.row[i + 1]; e++) {
                require(in.col[e] >= 0 && in.col[e] < N, TKA_E_WRONG_PARAMETER, "P column index out of range");
                require(std::isfinite(in.val[e]) && in.val[e] >= 0.0, TKA_E_WRONG_PARAMETER,
                        "P values must be finite and non-negative");
                sum += in.val[e];
            }
        }
        require(sum > 0.0, TKA_E_WRONG_PARAMETER, "P sums to zero");
    }
}

// Validation, normalisation, P and Y0: everything tsne.hpp:61-139 does
// before the main loop.  The session's stream is created on o.device.
tka_session* create_session(const SimilarityInput& in, int N, int no_dims, double perplexity, double theta,
                            const tka_tsne_options* opt_in)
{
    tka_tsne_options o;
    tka_default_options(&o);
    if (opt_in) {
        require(opt_in->struct_size == (int32_t)sizeof(tka_tsne_options), TKA_E_WRONG_PARAMETER,
                "tka_tsne_options.struct_size mismatch");
        o = *opt_in;
    }
    const bool features = in.kind == SimilarityInput::kFeatures;
    const int D = features ? in.D : 1;
    validate(N, D, no_dims, features || in.kind == SimilarityInput::kKnn ? perplexity : 0.0, theta);
    if (features) {
        require(in.X != nullptr, TKA_E_WRONG_PARAMETER, "null data pointer");
        check_finite(in.X, (size_t)N * D, "input data");
    }
    validate_graph(in, N);
    if (o.cancel && o.cancel()) fail(TKA_E_CANCELLED, "cancelled");  // methods.hpp:30-31
    const bool exact = (theta == 0.0);
    // a precomputed graph gives a sparse P also at theta = 0
    const bool sparse = !exact || o.exact_p_mode == TKA_EXACT_P_KNN || !features;
    const int K = in.kind == SimilarityInput::kKnn ? in.K : (o.knn_k > 0 ? o.knn_k : (int)(3 * perplexity));
    if (sparse && in.kind != SimilarityInput::kJointP) {
        require(K >= 1, TKA_E_WRONG_PARAMETER, "perplexity must be at least 1/3 (K = 3*perplexity neighbours)");
        require(K <= N - 1, TKA_E_WRONG_PARAMETER, "K must be below the number of points");
        if (perplexity > K) log_message(TKA_LOG_WARNING, "Perplexity should be lower than K!");  // tsne.hpp:604-605
    }
    if (o.y0_mode == TKA_Y0_EXPLICIT) require(o.y0 != nullptr, TKA_E_WRONG_PARAMETER, "TKA_Y0_EXPLICIT needs y0");
    log_message(TKA_LOG_INFO, exact ? "Using exact t-SNE algorithm" : "Using Barnes-Hut-SNE algorithm",  // tsne.hpp:66,68
                o.verbose != 0);
    TKA_HIP(hipSetDevice(o.device));
    std::unique_ptr<tka_session> sess;
    if (o.loop_precision == TKA_FP64)
        sess.reset(new Session<double>());
    else
        sess.reset(new Session<float>());
    tka_session& c = *sess;
    c.o = o;
    c.mb.cap = o.device_mem_cap;
    c.N = N;
    c.dims = no_dims;
    c.theta = theta;
    c.exact = exact;
    c.dd_mode = o.dd_mode;
    hipStream_t s = c.s();
    // every buffer of the build, and the session's, is used on s only: their
    // frees record the reuse event there (not on the null stream, which
    // would first wait for every other stream)
    AllocStream alloc_on_s(s);
    MemBudget* mb = &c.mb;
    tka_tsne_stats& st = c.st;
    const double t0 = now_ms();

    DevBuf<double> X;  // normalise (tsne.hpp:88-89)
    if (features) {
        upload(X, in.X, (size_t)N * D, s, mb);
        c.stream.sync();
        st.ms_upload = now_ms() - t0;
        normalize_device(X.get(), N, D, mb, s);
        c.stream.sync();
    }
    const double t1 = now_ms();
    st.ms_normalize = t1 - t0;

    DevBuf<uint32_t> order;  // the kNN's locality order of the points, if it made one
    bool ordered = false;
    auto fill = [&](auto* S) {
        using TT = typename std::remove_reference<decltype(*S->L.val.get())>::type;
        if (in.kind == SimilarityInput::kJointP) {
            // a precomputed joint P, divided by its sum (tsne.hpp:120-124)
            const int64_t nnz = in.row[N];
            Csr P;
            upload(P.row, in.row, (size_t)N + 1, s, mb);
            upload(P.col, in.col, (size_t)std::max<int64_t>(nnz, 1), s, mb);
            upload(P.val, in.val, (size_t)std::max<int64_t>(nnz, 1), s, mb);
            P.nnz = nnz;
            normalize_sum_device(P.val.get(), P.nnz, mb, s);
            c.stream.sync();
            st.ms_symmetrize = now_ms() - t1;
            st.nnz = P.nnz;
            S->L.row = std::move(P.row);
            S->L.col = std::move(P.col);
            S->L.val.alloc((size_t)std::max<int64_t>(P.nnz, 1), mb);
            cast_device<double, TT>(P.val.get(), S->L.val.get(), P.nnz, s);
        } else if (sparse) {
            DevBuf<int> idx((size_t)N * K, mb);
            DevBuf<double> dist((size_t)N * K, mb), val((size_t)N * K, mb);
            KnnStats ks;
            if (features) {
                const bool want_order = !exact;  // the Barnes-Hut loops renumber points
                if (want_order) order.alloc(N, mb);
                knn_device(X.get(), N, D, K, o.dist_dtype, o.candidates, idx.get(), dist.get(), &ks, mb, s,
                           want_order ? order.get() : nullptr);
                ordered = ks.order_written;
            } else {  // a precomputed neighbour graph
                upload(idx, in.idx, (size_t)N * K, s, mb);
                upload(dist, in.dist, (size_t)N * K, s, mb);
            }
            c.stream.sync();
            const double t2 = now_ms();
            X.release();
            const int unconv = perplexity_knn_device(dist.get(), N, K, perplexity, val.get(), nullptr, nullptr, mb, s);
            c.stream.sync();
            const double t3 = now_ms();
            Csr P;
            symmetrize_knn_device(idx.get(), val.get(), N, K, P, mb, s, ordered ? order.get() : nullptr);
            normalize_sum_device(P.val.get(), P.nnz, mb, s);  // tsne.hpp:120-124
            c.stream.sync();
            const double t4 = now_ms();
            st.ms_knn = t2 - t1;
            st.ms_perplexity = t3 - t2;
            st.ms_symmetrize = t4 - t3;
            st.nnz = P.nnz;
            st.knn_k = K;
            st.bisection_unconverged = unconv;
            st.certificate_failures = ks.certificate_failures;
            S->L.row = std::move(P.row);
            S->L.col = std::move(P.col);
            S->L.val.alloc(P.nnz, mb);
            cast_device<double, TT>(P.val.get(), S->L.val.get(), P.nnz, s);
        } else {
            DevBuf<double> DD((size_t)N * N, mb), P((size_t)N * N, mb);
            sqdist_dense_device(X.get(), N, D, o.dd_mode, DD.get(), s);
            const int unconv = perplexity_dense_device(DD.get(), N, perplexity, P.get(), mb, s);
            DD.release();
            symmetrize_dense_device(P.get(), N, mb, s);  // tsne.hpp:100-108
            c.stream.sync();
            st.ms_perplexity = now_ms() - t1;
            st.nnz = (int64_t)N * N;
            st.bisection_unconverged = unconv;
            S->L.dense = true;
            S->L.val.alloc((size_t)N * N, mb);
            cast_device<double, TT>(P.get(), S->L.val.get(), (size_t)N * N, s);
        }
        c.stream.sync();
    };
    if (o.loop_precision == TKA_FP64)
        fill(static_cast<Session<double>*>(sess.get()));
    else
        fill(static_cast<Session<float>*>(sess.get()));

    // Y0 (tsne.hpp:137-138): std::rand after the VP-tree's draws on the BH path.
    std::vector<double> Y0((size_t)N * no_dims);
    if (o.y0_mode == TKA_Y0_EXPLICIT) {
        std::memcpy(Y0.data(), o.y0, Y0.size() * sizeof(double));
        check_finite(Y0.data(), Y0.size(), "y0");
    } else {
        if (o.y0_mode == TKA_Y0_SEEDED) std::srand(o.seed);
        if (!exact && features) {  // the reference's VP-tree consumed these draws
            const long long burn = vptree_draws(N);
            for (long long i = 0; i < burn; i++) (void)std::rand();
        }
        for (auto& v : Y0) v = gaussian_random() * .0001;
    }
    // Barnes-Hut loop: the points start in the kNN's locality order, in which
    // each point's P neighbours have nearby numbers (the attraction's
    // gathers stay local from the first iteration; the tree's Morton order
    // takes over at the first periodic renumbering)
    if (o.loop_precision == TKA_FP64) {
        static_cast<Session<double>*>(sess.get())->init(Y0);
        if (ordered) {
            auto* S = static_cast<Session<double>*>(sess.get());
            S->renumber(order.get(), S->tree_fault());
        }
    } else {
        static_cast<Session<float>*>(sess.get())->init(Y0);
        if (ordered) {
            auto* S = static_cast<Session<float>*>(sess.get());
            S->renumber(order.get(), S->tree_fault());
        }
    }
    c.stream.sync();
    log_took("Input similarities computation", (now_ms() - t0) * 1e-3, o.verbose != 0);  // tsne.hpp:87
    return sess.release();
}

void tsne_run_impl(const double* Xh, int N, int D, double* Yh, int no_dims, double perplexity, double theta,
                   const tka_tsne_options* opt_in, tka_tsne_stats* st)
{
    require(Yh != nullptr, TKA_E_WRONG_PARAMETER, "null output pointer");
    const double t0 = now_ms();
    SimilarityInput in;
    in.X = Xh;
    in.D = D;
    std::unique_ptr<tka_session> sess(create_session(in, N, no_dims, perplexity, theta, opt_in));
    const double t5 = now_ms();
    sess->step(sess->o.max_iter);                       // tsne.hpp:143-193
    sess->embedding(Yh);
    const double t6 = now_ms();
    log_took("Main t-SNE loop", (t6 - t5) * 1e-3, sess->o.verbose != 0);  // tsne.hpp:142
    sess->st.kl_final_exact = sess->kl_exact();
    sess->st.ms_loop = t6 - t5;
    sess->st.ms_total = t6 - t0;
    if (st) *st = sess->st;
}

// Barnes-Hut repulsion of y (N x dims, dims 2 or 3; centred in place by the
// tree build) on a tree built for this call: neg and the Z partials; returns
// the number of partials.  Fails if the tree's Morton sort gave up.
template <class T>
int bh_forces_once(int dims, T* y, int N, double theta, T* neg, DevBuf<double>& zpart, MemBudget* mb, hipStream_t s)
{
    require(dims == 2 || dims == 3, TKA_E_WRONG_PARAMETER, "Barnes-Hut needs 2 or 3 dimensions");
    const int nz = bh_zblocks(N);
    zpart.alloc(nz, mb);
    if (dims == 3) {
        BhTree3<T> tree;
        tree.alloc(N, mb);
        bh3_build<T>(y, N, tree, s);
        bh3_repulsion<T>(tree, theta, neg, zpart.get(), nullptr, s);
        require(!bh3_sort_failed(tree, s), TKA_E_INTERNAL, "radix sort look-back gave up; the tree is not valid");
    } else {
        DevBuf<double> box((size_t)kBoxStride * bbox_blocks(N), mb);
        bbox_partials_device<T>(y, N, box.get(), s);
        BhTree<T> tree;
        tree.alloc(N, mb);
        bh_build<T>(y, N, box.get(), bbox_blocks(N), tree, s);
        bh_repulsion<T>(tree, y, theta, neg, zpart.get(), nullptr, s);
        require(!bh_sort_failed(tree, s), TKA_E_INTERNAL, "radix sort look-back gave up; the tree is not valid");
    }
    return nz;
}

template <class T>
void gradient_impl(const double* Yh, int N, int dims, const int* rowh, const int* colh, const double* valh,
                   double theta, double* dYh, double* sum_q)
{
    MemBudget mb;
    Stream st;
    hipStream_t s = st.s;
    const size_t nd = (size_t)N * dims;
    const int64_t nnz = rowh[N];
    DevBuf<double> tmp;
    DevBuf<T> Y(nd), neg(nd), dC(nd), val(nnz ? nnz : 1);
    upload(tmp, Yh, nd, s);
    cast_device<double, T>(tmp.get(), Y.get(), nd, s);
    DevBuf<int> row, col;
    upload(row, rowh, (size_t)N + 1, s);
    upload(col, colh, (size_t)nnz, s);
    DevBuf<double> vtmp;
    upload(vtmp, valh, (size_t)nnz, s);
    cast_device<double, T>(vtmp.get(), val.get(), nnz, s);
    int nz;
    DevBuf<double> zpart;
    if (theta > 0.0) {
        nz = bh_forces_once<T>(dims, Y.get(), N, theta, neg.get(), zpart, &mb, s);
        gradient_combine_device<T>(dims, Y.get(), neg.get(), nullptr, row.get(), col.get(), val.get(), 1.0,
                                   zpart.get(), nz, N, dC.get(), s);
    } else {
        nz = exact_zblocks(N);
        zpart.alloc(nz);
        DevBuf<T> work(exact_work_elems(N, dims, false), &mb);
        exact_forces_device<T>(dims, Y.get(), N, TKA_DD_TEXTBOOK, nullptr, 1.0, neg.get(), nullptr, zpart.get(),
                               work.get(), s);
        gradient_combine_device<T>(dims, Y.get(), neg.get(), nullptr, row.get(), col.get(), val.get(), 1.0,
                                   zpart.get(), nz, N, dC.get(), s);
    }
    DevBuf<double> out(nd);
    cast_device<T, double>(dC.get(), out.get(), nd, s);
    download(dYh, out.get(), nd, s);
    const double z = sum_device(zpart.get(), nz, &mb, s);
    st.sync();
    if (sum_q) *sum_q = z;
}

template <class T>
void gradient_dense_impl(const double* Ph, const double* Yh, int N, int dims, int dd_mode, double* dYh, double* sum_q)
{
    MemBudget mb;
    Stream st;
    hipStream_t s = st.s;
    const size_t nd = (size_t)N * dims;
    DevBuf<double> tmp, ptmp;
    DevBuf<T> Y(nd), P((size_t)N * N), neg(nd), pos(nd), dC(nd);
    upload(tmp, Yh, nd, s);
    cast_device<double, T>(tmp.get(), Y.get(), nd, s);
    upload(ptmp, Ph, (size_t)N * N, s);
    cast_device<double, T>(ptmp.get(), P.get(), (size_t)N * N, s);
    const int nz = exact_zblocks(N);
    DevBuf<double> zpart(nz);
    DevBuf<T> work(exact_work_elems(N, dims, true), &mb);
    exact_forces_device<T>(dims, Y.get(), N, dd_mode, P.get(), 1.0, neg.get(), pos.get(), zpart.get(), work.get(), s);
    gradient_combine_device<T>(dims, Y.get(), neg.get(), pos.get(), nullptr, nullptr, nullptr, 1.0, zpart.get(), nz,
                               N, dC.get(), s);
    DevBuf<double> out(nd);
    cast_device<T, double>(dC.get(), out.get(), nd, s);
    download(dYh, out.get(), nd, s);
    const double z = sum_device(zpart.get(), nz, &mb, s);
    st.sync();
    if (sum_q) *sum_q = z;
}

}  // namespace
}  // namespace tka

using namespace tka;

extern "C" {

TKA_API void tka_set_logger(tka_log_fn fn, void* user, uint32_t enabled_levels)
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    g_log_fn = fn;
    g_log_user = fn ? user : nullptr;
    g_log_levels = enabled_levels;
}

TKA_API void tka_default_options(tka_tsne_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->struct_size = (int32_t)sizeof(tka_tsne_options);
    o->max_iter = 1000;        // tsne.hpp:71
    o->stop_lying_iter = 250;
    o->mom_switch_iter = 250;
    o->momentum = 0.5;         // tsne.hpp:72
    o->final_momentum = 0.8;
    o->eta = 200.0;            // tsne.hpp:73
    o->exaggeration = 12.0;    // tsne.hpp:129-134
    o->knn_k = 0;
    o->dist_dtype = TKA_DIST_BF16;
    o->candidates = 0;
    o->loop_precision = TKA_FP32;
    o->exact_p_mode = TKA_EXACT_P_DENSE;
    o->dd_mode = TKA_DD_TEXTBOOK;
    o->y0_mode = TKA_Y0_RAND_COMPAT;
    o->seed = 1;
    o->y0 = nullptr;
    o->log_every = 50;         // tsne.hpp:184
    o->kl_mode = TKA_KL_AS_REFERENCE;
    o->verbose = 0;
    o->device = 0;
    o->device_mem_cap = 0;
    o->cancel = nullptr;
}

TKA_API const char* tka_last_error(void) { return last_error().c_str(); }

TKA_API int32_t tka_version(void) { return 10000; }

TKA_API tka_status tka_tsne_run(const double* X, int32_t N, int32_t D, double* Y, int32_t no_dims, double perplexity,
                                double theta, const tka_tsne_options* opt, tka_tsne_stats* stats)
{
    return guard([&] { tsne_run_impl(X, N, D, Y, no_dims, perplexity, theta, opt, stats); });
}

TKA_API tka_status tka_session_create(const double* X, int32_t N, int32_t D, int32_t no_dims, double perplexity,
                                      double theta, const tka_tsne_options* opt, tka_session** out)
{
    return guard([&] {
        require(out != nullptr, TKA_E_WRONG_PARAMETER, "null session pointer");
        *out = nullptr;
        SimilarityInput in;
        in.X = X;
        in.D = D;
        *out = create_session(in, N, no_dims, perplexity, theta, opt);
    });
}

TKA_API tka_status tka_session_create_knn(const int32_t* idx, const double* dist, int32_t N, int32_t K, int32_t no_dims,
                                          double perplexity, double theta, const tka_tsne_options* opt,
                                          tka_session** out)
{
    return guard([&] {
        require(out != nullptr, TKA_E_WRONG_PARAMETER, "null session pointer");
        *out = nullptr;
        SimilarityInput in;
        in.kind = SimilarityInput::kKnn;
        in.idx = idx;
        in.dist = dist;
        in.K = K;
        *out = create_session(in, N, no_dims, perplexity, theta, opt);
    });
}

TKA_API tka_status tka_session_create_p(const int32_t* row_ptr, const int32_t* col, const double* val, int32_t N,
                                        int32_t no_dims, double theta, const tka_tsne_options* opt, tka_session** out)
{
    return guard([&] {
        require(out != nullptr, TKA_E_WRONG_PARAMETER, "null session pointer");
        *out = nullptr;
        SimilarityInput in;
        in.kind = SimilarityInput::kJointP;
        in.row = row_ptr;
        in.col = col;
        in.val = val;
        *out = create_session(in, N, no_dims, 0.0, theta, opt);
    });
}

TKA_API tka_status tka_session_step(tka_session* s, int32_t iterations)
{
    return guard([&] {
        require(s != nullptr && iterations >= 0, TKA_E_WRONG_PARAMETER, "bad arguments");
        s->step(iterations);
    });
}

TKA_API tka_status tka_session_synchronize(tka_session* s)
{
    return guard([&] {
        require(s != nullptr, TKA_E_WRONG_PARAMETER, "null session");
        s->stream.sync();
    });
}

TKA_API tka_status tka_session_embedding(tka_session* s, double* Y)
{
    return guard([&] {
        require(s != nullptr && Y != nullptr, TKA_E_WRONG_PARAMETER, "bad arguments");
        s->embedding(Y);
    });
}

TKA_API tka_status tka_session_kl(tka_session* s, double* kl)
{
    return guard([&] {
        require(s != nullptr && kl != nullptr, TKA_E_WRONG_PARAMETER, "bad arguments");
        *kl = s->kl_exact();
    });
}

TKA_API tka_status tka_session_stats(tka_session* s, tka_tsne_stats* stats)
{
    return guard([&] {
        require(s != nullptr && stats != nullptr, TKA_E_WRONG_PARAMETER, "bad arguments");
        *stats = s->st;
    });
}

TKA_API tka_status tka_session_profile(tka_session* s, int32_t enable)
{
    return guard([&] {
        require(s != nullptr, TKA_E_WRONG_PARAMETER, "null session");
        s->profile = enable != 0;
    });
}

TKA_API tka_status tka_session_inject_sort_fault(tka_session* s)
{
    return guard([&] {
        require(s != nullptr, TKA_E_WRONG_PARAMETER, "null session");
        s->inject_sort_fault();
    });
}

TKA_API tka_status tka_session_kernel_timing(tka_session* s, tka_kernel_timing* t)
{
    return guard([&] {
        require(s != nullptr && t != nullptr, TKA_E_WRONG_PARAMETER, "bad arguments");
        s->stream.sync();
        std::memset(t, 0, sizeof(*t));
        for (auto& tm : s->timers) tm.collect();
        t->launches_build = s->timers[kTimBuild].launches;
        t->launches_repulsion = s->timers[kTimRepulsion].launches;
        t->launches_step = s->timers[kTimStep].launches;
        t->launches_exact = s->timers[kTimExact].launches;
        t->ms_build = s->timers[kTimBuild].ms;
        t->ms_repulsion = s->timers[kTimRepulsion].ms;
        t->ms_step = s->timers[kTimStep].ms;
        t->ms_exact = s->timers[kTimExact].ms;
        t->launches_attraction = s->timers[kTimAttract].launches;
        t->ms_attraction = s->timers[kTimAttract].ms;
        unsigned long long inter = 0;
        if (s->interactions.get()) {
            TKA_HIP(hipMemcpy(&inter, s->interactions.get(), sizeof(inter), hipMemcpyDeviceToHost));
            TKA_HIP(hipMemsetAsync(s->interactions.get(), 0, sizeof(inter), s->s()));
            s->stream.sync();
        }
        t->interactions = (int64_t)inter;
        t->nnz = s->st.nnz;
        t->n = s->N;
        t->dims = s->dims;
        for (auto& tm : s->timers) {
            tm.ms = 0.0;
            tm.launches = 0;
        }
    });
}

TKA_API void tka_session_destroy(tka_session* s) { delete s; }

TKA_API tka_status tka_normalize(double* X, int32_t N, int32_t D)
{
    return guard([&] {
        require(N > 0 && D > 0 && X, TKA_E_WRONG_PARAMETER, "bad shape");
        check_finite(X, (size_t)N * D, "input data");
        Stream st;
        DevBuf<double> d;
        upload(d, X, (size_t)N * D, st.s);
        normalize_device(d.get(), N, D, nullptr, st.s);
        download(X, d.get(), (size_t)N * D, st.s);
        st.sync();
    });
}

TKA_API tka_status tka_knn(const double* X, int32_t N, int32_t D, int32_t K, int32_t dist_dtype, int32_t candidates,
                           int32_t* idx, double* dist, int32_t* certificate_failures)
{
    return guard([&] {
        require(N > 0 && D > 0 && X && idx && dist, TKA_E_WRONG_PARAMETER, "bad arguments");
        check_finite(X, (size_t)N * D, "input data");
        Stream st;
        DevBuf<double> d;
        upload(d, X, (size_t)N * D, st.s);
        DevBuf<int> di((size_t)N * K);
        DevBuf<double> dd((size_t)N * K);
        KnnStats ks;
        knn_device(d.get(), N, D, K, dist_dtype, candidates, di.get(), dd.get(), &ks, nullptr, st.s);
        download(idx, di.get(), (size_t)N * K, st.s);
        download(dist, dd.get(), (size_t)N * K, st.s);
        st.sync();
        if (certificate_failures) *certificate_failures = ks.certificate_failures;
    });
}

TKA_API tka_status tka_perplexity(const double* dist, int32_t N, int32_t K, double perplexity, double* val,
                                  double* beta, int32_t* iters, int32_t* unconverged)
{
    return guard([&] {
        require(N > 0 && K > 0 && dist && val, TKA_E_WRONG_PARAMETER, "bad arguments");
        require(perplexity > 0.0, TKA_E_WRONG_PARAMETER, "perplexity must be positive");
        Stream st;
        DevBuf<double> d, v((size_t)N * K), b(N);
        DevBuf<int> it(N);
        upload(d, dist, (size_t)N * K, st.s);
        const int u = perplexity_knn_device(d.get(), N, K, perplexity, v.get(), b.get(), it.get(), nullptr, st.s);
        download(val, v.get(), (size_t)N * K, st.s);
        if (beta) download(beta, b.get(), N, st.s);
        if (iters) download(iters, it.get(), N, st.s);
        st.sync();
        if (unconverged) *unconverged = u;
    });
}

TKA_API tka_status tka_perplexity_dense(const double* X, int32_t N, int32_t D, double perplexity, int32_t dd_mode,
                                        double* P)
{
    return guard([&] {
        require(N > 1 && D > 0 && X && P, TKA_E_WRONG_PARAMETER, "bad arguments");
        Stream st;
        DevBuf<double> x, DD((size_t)N * N), Pd((size_t)N * N);
        upload(x, X, (size_t)N * D, st.s);
        sqdist_dense_device(x.get(), N, D, dd_mode, DD.get(), st.s);
        perplexity_dense_device(DD.get(), N, perplexity, Pd.get(), nullptr, st.s);
        download(P, Pd.get(), (size_t)N * N, st.s);
        st.sync();
    });
}

TKA_API tka_status tka_symmetrize(const int32_t* idx, const double* val, int32_t N, int32_t K, int32_t* row_ptr,
                                  int32_t* col, double* sval, int64_t* nnz)
{
    return guard([&] {
        require(N > 0 && K > 0 && idx && val && row_ptr && col && sval, TKA_E_WRONG_PARAMETER, "bad arguments");
        for (size_t e = 0; e < (size_t)N * K; e++)
            require(idx[e] >= 0 && idx[e] < N, TKA_E_WRONG_PARAMETER, "neighbour index out of range");
        Stream st;
        DevBuf<int> di;
        DevBuf<double> dv;
        upload(di, idx, (size_t)N * K, st.s);
        upload(dv, val, (size_t)N * K, st.s);
        Csr P;
        symmetrize_knn_device(di.get(), dv.get(), N, K, P, nullptr, st.s);
        download(row_ptr, P.row.get(), (size_t)N + 1, st.s);
        download(col, P.col.get(), (size_t)P.nnz, st.s);
        download(sval, P.val.get(), (size_t)P.nnz, st.s);
        st.sync();
        if (nnz) *nnz = P.nnz;
    });
}

TKA_API tka_status tka_build_p(const double* X, int32_t N, int32_t D, double perplexity, int32_t knn_k,
                               int32_t dist_dtype, int32_t* row_ptr, int32_t* col, double* val, int64_t* nnz)
{
    return guard([&] {
        require(N > 1 && D > 0 && X && row_ptr && col && val, TKA_E_WRONG_PARAMETER, "bad arguments");
        require(perplexity > 0.0 && perplexity <= (N - 1) / 3.0, TKA_E_WRONG_PARAMETER,
                "perplexity must lie in (0, (N-1)/3]");
        const int K = knn_k > 0 ? knn_k : (int)(3 * perplexity);
        require(K >= 1 && K <= N - 1, TKA_E_WRONG_PARAMETER, "K must lie in [1, N-1]");
        check_finite(X, (size_t)N * D, "input data");
        Stream st;
        hipStream_t s = st.s;
        DevBuf<double> x;
        upload(x, X, (size_t)N * D, s);
        normalize_device(x.get(), N, D, nullptr, s);
        DevBuf<int> idx((size_t)N * K);
        DevBuf<double> dist((size_t)N * K), pv((size_t)N * K);
        KnnStats ks;
        knn_device(x.get(), N, D, K, dist_dtype, 0, idx.get(), dist.get(), &ks, nullptr, s);
        perplexity_knn_device(dist.get(), N, K, perplexity, pv.get(), nullptr, nullptr, nullptr, s);
        Csr P;
        symmetrize_knn_device(idx.get(), pv.get(), N, K, P, nullptr, s);
        normalize_sum_device(P.val.get(), P.nnz, nullptr, s);
        download(row_ptr, P.row.get(), (size_t)N + 1, s);
        download(col, P.col.get(), (size_t)P.nnz, s);
        download(val, P.val.get(), (size_t)P.nnz, s);
        st.sync();
        if (nnz) *nnz = P.nnz;
    });
}

static void check_csr(const int32_t* row_ptr, const int32_t* col, int N)
{
    require(row_ptr[0] == 0, TKA_E_WRONG_PARAMETER, "row_ptr[0] must be 0");
    for (int i = 0; i < N; i++) require(row_ptr[i + 1] >= row_ptr[i], TKA_E_WRONG_PARAMETER, "row_ptr not monotone");
    for (int64_t e = 0; e < row_ptr[N]; e++)
        require(col[e] >= 0 && col[e] < N, TKA_E_WRONG_PARAMETER, "column index out of range");
}

TKA_API tka_status tka_gradient(const double* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr, const int32_t* col,
                                const double* val, double theta, int32_t precision, double* dY, double* sum_q)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && Y && row_ptr && col && val && dY, TKA_E_WRONG_PARAMETER,
                "bad arguments");
        require(theta >= 0.0, TKA_E_WRONG_PARAMETER, "theta must be non-negative");
        check_csr(row_ptr, col, N);
        check_finite(Y, (size_t)N * no_dims, "Y");
        if (precision == TKA_FP64)
            gradient_impl<double>(Y, N, no_dims, row_ptr, col, val, theta, dY, sum_q);
        else
            gradient_impl<float>(Y, N, no_dims, row_ptr, col, val, theta, dY, sum_q);
    });
}

TKA_API tka_status tka_gradient_dense(const double* P, const double* Y, int32_t N, int32_t no_dims, int32_t dd_mode,
                                      int32_t precision, double* dY, double* sum_q)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && P && Y && dY, TKA_E_WRONG_PARAMETER, "bad arguments");
        if (precision == TKA_FP64)
            gradient_dense_impl<double>(P, Y, N, no_dims, dd_mode, dY, sum_q);
        else
            gradient_dense_impl<float>(P, Y, N, no_dims, dd_mode, dY, sum_q);
    });
}

TKA_API tka_status tka_kl(const double* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr, const int32_t* col,
                          const double* val, double theta, int32_t kl_mode, double* kl)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && Y && row_ptr && col && val && kl, TKA_E_WRONG_PARAMETER,
                "bad arguments");
        check_csr(row_ptr, col, N);
        MemBudget mb;
        Stream st;
        hipStream_t s = st.s;
        const size_t nd = (size_t)N * no_dims;
        DevBuf<double> Yd, vd;
        DevBuf<int> row, colv;
        upload(Yd, Y, nd, s);
        upload(row, row_ptr, (size_t)N + 1, s);
        upload(colv, col, (size_t)row_ptr[N], s);
        upload(vd, val, (size_t)row_ptr[N], s);
        if (kl_mode == TKA_KL_AS_REFERENCE && theta > 0.0) {
            DevBuf<double> zpart, neg(nd);
            const int nz = bh_forces_once<double>(no_dims, Yd.get(), N, theta, neg.get(), zpart, &mb, s);
            const double Z = sum_device(zpart.get(), nz, &mb, s);
            *kl = kl_edges_reference_device<double>(no_dims, Yd.get(), row.get(), colv.get(), vd.get(), 1.0, Z, N, &mb, s);
        } else {
            const double Z = exact_z_device<double>(no_dims, Yd.get(), N, &mb, s);
            *kl = kl_edges_exact_device<double>(no_dims, Yd.get(), row.get(), colv.get(), vd.get(), 1.0, Z, N, &mb, s);
        }
    });
}

TKA_API tka_status tka_kl_dense(const double* P, const double* Y, int32_t N, int32_t no_dims, int32_t dd_mode,
                                double* kl)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && P && Y && kl, TKA_E_WRONG_PARAMETER, "bad arguments");
        MemBudget mb;
        Stream st;
        DevBuf<double> Yd, Pd;
        upload(Yd, Y, (size_t)N * no_dims, st.s);
        upload(Pd, P, (size_t)N * N, st.s);
        *kl = kl_dense_device<double>(no_dims, Yd.get(), Pd.get(), 1.0, N, dd_mode, &mb, st.s);
    });
}

TKA_API tka_status tka_update(const double* dY, double* uY, double* gains, double* Y, int32_t N, int32_t no_dims,
                              double momentum, double eta, int32_t precision)
{
    return guard([&] {
        require(N >= 1 && no_dims >= 1 && no_dims <= 3 && dY && uY && gains && Y, TKA_E_WRONG_PARAMETER,
                "bad arguments");
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Stream st;
            hipStream_t s = st.s;
            const size_t nd = (size_t)N * no_dims;
            DevBuf<double> t0, t1, t2, t3;
            DevBuf<T> d(nd), u(nd), g(nd), y(nd), yn(nd), zero(nd);
            upload(t0, dY, nd, s);
            upload(t1, uY, nd, s);
            upload(t2, gains, nd, s);
            upload(t3, Y, nd, s);
            cast_device<double, T>(t0.get(), d.get(), nd, s);
            cast_device<double, T>(t1.get(), u.get(), nd, s);
            cast_device<double, T>(t2.get(), g.get(), nd, s);
            cast_device<double, T>(t3.get(), y.get(), nd, s);
            fill_device(zero.get(), nd, 0.0, s);
            DevBuf<double> one(1), box((size_t)kBoxStride * tsne_step_blocks(N));
            fill_device(one.get(), 1, 1.0, s);
            StepParams sp{momentum, eta, 1.0};
            const int nb = tsne_step_device<T>(no_dims, y.get(), yn.get(), u.get(), g.get(), zero.get(), d.get(),
                                               nullptr, nullptr, nullptr, one.get(), 1, N, sp, box.get(), nullptr, s);
            center_device<T>(no_dims, yn.get(), N, box.get(), nb, s);
            cast_device<T, double>(u.get(), t1.get(), nd, s);
            cast_device<T, double>(g.get(), t2.get(), nd, s);
            cast_device<T, double>(yn.get(), t3.get(), nd, s);
            download(uY, t1.get(), nd, s);
            download(gains, t2.get(), nd, s);
            download(Y, t3.get(), nd, s);
            st.sync();
        };
        if (precision == TKA_FP64)
            run(double{});
        else
            run(float{});
    });
}

TKA_API tka_status tka_bh_tree(const double* Y, int32_t N, int32_t precision, int32_t* n_nodes, int32_t* level,
                               int32_t* count, int32_t* leaf_point, int32_t* skip, double* com_x, double* com_y,
                               double* root_box)
{
    return guard([&] {
        require(N >= 2 && Y && n_nodes, TKA_E_WRONG_PARAMETER, "bad arguments");
        check_finite(Y, (size_t)N * 2, "Y");
        auto run = [&](auto tag) {
            using T = decltype(tag);
            MemBudget mb;
            Stream st;
            hipStream_t s = st.s;
            DevBuf<double> tmp;
            DevBuf<T> y((size_t)N * 2, &mb);
            upload(tmp, Y, (size_t)N * 2, s);
            DevBuf<double> box((size_t)kBoxStride * bbox_blocks(N));
            BhTree<T> tree;
            tree.alloc(N, &mb);
            // One build per radix width from the same input (the key kernel
            // centres Y in place): 2, 3 and 4 passes up to 2^22 points (4
            // beyond), then one with the width the last build's run counts
            // ask for; the trees must agree node for node.
            std::vector<int> widths;
            if (N <= (1 << 22)) widths = {2, 3, 4};
            else widths = {4};
            widths.push_back(0);
            std::vector<BhNode<T>> first;
            uint32_t total = 0;
            RootBox rb;
            for (size_t b = 0; b < widths.size(); b++) {
                tree.forced_passes = widths[b];
                cast_device<double, T>(tmp.get(), y.get(), (size_t)N * 2, s);
                bbox_partials_device<T>(y.get(), N, box.get(), s);
                bh_build<T>(y.get(), N, box.get(), bbox_blocks(N), tree, s);
                require(!bh_sort_failed(tree, s), TKA_E_INTERNAL, "radix sort look-back gave up; the tree is not valid");
                download(&total, tree.pre.get() + N, 1, s);
                download(&rb, tree.root.get(), 1, s);
                st.sync();
                std::vector<BhNode<T>> nodes(total);
                download(nodes.data(), tree.nodes.get(), total, s);
                st.sync();
                if (b == 0) {
                    first.swap(nodes);
                } else {
                    require(nodes.size() == first.size() &&
                                std::memcmp(nodes.data(), first.data(), nodes.size() * sizeof(BhNode<T>)) == 0,
                            TKA_E_INTERNAL, "Barnes-Hut trees built with different radix widths differ");
                }
            }
            std::vector<BhNode<T>>& nodes = first;
            *n_nodes = (int32_t)total;
            for (uint32_t t = 0; t < total; t++) {
                const bool leaf = (nodes[t].link & kLeafBit) != 0;
                if (level) level[t] = leaf ? -1 : (int)(nodes[t].cl >> kCountBits);
                if (count) count[t] = leaf ? 1 : (int)(nodes[t].cl & ((1u << kCountBits) - 1u));
                if (leaf_point) leaf_point[t] = leaf ? (int)(nodes[t].link & ~kLeafBit) : -1;
                if (skip) skip[t] = leaf ? (int)t + 1 : (int)nodes[t].link;
                if (com_x) com_x[t] = (double)nodes[t].x;
                if (com_y) com_y[t] = (double)nodes[t].y;
            }
            if (root_box) {
                root_box[0] = rb.mean_x;
                root_box[1] = rb.mean_y;
                root_box[2] = rb.hw;
                root_box[3] = rb.hh;
            }
        };
        if (precision == TKA_FP64)
            run(double{});
        else
            run(float{});
    });
}

TKA_API tka_status tka_gradient_dev(const float* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                    const int32_t* col, const float* val, double theta, float* dY, double* sum_q,
                                    void* stream)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && Y && row_ptr && col && val && dY, TKA_E_WRONG_PARAMETER,
                "bad arguments");
        require(theta >= 0.0, TKA_E_WRONG_PARAMETER, "theta must be non-negative");
        MemBudget mb;
        Stream st(stream);
        hipStream_t s = st.s;
        AllocStream as(s);  // temporaries ordered on the caller's stream, which may be non-blocking
        const size_t nd = (size_t)N * no_dims;
        DevBuf<float> neg(nd);
        int nz;
        DevBuf<double> zpart;
        if (theta > 0.0) {
            // the tree build centres its input: work on a copy
            DevBuf<float> y(nd);
            TKA_HIP(hipMemcpyAsync(y.get(), Y, nd * sizeof(float), hipMemcpyDeviceToDevice, s));
            nz = bh_forces_once<float>(no_dims, y.get(), N, theta, neg.get(), zpart, &mb, s);
            gradient_combine_device<float>(no_dims, y.get(), neg.get(), nullptr, row_ptr, col, val, 1.0, zpart.get(),
                                           nz, N, dY, s);
            const double z = sum_device(zpart.get(), nz, &mb, s);
            if (sum_q) *sum_q = z;
        } else {
            nz = exact_zblocks(N);
            zpart.alloc(nz);
            DevBuf<float> work(exact_work_elems(N, no_dims, false));
            exact_forces_device<float>(no_dims, Y, N, TKA_DD_TEXTBOOK, nullptr, 1.0, neg.get(), nullptr, zpart.get(),
                                       work.get(), s);
            gradient_combine_device<float>(no_dims, Y, neg.get(), nullptr, row_ptr, col, val, 1.0, zpart.get(), nz, N,
                                           dY, s);
            const double z = sum_device(zpart.get(), nz, &mb, s);
            if (sum_q) *sum_q = z;
        }
    });
}

TKA_API tka_status tka_kl_dev(const float* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr, const int32_t* col,
                              const float* val, double theta, double* kl, double* sum_q, void* stream)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && Y && row_ptr && col && val && kl, TKA_E_WRONG_PARAMETER,
                "bad arguments");
        MemBudget mb;
        Stream st(stream);
        hipStream_t s = st.s;
        AllocStream as(s);  // temporaries ordered on the caller's stream, which may be non-blocking
        double Z;
        if (theta > 0.0) {
            const size_t nd = (size_t)N * no_dims;
            DevBuf<float> y(nd), neg(nd);
            TKA_HIP(hipMemcpyAsync(y.get(), Y, nd * sizeof(float), hipMemcpyDeviceToDevice, s));
            DevBuf<double> zpart;
            const int nz = bh_forces_once<float>(no_dims, y.get(), N, theta, neg.get(), zpart, &mb, s);
            Z = sum_device(zpart.get(), nz, &mb, s);
        } else {
            Z = exact_z_device<float>(no_dims, Y, N, &mb, s);
        }
        *kl = kl_edges_exact_device<float>(no_dims, Y, row_ptr, col, val, 1.0, Z, N, &mb, s);
        if (sum_q) *sum_q = Z;
    });
}

}  // extern "C"

namespace tka {
namespace {

// KL(P||Q) and dC = pos - neg/Z at one Y on the caller's stream, sharing Z.
template <class T>
void kl_grad_impl(const T* Y, int N, int dims, const int* row_ptr, const int* col, const T* val, double theta,
                  double* kl, T* dY, double* sum_q, hipStream_t s)
{
    MemBudget mb;
    const size_t nd = (size_t)N * dims;
    DevBuf<T> neg(nd);
    DevBuf<double> zpart;
    double Z;
    if (theta > 0.0) {
        DevBuf<T> y(nd);  // the tree build centres its input
        TKA_HIP(hipMemcpyAsync(y.get(), Y, nd * sizeof(T), hipMemcpyDeviceToDevice, s));
        const int nz = bh_forces_once<T>(dims, y.get(), N, theta, neg.get(), zpart, &mb, s);
        gradient_combine_device<T>(dims, y.get(), neg.get(), nullptr, row_ptr, col, val, 1.0, zpart.get(), nz, N, dY,
                                   s);
        Z = sum_device(zpart.get(), nz, &mb, s);
    } else {
        const int nz = exact_zblocks(N);
        zpart.alloc(nz);
        DevBuf<T> work(exact_work_elems(N, dims, false));
        exact_forces_device<T>(dims, Y, N, TKA_DD_TEXTBOOK, nullptr, 1.0, neg.get(), nullptr, zpart.get(), work.get(),
                               s);
        gradient_combine_device<T>(dims, Y, neg.get(), nullptr, row_ptr, col, val, 1.0, zpart.get(), nz, N, dY, s);
        Z = sum_device(zpart.get(), nz, &mb, s);
    }
    *kl = kl_edges_exact_device<T>(dims, Y, row_ptr, col, val, 1.0, Z, N, &mb, s);
    if (sum_q) *sum_q = Z;
}

}  // namespace
}  // namespace tka

extern "C" {

TKA_API tka_status tka_kl_grad_dev(const void* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                   const int32_t* col, const void* val, double theta, int32_t precision, double* kl,
                                   void* dY, double* sum_q, void* stream)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && Y && row_ptr && col && val && kl && dY,
                TKA_E_WRONG_PARAMETER, "bad arguments");
        require(theta >= 0.0, TKA_E_WRONG_PARAMETER, "theta must be non-negative");
        require(precision == TKA_FP32 || precision == TKA_FP64, TKA_E_WRONG_PARAMETER, "precision must be FP32 or FP64");
        Stream st(stream);
        AllocStream as(st.s);  // temporaries ordered on the caller's stream, which may be non-blocking
        if (precision == TKA_FP64)
            kl_grad_impl<double>(static_cast<const double*>(Y), N, no_dims, row_ptr, col,
                                 static_cast<const double*>(val), theta, kl, static_cast<double*>(dY), sum_q, st.s);
        else
            kl_grad_impl<float>(static_cast<const float*>(Y), N, no_dims, row_ptr, col, static_cast<const float*>(val),
                                theta, kl, static_cast<float*>(dY), sum_q, st.s);
    });
}

}  // extern "C"
