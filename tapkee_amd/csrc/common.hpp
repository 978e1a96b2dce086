// Shared helpers for the MI355X (gfx950) t-SNE engine: error mapping, device
// buffers, wave64 reductions.  Host and device code are compiled by hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/tapkee_amd.h"

namespace tka {

// Exception carrying a tka_status; the C-ABI layer converts it to a return
// code plus tka_last_error() text.
struct Error : std::runtime_error {
    tka_status status;
    Error(tka_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void throw_hip(hipError_t e, const char* what, const char* file, int line)
{
    const tka_status s = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? TKA_E_NOT_ENOUGH_MEMORY
                                                                                      : TKA_E_DEVICE;
    throw Error(s, std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")");
}

#define TKA_HIP(x)                                                          \
    do {                                                                    \
        hipError_t tka_e_ = (x);                                            \
        if (tka_e_ != hipSuccess) ::tka::throw_hip(tka_e_, #x, __FILE__, __LINE__); \
    } while (0)

#define TKA_LAUNCH_CHECK() TKA_HIP(hipGetLastError())

[[noreturn]] inline void fail(tka_status s, const std::string& m) { throw Error(s, m); }

inline void require(bool ok, tka_status s, const std::string& m)
{
    if (!ok) fail(s, m);
}

// The engine's messages (the reference's Logging singleton,
// include/tapkee/utils/logging.hpp:89-144): delivered to the installed
// tka_log_fn, or written to stderr as "[level] message", when the level is
// enabled or `force` is set (a call's verbose option).
void log_message(tka_log_level level, const std::string& msg, bool force = false);

// "<operation> took <seconds> seconds." at the benchmark level, as the
// reference's timed_context (include/tapkee/utils/time.hpp:29-42).
void log_took(const char* operation, double seconds, bool force = false);

// Message of the last failed entry point on this thread (tka_last_error).
std::string& last_error();

// Runs an entry point's body: exceptions become a tka_status and the message
// of tka_last_error().
template <class F>
tka_status guard(F&& f)
{
    try {
        f();
        last_error().clear();
        return TKA_OK;
    } catch (const Error& e) {
        last_error() = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        last_error() = "Not enough memory";
        return TKA_E_NOT_ENOUGH_MEMORY;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return TKA_E_INTERNAL;
    }
}

// Device memory comes from a process-wide cache of blocks (engine.hip):
// freeing a block records an event on its allocation stream and keeps the
// block; an allocation of the same size class on the same device reuses a
// cached block whose event has completed, else calls hipMalloc.  A hipFree of
// a large block costs hundreds of microseconds (and synchronises the device);
// the setup of a t-SNE run allocates and frees sort buffers of hundreds of
// MB, so repeated runs in one process would pay that on every phase.
void* dev_alloc(size_t bytes, size_t* block_bytes, int* device);
void dev_free(void* p, size_t block_bytes, int device, hipStream_t stream) noexcept;

// Stream whose work a DevBuf's block must outlive (its free records the
// reuse event there): the null stream -- ordered after all work on the
// engine's blocking streams -- unless an entry point that works on a
// caller's stream, which may be non-blocking, names that stream for its
// scope.
inline hipStream_t& alloc_stream()
{
    static thread_local hipStream_t s = nullptr;
    return s;
}
struct AllocStream {
    hipStream_t prev;
    explicit AllocStream(hipStream_t s) : prev(alloc_stream()) { alloc_stream() = s; }
    ~AllocStream() { alloc_stream() = prev; }
    AllocStream(const AllocStream&) = delete;
    AllocStream& operator=(const AllocStream&) = delete;
};

// Owning device allocation.  Allocation size is tracked against an optional
// cap so tests can force the out-of-memory path from the host.
struct MemBudget {
    int64_t cap = 0;   // 0 = unlimited
    int64_t used = 0;
};

template <class T>
class DevBuf {
  public:
    DevBuf() = default;
    DevBuf(size_t n, MemBudget* b = nullptr) { alloc(n, b); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { swap(o); }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            swap(o);
        }
        return *this;
    }
    void alloc(size_t n, MemBudget* b = nullptr)
    {
        release();
        const int64_t bytes = (int64_t)(n ? n : 1) * (int64_t)sizeof(T);
        if (b && b->cap > 0 && b->used + bytes > b->cap)
            fail(TKA_E_NOT_ENOUGH_MEMORY, "device allocation of " + std::to_string(bytes) +
                                              " bytes exceeds the configured cap");
        stream_ = alloc_stream();
        p_ = static_cast<T*>(dev_alloc((size_t)bytes, &block_, &dev_));
        n_ = n;
        budget_ = b;
        bytes_ = bytes;
        if (b) b->used += bytes;
    }
    // Non-owning view of n elements at p (memory the caller keeps alive);
    // release() forgets it without freeing.
    void view(T* p, size_t n)
    {
        release();
        p_ = p;
        n_ = n;
        block_ = 0;
    }
    void release()
    {
        if (p_ && block_ != 0) {
            dev_free(p_, block_, dev_, stream_);
            if (budget_) budget_->used -= bytes_;
        }
        p_ = nullptr;
        n_ = 0;
        budget_ = nullptr;
        bytes_ = 0;
    }
    T* get() const { return p_; }
    operator T*() const { return p_; }
    size_t size() const { return n_; }

  private:
    void swap(DevBuf& o)
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        std::swap(budget_, o.budget_);
        std::swap(bytes_, o.bytes_);
        std::swap(stream_, o.stream_);
        std::swap(block_, o.block_);
        std::swap(dev_, o.dev_);
    }
    T* p_ = nullptr;
    hipStream_t stream_ = nullptr;
    size_t block_ = 0;
    int dev_ = 0;
    size_t n_ = 0;
    MemBudget* budget_ = nullptr;
    int64_t bytes_ = 0;
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Copies `bytes` (at most 256) of device memory to the host after the work
// queued on s, through a pinned per-thread buffer (a pageable destination
// makes the runtime stage the copy), and waits for it.
void read_to_host(void* dst, const void* src, size_t bytes, hipStream_t s);

// Buffers carved one after another (256-byte aligned) out of one region that
// the caller owns.  With a null base nothing is carved and `used` counts the
// bytes a region would need.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    template <class T>
    void take(DevBuf<T>& b, size_t n)
    {
        used = (used + 255) & ~(size_t)255;
        const size_t bytes = (n ? n : 1) * sizeof(T);
        if (base) {
            require(used + bytes <= cap, TKA_E_INTERNAL, "scratch region too small");
            b.view(reinterpret_cast<T*>(base + used), n);
        }
        used += bytes;
    }
};

// Compute units of the current device (256 on MI355X).
int device_cus();

// ---------------------------------------------------------------- device side
#ifdef __HIP_DEVICE_COMPILE__
#define TKA_DEV __device__ __forceinline__
#else
#define TKA_DEV __device__ inline
#endif

constexpr int kWave = 64;
constexpr int kBoxStride = 8;  // doubles per block of Y statistics

// FP64 multiply and add that the compiler never fuses into an FMA, so a sum
// of products rounds exactly like a scalar loop compiled without contraction
// (the reference's euclidean_distance, vptree.hpp:95-101).
TKA_DEV double dmul_sep(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}
TKA_DEV double dadd_sep(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}

template <class T>
TKA_DEV T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
TKA_DEV T wave_min(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

template <class T>
TKA_DEV T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Block-wide sum of a double for blockDim.x <= 1024 (multiple of 64).
// Result valid in every thread.  `scratch` needs blockDim.x/64 doubles.
TKA_DEV double block_sum(double v, double* scratch)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; i++) t += scratch[i];
    return t;
}

}  // namespace tka
