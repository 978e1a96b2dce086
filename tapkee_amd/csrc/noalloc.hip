// Allocation-free variant of tka_gradient_dev (tka_gradient_dev_noalloc).
//
// Every device buffer the gradient needs beyond its inputs and outputs (the
// Barnes-Hut tree's keys, sort state and nodes, a centred copy of Y, the
// repulsive forces, the Z partials) is carved out of one file-scope device
// array, `scratch`, of fixed size: no hipMalloc, hipFree or allocating
// library call happens inside the entry point, so callers that cannot
// allocate (code under HIP graph capture, tight serving loops) can use it.
// Each GPU holds its own copy of the array, at its own address, until the
// process ends; that address is looked up once per device with
// hipGetSymbolAddress, and only the looked-up pointer is ever handed to
// kernels or copies (the array's name in host code is a host-side stand-in).
//
// The kernels and their order are those of tka_gradient_dev, so the results
// are the same bit for bit.  Calls are serialised by a process-wide lock and
// each call synchronises its stream before it returns: two calls of one
// process never use the array at the same time.
#include <mutex>

#include "bh.hpp"
#include "common.hpp"
#include "forces.hpp"

namespace tka {
namespace {

// Room for TKA_NOALLOC_MAX_POINTS points in 2-D at theta > 0 (the tree's
// buffers plus Y's copy, the forces and the partials: about 140 bytes per
// point and the sort's tile state); checked against the carve below.
constexpr size_t kScratchBytes = (size_t)20 << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on the first call on
// that device, with that device current).
char* device_scratch()
{
    int dev = 0;
    TKA_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < kMaxDevices, TKA_E_DEVICE, "device ordinal out of range");
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            fail(TKA_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = p;
    }
    return static_cast<char*>(g_scratch_ptr[dev]);
}

struct Views {
    DevBuf<float> y, neg, work;
    DevBuf<double> zpart, box, zsum;
    BhTree<float> tree;
};

// The layout of the region for one call; with a null arena base it only
// measures it.
void carve(int N, int dims, bool bh, Arena& a, Views& v)
{
    const size_t nd = (size_t)N * dims;
    a.take(v.neg, nd);
    a.take(v.zsum, 1);
    if (bh) {
        a.take(v.y, nd);
        a.take(v.zpart, bh_zblocks(N));
        a.take(v.box, (size_t)kBoxStride * bbox_blocks(N));
        v.tree.carve(N, a);
    } else {
        a.take(v.zpart, exact_zblocks(N));
        a.take(v.work, exact_work_elems(N, dims, false));
    }
}

}  // namespace
}  // namespace tka

using namespace tka;

extern "C" {

TKA_API tka_status tka_gradient_dev_noalloc(const float* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                            const int32_t* col, const float* val, double theta, float* dY,
                                            double* sum_q, void* stream)
{
    return guard([&] {
        require(N >= 2 && no_dims >= 1 && no_dims <= 3 && Y && row_ptr && col && val && dY, TKA_E_WRONG_PARAMETER,
                "bad arguments");
        require(theta >= 0.0, TKA_E_WRONG_PARAMETER, "theta must be non-negative");
        require(N <= TKA_NOALLOC_MAX_POINTS, TKA_E_WRONG_PARAMETER,
                "tka_gradient_dev_noalloc: more than TKA_NOALLOC_MAX_POINTS (" +
                    std::to_string(TKA_NOALLOC_MAX_POINTS) + ") points");
        const bool bh = theta > 0.0;
        if (bh) require(no_dims == 2, TKA_E_WRONG_PARAMETER, "Barnes-Hut gradient needs 2 dimensions");
        std::lock_guard<std::mutex> lk(g_mu);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Views v;
        {
            Arena probe;  // the region must hold the largest call it accepts
            Views w;
            carve(TKA_NOALLOC_MAX_POINTS, 2, true, probe, w);
            require(probe.used <= kScratchBytes, TKA_E_INTERNAL, "no-alloc scratch region too small");
            Arena probe0;
            Views w0;
            carve(TKA_NOALLOC_MAX_POINTS, 3, false, probe0, w0);
            require(probe0.used <= kScratchBytes, TKA_E_INTERNAL, "no-alloc scratch region too small");
        }
        Arena a;
        a.base = device_scratch();
        a.cap = kScratchBytes;
        carve(N, no_dims, bh, a, v);
        const size_t nd = (size_t)N * no_dims;
        if (bh) {
            // the tree build centres its input: a copy of Y in the region
            TKA_HIP(hipMemcpyAsync(v.y.get(), Y, nd * sizeof(float), hipMemcpyDeviceToDevice, s));
            v.tree.clear_sort_state(s);
            bbox_partials_device<float>(v.y.get(), N, v.box.get(), s);
            bh_build<float>(v.y.get(), N, v.box.get(), bbox_blocks(N), v.tree, s);
            bh_repulsion<float>(v.tree, v.y.get(), theta, v.neg.get(), v.zpart.get(), nullptr, s);
            gradient_combine_device<float>(2, v.y.get(), v.neg.get(), nullptr, row_ptr, col, val, 1.0,
                                           v.zpart.get(), bh_zblocks(N), N, dY, s);
            sum_into_device(v.zpart.get(), bh_zblocks(N), v.zsum.get(), s);
        } else {
            const int nz = exact_forces_device<float>(no_dims, Y, N, TKA_DD_TEXTBOOK, nullptr, 1.0, v.neg.get(),
                                                      nullptr, v.zpart.get(), v.work.get(), s);
            gradient_combine_device<float>(no_dims, Y, v.neg.get(), nullptr, row_ptr, col, val, 1.0, v.zpart.get(),
                                           nz, N, dY, s);
            sum_into_device(v.zpart.get(), nz, v.zsum.get(), s);
        }
        double z = 0.0;
        TKA_HIP(hipMemcpyAsync(&z, v.zsum.get(), sizeof(double), hipMemcpyDeviceToHost, s));
        TKA_HIP(hipStreamSynchronize(s));
        if (bh) require(!bh_sort_failed(v.tree, s), TKA_E_INTERNAL, "radix sort look-back gave up; the tree is not valid");
        if (sum_q) *sum_q = z;
    });
}

}  // extern "C"
