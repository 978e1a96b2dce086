// Attractive forces of CSR rows (quadtree.hpp:357-384), shared by the
// standalone attraction launch and by the tree-build kernels that run rows
// of it in workgroups past their own grid:
//   pos_i = sum_e exag*P_e (y_i - y_col(e)) / (1 + |y_i - y_col(e)|^2).
#pragma once

#include "common.hpp"

namespace tka {

constexpr int kG = 16;                 // lanes per CSR row
constexpr int kRowsPerBlock = 256 / kG;

// Sum over the kG lanes of a row group (lanes g*kG .. g*kG+kG-1).
template <class T>
TKA_DEV T group_sum(T v)
{
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Each lane takes its row's edges e, e+kG, e+2kG, ..., loading kBatch column
// indices, then kBatch rows of Y, before the arithmetic, so a lane waits on
// two memory latencies per kBatch edges rather than per edge.  Padding
// entries use the row itself with P = 0 and add exactly zero.
constexpr int kBatch = 4;

template <class T, int DIMS>
TKA_DEV void attract_row(const T* __restrict__ Y, const int* __restrict__ row_ptr, const int* __restrict__ col,
                         const T* __restrict__ val, T exag, int i, int gl, T* __restrict__ pos)
{
    T yi[DIMS], acc[DIMS];
#pragma unroll
    for (int d = 0; d < DIMS; d++) {
        yi[d] = Y[(size_t)i * DIMS + d];
        acc[d] = (T)0;
    }
    const int e1 = row_ptr[i + 1];
    for (int base = row_ptr[i] + gl; base < e1; base += kBatch * kG) {
        int j[kBatch];
        T v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int e = base + u * kG;
            const bool ok = e < e1;
            j[u] = ok ? col[e] : i;
            v[u] = ok ? val[e] : (T)0;
        }
        T yj[kBatch][DIMS];
#pragma unroll
        for (int u = 0; u < kBatch; u++)
#pragma unroll
            for (int d = 0; d < DIMS; d++) yj[u][d] = Y[(size_t)j[u] * DIMS + d];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            T diff[DIMS];
            T D = (T)0;
#pragma unroll
            for (int d = 0; d < DIMS; d++) {
                diff[d] = yi[d] - yj[u][d];
                D += diff[d] * diff[d];
            }
            const T m = (exag * v[u]) / ((T)1 + D);
#pragma unroll
            for (int d = 0; d < DIMS; d++) acc[d] += m * diff[d];
        }
    }
#pragma unroll
    for (int d = 0; d < DIMS; d++) acc[d] = group_sum(acc[d]);
    if (gl == 0)
#pragma unroll
        for (int d = 0; d < DIMS; d++) pos[(size_t)i * DIMS + d] = acc[d];
}


// Rows [row0, row1) of the attraction of a DIMS-D embedding, run as extra
// work by workgroups of 256 threads of another kernel: workgroup wg of nwg
// takes row groups wg, wg + nwg, ...  Each row is computed exactly as by the
// standalone launch, so the result does not depend on which kernel ran it.
template <class T, int DIMS = 2>
struct AttractSlice {
    const T* Y = nullptr;
    const int* row_ptr = nullptr;
    const int* col = nullptr;
    const T* val = nullptr;
    T exag = (T)0;
    T* pos = nullptr;
    int row0 = 0, row1 = 0;

    TKA_DEV void run(int wg, int nwg) const
    {
        const int g = threadIdx.x / kG, gl = threadIdx.x % kG;
        for (int r = row0 + wg * kRowsPerBlock; r < row1; r += nwg * kRowsPerBlock) {
            const int i = r + g;
            if (i < row1)  // whole row groups skip together; group sums stay inside a group
                attract_row<T, DIMS>(Y, row_ptr, col, val, exag, i, gl, pos);
        }
    }
};

}  // namespace tka
