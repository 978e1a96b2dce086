// Input similarities on the device: FP64 perplexity calibration and the
// symmetric, normalised P (CSR for the kNN form, dense for theta = 0).
#pragma once

#include "common.hpp"

namespace tka {

// Row-wise bisection on beta for kNN rows (tsne.hpp:648-711).  dist/val:
// N x K.  beta/iters may be null.  Returns the number of rows that hit the
// 200-evaluation cap.
int perplexity_knn_device(const double* dist, int N, int K, double perplexity, double* val, double* beta,
                          int* iters, MemBudget* mb, hipStream_t s);

// Pairwise squared distances (N x N) with the tka_dd_mode convention.
void sqdist_dense_device(const double* X, int N, int D, int dd_mode, double* DD, hipStream_t s);

// Dense rows over all points (tsne.hpp:520-599); DD: N x N, P: N x N.
int perplexity_dense_device(const double* DD, int N, double perplexity, double* P, MemBudget* mb, hipStream_t s);

// Dense symmetrisation and normalisation (tsne.hpp:100-108).
void symmetrize_dense_device(double* P, int N, MemBudget* mb, hipStream_t s);

// CSR symmetrisation (tsne.hpp:207-322): (P + P^T)/2 over the union
// pattern, columns ascending in each row.  Output buffers are allocated here.
// order (optional, N entries): a locality order of the points in which to
// visit the rows; the result does not depend on it.
struct Csr {
    DevBuf<int> row;     // N+1
    DevBuf<int> col;     // nnz
    DevBuf<double> val;  // nnz
    int64_t nnz = 0;
};
void symmetrize_knn_device(const int* idx, const double* val, int N, int K, Csr& out, MemBudget* mb, hipStream_t s,
                           const uint32_t* order = nullptr);

// Divides a device array by the FP64 sum of its entries (tsne.hpp:120-124).
void normalize_sum_device(double* v, int64_t n, MemBudget* mb, hipStream_t s);

}  // namespace tka
