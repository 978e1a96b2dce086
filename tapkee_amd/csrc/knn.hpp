// Normalisation and brute-force kNN on the device.
#pragma once

#include "common.hpp"

namespace tka {

// zeroMean + division by the signed maximum coefficient, in place on a
// device N x D FP64 array (tsne.hpp:88-89, :488-518).  Column sums are
// accumulated row by row in index order, as the reference does, so the
// normalised data is bit-identical to a sequential FP64 computation.
void normalize_device(double* X, int N, int D, MemBudget* mb, hipStream_t s);

struct KnnStats {
    int certificate_failures = 0;  // rows whose error bound failed on the first pass
    int unresolved = 0;            // rows still failing after the wide second pass
    bool order_written = false;    // the locality order went to order_out
};

// Exact K nearest neighbours (self excluded) of the rows of Xn (device,
// N x D FP64), sorted by (FP64 squared distance, index).  The distance tiles
// run in `dtype` (TKA_DIST_FP32 / TKA_DIST_BF16 / TKA_DIST_FP8), keep
// `kprime` candidates per row, and are re-ranked in FP64.  order_out
// (optional, N): receives the rows' locality order (the Morton order of
// their projection onto the estimated top-3 principal subspace, in which
// rows that are near in the data are near in the order) when the tiles
// computed one; st->order_written says whether they did.
void knn_device(const double* Xn, int N, int D, int K, int dtype, int kprime, int* idx, double* dist,
                KnnStats* st, MemBudget* mb, hipStream_t s, uint32_t* order_out = nullptr);

}  // namespace tka
