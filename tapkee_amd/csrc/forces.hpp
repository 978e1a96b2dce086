// Gradient pieces other than the Barnes-Hut traversal: the CSR attractive
// pass fused with the optimiser step, the exact O(N^2) repulsion for
// theta = 0, and the KL reductions.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace tka {

struct StepParams {
    double momentum;
    double eta;
    double exag;        // multiplies P (12 while "lying", then 1)
};

// One optimiser step for all rows (tsne.hpp:152-167, attraction
// quadtree.hpp:357-384, combine tsne.hpp:344-347):
//   pos_i = sum_e exag*P_e * (y_i - y_col)/(1 + |y_i - y_col|^2)   (CSR), or
//           pos_pre (dense theta = 0 form) when row_ptr == nullptr
//   dC    = pos - neg / Z,  Z = sum(zpart[0..nz))
//   gains, uY, Ynext = Y + uY.
// Writes box partials of Ynext (kBoxStride doubles per block) and returns the number
// of partial blocks.  dC_out (optional) receives dC.
template <class T>
int tsne_step_device(int dims, const T* Y, T* Ynext, T* uY, T* gains, const T* neg, const T* pos_pre,
                     const int* row_ptr, const int* col, const T* val, const double* zpart, int nz, int N,
                     StepParams sp, double* box_part, T* dC_out, hipStream_t s);
int tsne_step_blocks(int N);

// The same step in two parts, so that the attraction can run inside the tree
// build's launches (AttractSlice, bh_build_tree) or alone: pos (N x dims) from
// the CSR, then the update from pos, neg and the zpart partials of Z.  The update returns the number of
// box partials it wrote (tsne_update_blocks(N), a few per thousand points, so
// that the next tree build can reduce them cheaply).
template <class T>
void attraction_device(int dims, const T* Y, const int* row_ptr, const int* col, const T* val, double exag, int N,
                       T* pos, hipStream_t s);
template <class T>
int tsne_update_device(int dims, const T* Y, T* Ynext, T* uY, T* gains, const T* neg, const T* pos,
                       const double* zpart, int nz, int N, StepParams sp, double* box_part, T* dC_out, hipStream_t s);
int tsne_update_blocks(int N);

// zeroMean of Y (N x dims) from the box partials of the step that produced it.
template <class T>
void center_device(int dims, T* Y, int N, const double* box_part, int nb, hipStream_t s);

// zeroMean of Y (N x dims) computed from Y itself.
template <class T>
void center_fresh_device(int dims, T* Y, int N, MemBudget* mb, hipStream_t s);

// Gradient only (no update): dC = pos - neg/Z.
template <class T>
void gradient_combine_device(int dims, const T* Y, const T* neg, const T* pos_pre, const int* row_ptr,
                             const int* col, const T* val, double exag, const double* zpart, int nz, int N, T* dC,
                             hipStream_t s);

// Grid of the all-pairs kernels: row blocks of 256 rows times column slices
// of `cols` columns, about kExactTargetWgs workgroups in all (every CU busy
// at N = 2000 and N = 10000 alike), slices of at least kExactMinCols columns.
constexpr int kExactTargetWgs = 2048;
constexpr int kExactMinCols = 128;
struct ExactGrid {
    int row_blocks, slices, cols;
};
inline ExactGrid exact_grid(int N)
{
    const int rb = (int)ceil_div(N, 256);
    int slices = std::min((int)ceil_div(N, kExactMinCols), std::max(1, (int)ceil_div(kExactTargetWgs, rb)));
    slices = std::max(1, slices);
    const int cols = (int)ceil_div(N, slices);
    return {rb, (int)ceil_div(N, cols), cols};
}
// Z partials of the all-pairs kernels (one per workgroup).
inline int exact_zblocks(int N)
{
    const ExactGrid g = exact_grid(N);
    return g.row_blocks * g.slices;
}
// Elements of T the all-pairs forces need as workspace (per-slice row sums).
inline size_t exact_work_elems(int N, int dims, bool dense)
{
    return (size_t)exact_grid(N).slices * N * dims * (dense ? 2 : 1);
}

// Exact repulsion over all pairs (tsne.hpp:353-407): neg_i = sum_j q^2 (y_i -
// y_j), zpart = per-workgroup sums of q (exact_zblocks(N) of them).  With a
// dense symmetric P (N x N) also pos_i = sum_j exag*P_ij q (y_i - y_j).
// dd_mode as tka_dd_mode.  work: exact_work_elems(N, dims, P != nullptr).
template <class T>
int exact_forces_device(int dims, const T* Y, int N, int dd_mode, const T* P, double exag, T* neg, T* pos,
                        double* zpart, T* work, hipStream_t s);

// Sum of a double array (deterministic order); result copied to the host.
double sum_device(const double* v, int n, MemBudget* mb, hipStream_t s);
// The same sum (same order) into out[0] on the device, asynchronously.
void sum_into_device(const double* v, int n, double* out, hipStream_t s);

// KL over CSR edges as tsne.hpp:454-486: sum p*log((p + FLT_MIN)/(q/Z + FLT_MIN)),
// p = exag*val.  Z given.
template <class T>
double kl_edges_reference_device(int dims, const T* Y, const int* row_ptr, const int* col, const T* val,
                                 double exag, double Z, int N, MemBudget* mb, hipStream_t s);

// Exact Z (FP64, all pairs) and KL over CSR edges with it: sum p log(p Z (1+d^2)).
template <class T>
double exact_z_device(int dims, const T* Y, int N, MemBudget* mb, hipStream_t s);
template <class T>
double kl_edges_exact_device(int dims, const T* Y, const int* row_ptr, const int* col, const T* val, double exag,
                             double Z, int N, MemBudget* mb, hipStream_t s);

// Renumbering of the points for memory locality: new index j holds old
// point sorted[j] (the identity instead when *sort_fault is set: the sort
// that produced `sorted` gave up).  Gathers the per-point state (dims values
// each) and rebuilds the CSR with rows and columns in the new numbering;
// orig maps new indices to the caller's original indices and is permuted
// alongside.
template <class T>
void renumber_points_device(int N, int dims, const uint32_t* sorted, const uint32_t* sort_fault, T* const* state,
                            int nstate, T* tmp, const int* row, const int* col, const T* val, int* row2, int* col2,
                            T* val2, int* orig, int* orig_tmp, int* inv, uint32_t* scan_scratch, hipStream_t s);
size_t renumber_scan_scratch_words(int N);

// Y_out[orig[j]] = Y[j] (N x dims).
template <class T>
void unpermute_rows_device(int N, int dims, const T* Y, const int* orig, T* out, hipStream_t s);

// Dense-P KL as tsne.hpp:409-452 (FP64).
template <class T>
double kl_dense_device(int dims, const T* Y, const T* P, double exag, int N, int dd_mode, MemBudget* mb,
                       hipStream_t s);

}  // namespace tka
