// tapkee's nearest-neighbour search (include/tapkee/neighbors/neighbors.hpp:
// 119-216) on the MI355X engine, for feature vectors under the Euclidean
// distance.
//
//   #include <tapkee_amd/neighbors.hpp>
//   tapkee::tapkee_internal::Neighbors nb =
//       tapkee::tapkee_internal::find_neighbors(tapkee::Brute, features, k, true);
//
// features: D x N column-major (one column per object), as tapkee's
// DenseMatrix.  Every method (Brute, VpTree, CoverTree) is answered by the
// GPU's exact brute force (tka_knn: distance tiles on the matrix cores, an
// FP64 re-rank, ties broken by index), so each object's list holds its k
// nearest other objects, nearest first.  The reference's contract is kept:
// k above N-1 is lowered to N-1 with its warning, and with
// check_connectivity a neighbour graph that is not connected is recomputed
// with 2k neighbours (neighbors.hpp:196-205), with the same log messages.
// Limit of the engine: k <= 384.
#pragma once

#include <string>
#include <vector>

#include "tapkee.hpp"

namespace tapkee {

// include/tapkee/defines/methods.hpp:150-173
struct NeighborsMethod {
    explicit NeighborsMethod(const char* n) : name_(n) {}
    bool operator==(const NeighborsMethod& o) const { return name_ == o.name_; }
    bool is(const NeighborsMethod& o) const { return *this == o; }
    std::string name() const { return name_; }
    std::string name_;
};
static const NeighborsMethod Brute("Brute-force");
static const NeighborsMethod VpTree("Vantage point tree");
static const NeighborsMethod CoverTree("Cover tree");

namespace tapkee_internal {

typedef std::vector<IndexType> LocalNeighbors;
typedef std::vector<LocalNeighbors> Neighbors;

// connected.hpp:19-61 (tka_neighbors_connected: DFS from object 0)
inline bool is_connected(const Neighbors& neighbors)
{
    const IndexType N = (IndexType)neighbors.size();
    if (N == 0) return true;
    const IndexType k = (IndexType)neighbors[0].size();
    std::vector<int32_t> flat((size_t)N * k);
    for (IndexType i = 0; i < N; i++)
        for (IndexType j = 0; j < k; j++) flat[(size_t)i * k + j] = neighbors[i][j];
    int32_t connected = 0;
    throw_status(tka_neighbors_connected(flat.data(), N, k, &connected));
    return connected != 0;
}

template <class Matrix>
Neighbors find_neighbors(const NeighborsMethod& method, const Matrix& features, IndexType k, bool check_connectivity,
                         int32_t dist_dtype = TKA_DIST_BF16)
{
    const IndexType N = (IndexType)features.cols(), D = (IndexType)features.rows();
    if (N == 0) throw no_data_error();
    if (k > N - 1) {
        Logging::instance().message_warning("Number of neighbors is greater than number of objects to embed. "
                                            "Using greatest possible number of neighbors.");
        k = N - 1;
    }
    Logging::instance().message_info("Using the " + method.name() + " neighbors computation method.");
    std::vector<int32_t> idx((size_t)N * (k > 0 ? k : 1));
    std::vector<double> dist(idx.size());
    int32_t fails = 0;
    throw_status(tka_knn(features.data(), N, D, k, dist_dtype, 0, idx.data(), dist.data(), &fails));
    Neighbors neighbors((size_t)N);
    for (IndexType i = 0; i < N; i++) neighbors[i].assign(idx.begin() + (size_t)i * k, idx.begin() + (size_t)(i + 1) * k);
    if (check_connectivity && !is_connected(neighbors)) {
        Logging::instance().message_warning("The neighborhood graph with " + std::to_string(k) +
                                            " neighbors is not connected. Recomputing with a larger number of "
                                            "neighbors " + std::to_string(2 * k) + ".");
        return find_neighbors(method, features, 2 * k, check_connectivity, dist_dtype);
    }
    Logging::instance().message_info("The neighborhood graph with " + std::to_string(k) + " neighbors is connected.");
    return neighbors;
}

}  // namespace tapkee_internal
}  // namespace tapkee
