// FP64 CPU restatement of tapkee's t-SNE path -- TEST INFRASTRUCTURE ONLY.
//
// This file is the oracle that the HIP engine is checked against.  It is not
// part of the product: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it (through oracle/oracle.py), and only as the checker
// or as the timed CPU baseline.
//
// It restates, from the algorithm (not from the source text), the code in
// tapkee 1.3.4:
//   include/tapkee/external/barnes_hut_sne/tsne.hpp      (TSNE::run and helpers)
//   include/tapkee/external/barnes_hut_sne/quadtree.hpp  (pointer quadtree)
//   include/tapkee/external/barnes_hut_sne/vptree.hpp    (VP-tree kNN)
//   include/tapkee/defines/random.hpp                    (std::rand based RNG)
// Each function cites the lines it follows.  Everything is FP64, single
// threaded, and compiled with -ffp-contract=off so that sums of products round
// like the reference's separate multiply/add.
//
// Parity pins (what checks each function, see tests/test_oracle.py):
//   zero_mean/normalize      closed form (numpy)
//   sqdist (both modes)      closed form; compat mode == -2*X X^T
//   knn_brute                scipy cdist + stable argsort
//   vptree_knn               subset-of-brute-force + distances exact; recall
//   perplexity_sparse/dense  row sums, entropy vs log(perp), brentq root of H
//   symmetrize               scipy.sparse (P + P^T)/2
//   quadtree                 cell COM/count vs numpy cell membership
//   gradient_exact(_sparse)  torch.autograd of KL (== 4 * dY) in FP64
//   gradient_bh              theta=0 equals exact; error shrinks with theta
//   kl_*                     numpy closed form; central finite differences
//   update                   scalar recurrence in numpy
//   gaussian_random/rand     glibc rand() known-answer sequence (seed 1)
#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <vector>

#define OR_API extern "C" __attribute__((visibility("default")))

// ---------------------------------------------------------------------------
// RNG: defines/random.hpp:29-54
// ---------------------------------------------------------------------------
static double or_uniform_random_impl() { return std::rand() / ((double)RAND_MAX + 1); }

OR_API double or_uniform_random(void) { return or_uniform_random_impl(); }

// Marsaglia polar method, returns only x (random.hpp:38-54).
OR_API double or_gaussian_random(void)
{
    double x, y, r;
    do {
        x = 2 * or_uniform_random_impl() - 1;
        y = 2 * or_uniform_random_impl() - 1;
        r = x * x + y * y;
    } while (r >= 1.0 || r == 0.0);
    r = std::sqrt(-2 * std::log(r) / r);
    return x * r;
}

OR_API void or_srand(unsigned seed) { std::srand(seed); }
OR_API int or_rand(void) { return std::rand(); }

// Number of std::rand() calls the VP-tree build makes for n points: one per
// node whose range has more than one item; children are [lower+1, median)
// and [median, upper) with median = (lower+upper)/2 (vptree.hpp:211-245).
OR_API long long or_vptree_rand_calls(long long n)
{
    // iterative to avoid deep recursion on large n
    long long calls = 0;
    std::vector<long long> stack;
    stack.push_back(n);
    while (!stack.empty()) {
        long long m = stack.back();
        stack.pop_back();
        if (m <= 1) continue;
        calls++;
        long long half = m / 2;
        stack.push_back(half - 1);
        stack.push_back(m - half);
    }
    return calls;
}

// ---------------------------------------------------------------------------
// Normalisation: tsne.hpp:488-518 (zeroMean) and :88-89 (divide by max coeff)
// ---------------------------------------------------------------------------
OR_API void or_zero_mean(double* X, int N, int D)
{
    std::vector<double> mean(D, 0.0);
    for (int n = 0; n < N; n++)
        for (int d = 0; d < D; d++) mean[d] += X[(size_t)n * D + d];
    for (int d = 0; d < D; d++) mean[d] /= (double)N;
    for (int n = 0; n < N; n++)
        for (int d = 0; d < D; d++) X[(size_t)n * D + d] -= mean[d];
}

OR_API void or_normalize(double* X, int N, int D)
{
    or_zero_mean(X, N, D);
    double mx = -DBL_MAX;
    for (size_t i = 0; i < (size_t)N * D; i++) mx = std::max(mx, X[i]);
    for (size_t i = 0; i < (size_t)N * D; i++) X[i] /= mx;
}

// ---------------------------------------------------------------------------
// Squared distances.  tsne.hpp:927-955.  mode 0 = textbook ||x-y||^2 (what the
// commented cblas_dgemm with beta=1 intends); mode 1 = tapkee_compat, which
// reproduces the reference's overwrite: DD = -2 * X^T X (:948-950 assigns over
// the norm sums written at :941-947).
// ---------------------------------------------------------------------------
static inline double sqdist_rows(const double* a, const double* b, int D)
{
    double dd = 0.0;  // vptree.hpp:95-101 summation order
    for (int d = 0; d < D; d++) {
        double t = a[d] - b[d];
        dd += t * t;
    }
    return dd;
}

OR_API void or_sqdist(const double* X, int N, int D, double* DD, int compat)
{
    for (int n = 0; n < N; n++) {
        for (int m = 0; m < N; m++) {
            if (compat) {
                double g = 0.0;
                for (int d = 0; d < D; d++) g += X[(size_t)n * D + d] * X[(size_t)m * D + d];
                DD[(size_t)n * N + m] = -2.0 * g;
            } else {
                DD[(size_t)n * N + m] = sqdist_rows(X + (size_t)n * D, X + (size_t)m * D, D);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// kNN.  Brute force ground truth: K nearest by (squared distance, index),
// self excluded.  This is the exact answer tapkee's VP-tree approximates
// (tsne.hpp:628-646 searches K+1 and drops the first hit as "self").
// ---------------------------------------------------------------------------
static void knn_brute_row(const double* X, int N, int D, int K, int n, std::vector<std::pair<double, int>>& row,
                          int* idx, double* dist)
{
    int c = 0;
    for (int m = 0; m < N; m++) {
        if (m == n) continue;
        row[c++] = std::make_pair(sqdist_rows(X + (size_t)n * D, X + (size_t)m * D, D), m);
    }
    std::partial_sort(row.begin(), row.begin() + K, row.begin() + c);
    for (int k = 0; k < K; k++) {
        idx[k] = row[k].second;
        dist[k] = row[k].first;
    }
}

OR_API void or_knn_brute(const double* X, int N, int D, int K, int* idx, double* dist)
{
    std::vector<std::pair<double, int>> row(N > 0 ? N - 1 : 0);
    for (int n = 0; n < N; n++) knn_brute_row(X, N, D, K, n, row, idx + (size_t)n * K, dist + (size_t)n * K);
}

// The same for the query rows rows[0..nrows) only (output row r belongs to
// query rows[r]): exact neighbours of a sample of a large problem.
OR_API void or_knn_brute_rows(const double* X, int N, int D, int K, const int* rows, int nrows, int* idx,
                              double* dist)
{
    std::vector<std::pair<double, int>> row(N > 0 ? N - 1 : 0);
    for (int r = 0; r < nrows; r++)
        knn_brute_row(X, N, D, K, rows[r], row, idx + (size_t)r * K, dist + (size_t)r * K);
}

// VP-tree restatement (vptree.hpp:103-293).  Items are point indices into X;
// the reference deep-copies DataPoint objects but orders and compares them
// identically, so the partitions (std::nth_element with the same comparator)
// and the heap order (std::priority_queue on distance) are the same.
namespace {
struct VpNode {
    int index = 0;
    double threshold = 0.0;
    VpNode* left = nullptr;
    VpNode* right = nullptr;
    ~VpNode()
    {
        delete left;
        delete right;
    }
};

struct VpHeapItem {
    int index;
    double dist;
    bool operator<(const VpHeapItem& o) const { return dist < o.dist; }
};

struct VpTree {
    const double* X;
    int D;
    std::vector<int> items;  // original point index per slot
    VpNode* root = nullptr;
    double tau = 0.0;

    double dist(int a, int b) const { return sqdist_rows(X + (size_t)a * D, X + (size_t)b * D, D); }

    VpNode* build(int lower, int upper)  // vptree.hpp:211-245
    {
        if (upper == lower) return nullptr;
        VpNode* node = new VpNode();
        node->index = lower;
        if (upper - lower > 1) {
            int i = (int)(or_uniform_random_impl() * (upper - lower - 1)) + lower;
            std::swap(items[lower], items[i]);
            int median = (upper + lower) / 2;
            const int vp = items[lower];
            std::nth_element(items.begin() + lower + 1, items.begin() + median, items.begin() + upper,
                             [&](int a, int b) { return dist(vp, a) < dist(vp, b); });
            node->threshold = dist(items[lower], items[median]);
            node->left = build(lower + 1, median);
            node->right = build(median, upper);
        }
        return node;
    }

    void search(VpNode* node, int target, int k, std::priority_queue<VpHeapItem>& heap)  // :248-293
    {
        if (node == nullptr) return;
        double d = dist(items[node->index], target);
        if (d < tau) {
            if (heap.size() == (size_t)k) heap.pop();
            heap.push(VpHeapItem{node->index, d});
            if (heap.size() == (size_t)k) tau = heap.top().dist;
        }
        if (node->left == nullptr && node->right == nullptr) return;
        if (d < node->threshold) {
            search(node->left, target, k, heap);
            if (d + tau >= node->threshold) search(node->right, target, k, heap);
        } else {
            search(node->right, target, k, heap);
            if (d - tau <= node->threshold) search(node->left, target, k, heap);
        }
    }
};
}  // namespace

// Searches K+1 neighbours per point and drops the first (tsne.hpp:646,663,709).
// Consumes std::rand() exactly like the reference build.
OR_API void or_vptree_knn(const double* X, int N, int D, int K, int* idx, double* dist)
{
    VpTree t;
    t.X = X;
    t.D = D;
    t.items.resize(N);
    for (int n = 0; n < N; n++) t.items[n] = n;
    t.root = t.build(0, N);
    std::vector<VpHeapItem> res;
    for (int n = 0; n < N; n++) {
        std::priority_queue<VpHeapItem> heap;
        t.tau = DBL_MAX;
        t.search(t.root, n, K + 1, heap);
        res.clear();
        while (!heap.empty()) {
            res.push_back(heap.top());
            heap.pop();
        }
        std::reverse(res.begin(), res.end());
        for (int k = 0; k < K; k++) {
            idx[(size_t)n * K + k] = t.items[res[k + 1].index];
            dist[(size_t)n * K + k] = res[k + 1].dist;
        }
    }
    delete t.root;
}

// Timing sample of or_vptree_knn for the CPU baseline: the tree over all N
// points, then the searches of queries 0, stride, 2*stride, ... only.  Rows of
// idx/dist hold those queries in order.  Returns the number of queries;
// *sec_build / *sec_search receive the wall time of each phase.
OR_API int or_vptree_knn_sample(const double* X, int N, int D, int K, int stride, int* idx, double* dist,
                                double* sec_build, double* sec_search)
{
    VpTree t;
    t.X = X;
    t.D = D;
    t.items.resize(N);
    for (int n = 0; n < N; n++) t.items[n] = n;
    const auto t0 = std::chrono::steady_clock::now();
    t.root = t.build(0, N);
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<VpHeapItem> res;
    int q = 0;
    for (int n = 0; n < N; n += stride, q++) {
        std::priority_queue<VpHeapItem> heap;
        t.tau = DBL_MAX;
        t.search(t.root, n, K + 1, heap);
        res.clear();
        while (!heap.empty()) {
            res.push_back(heap.top());
            heap.pop();
        }
        std::reverse(res.begin(), res.end());
        for (int k = 0; k < K; k++) {
            idx[(size_t)q * K + k] = t.items[res[k + 1].index];
            dist[(size_t)q * K + k] = res[k + 1].dist;
        }
    }
    const auto t2 = std::chrono::steady_clock::now();
    *sec_build = std::chrono::duration<double>(t1 - t0).count();
    *sec_search = std::chrono::duration<double>(t2 - t1).count();
    delete t.root;
    return q;
}

// ---------------------------------------------------------------------------
// Perplexity calibration.  Bisection on beta exactly as tsne.hpp:648-711
// (kNN rows) and :531-594 (dense rows): beta starts at 1, doubles/halves
// until bracketed, then bisects; tol 1e-5 on H - log(perplexity); <= 200
// evaluations; sum_P seeded with DBL_MIN; the returned row is the one
// evaluated last (the update after a failed test is not re-evaluated).
// ---------------------------------------------------------------------------
struct BisectResult {
    double beta;
    int iters;
    int found;
};

template <class Eval>
static BisectResult bisect_beta(double perplexity, Eval eval)
{
    bool found = false;
    double beta = 1.0, min_beta = -DBL_MAX, max_beta = DBL_MAX;
    const double tol = 1e-5;
    int iter = 0;
    double last_beta = beta;
    while (!found && iter < 200) {
        last_beta = beta;
        double H = eval(beta);
        double Hdiff = H - std::log(perplexity);
        if (Hdiff < tol && -Hdiff < tol) {
            found = true;
        } else if (Hdiff > 0) {
            min_beta = beta;
            if (max_beta == DBL_MAX || max_beta == -DBL_MAX)
                beta *= 2.0;
            else
                beta = (beta + max_beta) / 2.0;
        } else {
            max_beta = beta;
            if (min_beta == -DBL_MAX || min_beta == DBL_MAX)
                beta /= 2.0;
            else
                beta = (beta + min_beta) / 2.0;
        }
        iter++;
    }
    return BisectResult{last_beta, iter, found ? 1 : 0};
}

// dist: N x K squared distances of the K neighbours (self already dropped).
// val:  N x K output, row normalised.  beta/iters optional (may be NULL).
OR_API int or_perplexity_sparse(const double* dist, int N, int K, double perplexity, double* val, double* beta_out,
                                int* iters_out)
{
    int unconverged = 0;
    std::vector<double> cur(K);
    for (int n = 0; n < N; n++) {
        const double* dr = dist + (size_t)n * K;
        double sum_P = DBL_MIN;
        auto eval = [&](double beta) {
            for (int m = 0; m < K; m++) cur[m] = std::exp(-beta * dr[m]);
            sum_P = DBL_MIN;
            for (int m = 0; m < K; m++) sum_P += cur[m];
            double H = 0.0;
            for (int m = 0; m < K; m++) H += beta * (dr[m] * cur[m]);
            return (H / sum_P) + std::log(sum_P);
        };
        BisectResult r = bisect_beta(perplexity, eval);
        for (int m = 0; m < K; m++) val[(size_t)n * K + m] = cur[m] / sum_P;
        if (beta_out) beta_out[n] = r.beta;
        if (iters_out) iters_out[n] = r.iters;
        unconverged += !r.found;
    }
    return unconverged;
}

// Dense rows over all N points with the diagonal forced to DBL_MIN
// (tsne.hpp:548-550).  DD: N x N from or_sqdist.  P: N x N output.
OR_API int or_perplexity_dense(const double* DD, int N, double perplexity, double* P, double* beta_out)
{
    int unconverged = 0;
    for (int n = 0; n < N; n++) {
        const double* dr = DD + (size_t)n * N;
        double* pr = P + (size_t)n * N;
        double sum_P = DBL_MIN;
        auto eval = [&](double beta) {
            for (int m = 0; m < N; m++) pr[m] = std::exp(-beta * dr[m]);
            pr[n] = DBL_MIN;
            sum_P = DBL_MIN;
            for (int m = 0; m < N; m++) sum_P += pr[m];
            double H = 0.0;
            for (int m = 0; m < N; m++) H += beta * (dr[m] * pr[m]);
            return (H / sum_P) + std::log(sum_P);
        };
        BisectResult r = bisect_beta(perplexity, eval);
        for (int m = 0; m < N; m++) pr[m] /= sum_P;
        if (beta_out) beta_out[n] = r.beta;
        unconverged += !r.found;
    }
    return unconverged;
}

// ---------------------------------------------------------------------------
// Symmetrisation: tsne.hpp:207-322.  Output has the reference's row layout
// (insertion order within rows) and values (p_ij + p_ji)/2 over the union
// pattern.  srow must hold N+1 ints; scol/sval must hold 2*nnz_in entries.
// Returns the number of symmetric entries.
// ---------------------------------------------------------------------------
OR_API long long or_symmetrize(const int* row_P, const int* col_P, const double* val_P, int N, int* srow, int* scol,
                               double* sval)
{
    std::vector<int> row_counts(N, 0);
    for (int n = 0; n < N; n++) {
        for (int i = row_P[n]; i < row_P[n + 1]; i++) {
            bool present = false;
            for (int m = row_P[col_P[i]]; m < row_P[col_P[i] + 1]; m++)
                if (col_P[m] == n) present = true;
            if (present)
                row_counts[n]++;
            else {
                row_counts[n]++;
                row_counts[col_P[i]]++;
            }
        }
    }
    srow[0] = 0;
    for (int n = 0; n < N; n++) srow[n + 1] = srow[n] + row_counts[n];
    std::vector<int> offset(N, 0);
    for (int n = 0; n < N; n++) {
        for (int i = row_P[n]; i < row_P[n + 1]; i++) {
            const int c = col_P[i];
            bool present = false;
            for (int m = row_P[c]; m < row_P[c + 1]; m++) {
                if (col_P[m] == n) {
                    present = true;
                    if (n <= c) {
                        scol[srow[n] + offset[n]] = c;
                        scol[srow[c] + offset[c]] = n;
                        sval[srow[n] + offset[n]] = val_P[i] + val_P[m];
                        sval[srow[c] + offset[c]] = val_P[i] + val_P[m];
                    }
                }
            }
            if (!present) {
                scol[srow[n] + offset[n]] = c;
                scol[srow[c] + offset[c]] = n;
                sval[srow[n] + offset[n]] = val_P[i];
                sval[srow[c] + offset[c]] = val_P[i];
            }
            if (!present || (present && n <= c)) {
                offset[n]++;
                if (c != n) offset[c]++;
            }
        }
    }
    const long long nnz = srow[N];
    for (long long i = 0; i < nnz; i++) sval[i] /= 2.0;
    return nnz;
}

// ---------------------------------------------------------------------------
// Pointer quadtree: quadtree.hpp:44-384.  Rectangle root centred on the mean
// with half-widths max(max-mean, mean-min)+1e-5 per axis; capacity-1 leaves;
// online centre of mass; duplicates counted but not stored; children tried
// in NW, NE, SW, SE order with closed-interval containment.
// ---------------------------------------------------------------------------
namespace {
struct QNode {
    double cx, cy, hw, hh;  // Cell (quadtree.hpp:44-65)
    bool is_leaf = true;
    int size = 0;
    int cum_size = 0;
    double com[2] = {0.0, 0.0};
    int index = -1;
    QNode* child[4] = {nullptr, nullptr, nullptr, nullptr};  // NW, NE, SW, SE
    int level = 0;

    bool contains(const double* p) const
    {
        if (cx - hw > p[0]) return false;
        if (cx + hw < p[0]) return false;
        if (cy - hh > p[1]) return false;
        if (cy + hh < p[1]) return false;
        return true;
    }
};

struct QTree {
    const double* Y;
    QNode* root = nullptr;

    static QNode* make(double x, double y, double hw, double hh, int level)
    {
        QNode* q = new QNode();
        q->cx = x;
        q->cy = y;
        q->hw = hw;
        q->hh = hh;
        q->level = level;
        return q;
    }

    void subdivide(QNode* q)  // quadtree.hpp:255-285
    {
        const double h = .5 * q->hw, v = .5 * q->hh;
        q->child[0] = make(q->cx - h, q->cy - v, h, v, q->level + 1);
        q->child[1] = make(q->cx + h, q->cy - v, h, v, q->level + 1);
        q->child[2] = make(q->cx - h, q->cy + v, h, v, q->level + 1);
        q->child[3] = make(q->cx + h, q->cy + v, h, v, q->level + 1);
        for (int i = 0; i < q->size; i++) {
            bool ok = false;
            for (int c = 0; c < 4 && !ok; c++) ok = insert(q->child[c], q->index);
            q->index = -1;
        }
        q->size = 0;
        q->is_leaf = false;
    }

    bool insert(QNode* q, int idx)  // quadtree.hpp:194-252
    {
        const double* p = Y + (size_t)idx * 2;
        if (!q->contains(p)) return false;
        q->cum_size++;
        const double mult1 = (double)(q->cum_size - 1) / (double)q->cum_size;
        const double mult2 = 1.0 / (double)q->cum_size;
        for (int d = 0; d < 2; d++) q->com[d] *= mult1;
        for (int d = 0; d < 2; d++) q->com[d] += mult2 * p[d];
        if (q->is_leaf && q->size < 1) {
            q->index = idx;
            q->size++;
            return true;
        }
        if (q->size == 1) {
            const double* o = Y + (size_t)q->index * 2;
            if (o[0] == p[0] && o[1] == p[1]) return true;  // duplicate
        }
        if (q->is_leaf) subdivide(q);
        for (int c = 0; c < 4; c++)
            if (insert(q->child[c], idx)) return true;
        return false;
    }

    void build(const double* Yin, int N)  // quadtree.hpp:99-134
    {
        Y = Yin;
        double mean[2] = {0, 0}, mn[2] = {DBL_MAX, DBL_MAX}, mx[2] = {-DBL_MAX, -DBL_MAX};
        for (int n = 0; n < N; n++)
            for (int d = 0; d < 2; d++) {
                const double v = Y[(size_t)n * 2 + d];
                mean[d] += v;
                if (v < mn[d]) mn[d] = v;
                if (v > mx[d]) mx[d] = v;
            }
        for (int d = 0; d < 2; d++) mean[d] /= (double)N;
        root = make(mean[0], mean[1], std::max(mx[0] - mean[0], mean[0] - mn[0]) + 1e-5,
                    std::max(mx[1] - mean[1], mean[1] - mn[1]) + 1e-5, 0);
        for (int i = 0; i < N; i++) insert(root, i);
    }

    static void destroy(QNode* q)
    {
        if (!q) return;
        for (int c = 0; c < 4; c++) destroy(q->child[c]);
        delete q;
    }

    // quadtree.hpp:317-354
    void non_edge(const QNode* q, int pi, double theta, double* neg_f, double* sum_Q) const
    {
        if (q->cum_size == 0 || (q->is_leaf && q->size == 1 && q->index == pi)) return;
        double buff[2];
        double D = 0.0;
        for (int d = 0; d < 2; d++) buff[d] = Y[(size_t)pi * 2 + d] - q->com[d];
        for (int d = 0; d < 2; d++) D += buff[d] * buff[d];
        if (q->is_leaf || std::max(q->hh, q->hw) / std::sqrt(D) < theta) {
            const double Q = 1.0 / (1.0 + D);
            *sum_Q += q->cum_size * Q;
            const double mult = q->cum_size * Q * Q;
            for (int d = 0; d < 2; d++) neg_f[d] += mult * buff[d];
        } else {
            for (int c = 0; c < 4; c++) non_edge(q->child[c], pi, theta, neg_f, sum_Q);
        }
    }
};
}  // namespace

// quadtree.hpp:357-384.  Y is N x 2.
static void edge_forces(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N, int dims,
                        double* pos_f)
{
    std::vector<double> buff(dims);
    for (int n = 0; n < N; n++) {
        for (int i = row_P[n]; i < row_P[n + 1]; i++) {
            double D = 0.0;
            for (int d = 0; d < dims; d++) buff[d] = Y[(size_t)n * dims + d] - Y[(size_t)col_P[i] * dims + d];
            for (int d = 0; d < dims; d++) D += buff[d] * buff[d];
            D = val_P[i] / (1.0 + D);
            for (int d = 0; d < dims; d++) pos_f[(size_t)n * dims + d] += D * buff[d];
        }
    }
}

// Barnes-Hut gradient, tsne.hpp:325-351 (no factor 4: dC = pos_f - neg_f/sum_Q).
OR_API double or_gradient_bh(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N,
                             double theta, double* dC)
{
    QTree t;
    t.build(Y, N);
    std::vector<double> pos(2 * (size_t)N, 0.0), neg(2 * (size_t)N, 0.0);
    double sum_Q = 0.0;
    edge_forces(row_P, col_P, val_P, Y, N, 2, pos.data());
    for (int n = 0; n < N; n++) t.non_edge(t.root, n, theta, neg.data() + 2 * (size_t)n, &sum_Q);
    for (size_t i = 0; i < 2 * (size_t)N; i++) dC[i] = pos[i] - (neg[i] / sum_Q);
    QTree::destroy(t.root);
    return sum_Q;
}

// ---------------------------------------------------------------------------
// Pointer octree: the quadtree above (quadtree.hpp:44-384) with a third axis.
// The reference's tree is 2-D only (QT_NO_DIMS = 2, quadtree.hpp:71); this
// is its direct generalisation, the definition the 3-D Barnes-Hut engine is
// checked against.  Root box centred on the mean with half-widths
// max(max-mean, mean-min)+1e-5 per axis; capacity-1 leaves; online centre of
// mass; duplicates counted but not stored; child c = 4 zbit + 2 ybit + xbit
// (bit 1 = upper half), tried in order 0..7 with closed-interval containment
// (a point on a split plane goes to the lower half); a cell is a summary when
// it is a leaf or max(half-widths)/sqrt(D) < theta.
// ---------------------------------------------------------------------------
namespace {
struct ONode {
    double c[3], h[3];
    bool is_leaf = true;
    int size = 0;
    int cum_size = 0;
    double com[3] = {0.0, 0.0, 0.0};
    int index = -1;
    ONode* child[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int level = 0;

    bool contains(const double* p) const
    {
        for (int d = 0; d < 3; d++) {
            if (c[d] - h[d] > p[d]) return false;
            if (c[d] + h[d] < p[d]) return false;
        }
        return true;
    }
};

struct OTree {
    const double* Y;
    ONode* root = nullptr;

    static ONode* make(const double* c, const double* h, int level)
    {
        ONode* q = new ONode();
        for (int d = 0; d < 3; d++) {
            q->c[d] = c[d];
            q->h[d] = h[d];
        }
        q->level = level;
        return q;
    }

    void subdivide(ONode* q)
    {
        double h[3];
        for (int d = 0; d < 3; d++) h[d] = .5 * q->h[d];
        for (int k = 0; k < 8; k++) {
            double c[3];
            for (int d = 0; d < 3; d++) c[d] = ((k >> d) & 1) ? q->c[d] + h[d] : q->c[d] - h[d];
            q->child[k] = make(c, h, q->level + 1);
        }
        for (int i = 0; i < q->size; i++) {
            bool ok = false;
            for (int k = 0; k < 8 && !ok; k++) ok = insert(q->child[k], q->index);
            q->index = -1;
        }
        q->size = 0;
        q->is_leaf = false;
    }

    bool insert(ONode* q, int idx)
    {
        const double* p = Y + (size_t)idx * 3;
        if (!q->contains(p)) return false;
        q->cum_size++;
        const double mult1 = (double)(q->cum_size - 1) / (double)q->cum_size;
        const double mult2 = 1.0 / (double)q->cum_size;
        for (int d = 0; d < 3; d++) q->com[d] *= mult1;
        for (int d = 0; d < 3; d++) q->com[d] += mult2 * p[d];
        if (q->is_leaf && q->size < 1) {
            q->index = idx;
            q->size++;
            return true;
        }
        if (q->size == 1) {
            const double* o = Y + (size_t)q->index * 3;
            if (o[0] == p[0] && o[1] == p[1] && o[2] == p[2]) return true;  // duplicate
        }
        if (q->is_leaf) subdivide(q);
        for (int k = 0; k < 8; k++)
            if (insert(q->child[k], idx)) return true;
        return false;
    }

    void build(const double* Yin, int N)
    {
        Y = Yin;
        double mean[3] = {0, 0, 0}, mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int n = 0; n < N; n++)
            for (int d = 0; d < 3; d++) {
                const double v = Y[(size_t)n * 3 + d];
                mean[d] += v;
                if (v < mn[d]) mn[d] = v;
                if (v > mx[d]) mx[d] = v;
            }
        double h[3];
        for (int d = 0; d < 3; d++) {
            mean[d] /= (double)N;
            h[d] = std::max(mx[d] - mean[d], mean[d] - mn[d]) + 1e-5;
        }
        root = make(mean, h, 0);
        for (int i = 0; i < N; i++) insert(root, i);
    }

    static void destroy(ONode* q)
    {
        if (!q) return;
        for (int k = 0; k < 8; k++) destroy(q->child[k]);
        delete q;
    }

    void non_edge(const ONode* q, int pi, double theta, double* neg_f, double* sum_Q) const
    {
        if (q->cum_size == 0 || (q->is_leaf && q->size == 1 && q->index == pi)) return;
        double buff[3];
        double D = 0.0;
        for (int d = 0; d < 3; d++) buff[d] = Y[(size_t)pi * 3 + d] - q->com[d];
        for (int d = 0; d < 3; d++) D += buff[d] * buff[d];
        if (q->is_leaf || std::max(std::max(q->h[0], q->h[1]), q->h[2]) / std::sqrt(D) < theta) {
            const double Q = 1.0 / (1.0 + D);
            *sum_Q += q->cum_size * Q;
            const double mult = q->cum_size * Q * Q;
            for (int d = 0; d < 3; d++) neg_f[d] += mult * buff[d];
        } else {
            for (int k = 0; k < 8; k++) non_edge(q->child[k], pi, theta, neg_f, sum_Q);
        }
    }
};
}  // namespace

// Barnes-Hut gradient on the octree (3-D counterpart of or_gradient_bh; no
// factor 4: dC = pos_f - neg_f/sum_Q).  Y is N x 3.
OR_API double or_gradient_bh3(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N,
                              double theta, double* dC)
{
    OTree t;
    t.build(Y, N);
    std::vector<double> pos(3 * (size_t)N, 0.0), neg(3 * (size_t)N, 0.0);
    double sum_Q = 0.0;
    edge_forces(row_P, col_P, val_P, Y, N, 3, pos.data());
    for (int n = 0; n < N; n++) t.non_edge(t.root, n, theta, neg.data() + 3 * (size_t)n, &sum_Q);
    for (size_t i = 0; i < 3 * (size_t)N; i++) dC[i] = pos[i] - (neg[i] / sum_Q);
    OTree::destroy(t.root);
    return sum_Q;
}

// Octree node count, depth and the number of non-empty cells per level (up
// to cap levels), for structural checks.
OR_API int or_octree_stats(const double* Y, int N, int cap, int* per_level)
{
    OTree t;
    t.build(Y, N);
    for (int l = 0; l < cap; l++) per_level[l] = 0;
    int nodes = 0;
    std::vector<const ONode*> st{t.root};
    while (!st.empty()) {
        const ONode* q = st.back();
        st.pop_back();
        if (q->cum_size == 0) continue;
        nodes++;
        if (q->level < cap) per_level[q->level]++;
        for (int k = 0; k < 8; k++)
            if (q->child[k]) st.push_back(q->child[k]);
    }
    OTree::destroy(t.root);
    return nodes;
}

// Flattened export of the quadtree for structural comparison.  For every node
// (pre-order, children NW,NE,SW,SE): level, cum_size, is_leaf, stored index,
// com x/y, number of non-empty children.  Returns the node count; writes at
// most cap nodes.
namespace {
struct Exporter {
    int cap, n = 0;
    int *level, *cum, *leaf, *index, *nchild;
    double *comx, *comy;
    void walk(const QNode* q)
    {
        if (!q) return;
        if (n < cap) {
            level[n] = q->level;
            cum[n] = q->cum_size;
            leaf[n] = q->is_leaf ? 1 : 0;
            index[n] = q->index;
            comx[n] = q->com[0];
            comy[n] = q->com[1];
            int c = 0;
            for (int k = 0; k < 4; k++) c += (q->child[k] && q->child[k]->cum_size > 0);
            nchild[n] = c;
        }
        n++;
        for (int k = 0; k < 4; k++) walk(q->child[k]);
    }
};
}  // namespace

OR_API int or_quadtree_export(const double* Y, int N, int cap, int* level, int* cum, int* leaf, int* index,
                              double* comx, double* comy, int* nchild, double* root_box)
{
    QTree t;
    t.build(Y, N);
    Exporter e{cap, 0, level, cum, leaf, index, nchild, comx, comy};
    e.walk(t.root);
    if (root_box) {
        root_box[0] = t.root->cx;
        root_box[1] = t.root->cy;
        root_box[2] = t.root->hw;
        root_box[3] = t.root->hh;
    }
    QTree::destroy(t.root);
    return e.n;
}

// Exact gradient with a dense P, tsne.hpp:353-407.  compat selects the
// reference's DD (see or_sqdist).  dims = no_dims.
OR_API double or_gradient_exact_dense(const double* P, const double* Y, int N, int dims, double* dC, int compat)
{
    std::vector<double> DD((size_t)N * N), Q((size_t)N * N);
    or_sqdist(Y, N, dims, DD.data(), compat);
    double sum_Q = 0.0;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < N; m++)
            if (n != m) {
                Q[(size_t)n * N + m] = 1 / (1 + DD[(size_t)n * N + m]);
                sum_Q += Q[(size_t)n * N + m];
            }
    for (size_t i = 0; i < (size_t)N * dims; i++) dC[i] = 0.0;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < N; m++)
            if (n != m) {
                const double mult = (P[(size_t)n * N + m] - (Q[(size_t)n * N + m] / sum_Q)) * Q[(size_t)n * N + m];
                for (int d = 0; d < dims; d++)
                    dC[(size_t)n * dims + d] += (Y[(size_t)n * dims + d] - Y[(size_t)m * dims + d]) * mult;
            }
    return sum_Q;
}

// Exact repulsion with a sparse (CSR) P: the same objective as the dense form
// with P zero off the pattern; textbook distances.  Used for "theta = 0 with
// a kNN P" and as the theta -> 0 limit of the Barnes-Hut gradient.
OR_API double or_gradient_exact_sparse(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N,
                                       int dims, double* dC)
{
    std::vector<double> pos((size_t)N * dims, 0.0), neg((size_t)N * dims, 0.0);
    edge_forces(row_P, col_P, val_P, Y, N, dims, pos.data());
    double sum_Q = 0.0;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < N; m++) {
            if (n == m) continue;
            double D = 0.0;
            for (int d = 0; d < dims; d++) {
                double t = Y[(size_t)n * dims + d] - Y[(size_t)m * dims + d];
                D += t * t;
            }
            const double Q = 1.0 / (1.0 + D);
            sum_Q += Q;
            for (int d = 0; d < dims; d++)
                neg[(size_t)n * dims + d] += Q * Q * (Y[(size_t)n * dims + d] - Y[(size_t)m * dims + d]);
        }
    for (size_t i = 0; i < (size_t)N * dims; i++) dC[i] = pos[i] - neg[i] / sum_Q;
    return sum_Q;
}

// ---------------------------------------------------------------------------
// KL divergence.
// ---------------------------------------------------------------------------
// tsne.hpp:409-452 (dense P; the reference hard-codes 2 dims at :418).
OR_API double or_kl_exact_dense(const double* P, const double* Y, int N, int dims, int compat)
{
    std::vector<double> DD((size_t)N * N), Q((size_t)N * N);
    or_sqdist(Y, N, dims, DD.data(), compat);
    double sum_Q = DBL_MIN;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < N; m++) {
            if (n != m) {
                Q[(size_t)n * N + m] = 1 / (1 + DD[(size_t)n * N + m]);
                sum_Q += Q[(size_t)n * N + m];
            } else
                Q[(size_t)n * N + m] = DBL_MIN;
        }
    for (size_t i = 0; i < (size_t)N * N; i++) Q[i] /= sum_Q;
    double C = 0.0;
    for (size_t i = 0; i < (size_t)N * N; i++) C += P[i] * std::log((P[i] + 1e-9) / (Q[i] + 1e-9));
    return C;
}

// tsne.hpp:454-486 (sparse P, Barnes-Hut estimate of sum_Q, FLT_MIN guards).
OR_API double or_kl_bh(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N, double theta)
{
    QTree t;
    t.build(Y, N);
    double buff[2] = {0, 0};
    double sum_Q = 0.0;
    for (int n = 0; n < N; n++) t.non_edge(t.root, n, theta, buff, &sum_Q);
    QTree::destroy(t.root);
    double C = 0.0;
    for (int n = 0; n < N; n++)
        for (int i = row_P[n]; i < row_P[n + 1]; i++) {
            double Q = 0.0;
            for (int d = 0; d < 2; d++) {
                double b = Y[(size_t)n * 2 + d] - Y[(size_t)col_P[i] * 2 + d];
                Q += b * b;
            }
            Q = (1.0 / (1.0 + Q)) / sum_Q;
            C += val_P[i] * std::log((val_P[i] + FLT_MIN) / (Q + FLT_MIN));
        }
    return C;
}

// or_kl_bh with the octree's estimate of Z (Y is N x 3).
OR_API double or_kl_bh3(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N, double theta)
{
    OTree t;
    t.build(Y, N);
    double buff[3] = {0, 0, 0};
    double sum_Q = 0.0;
    for (int n = 0; n < N; n++) t.non_edge(t.root, n, theta, buff, &sum_Q);
    OTree::destroy(t.root);
    double C = 0.0;
    for (int n = 0; n < N; n++)
        for (int i = row_P[n]; i < row_P[n + 1]; i++) {
            double Q = 0.0;
            for (int d = 0; d < 3; d++) {
                double b = Y[(size_t)n * 3 + d] - Y[(size_t)col_P[i] * 3 + d];
                Q += b * b;
            }
            Q = (1.0 / (1.0 + Q)) / sum_Q;
            C += val_P[i] * std::log((val_P[i] + FLT_MIN) / (Q + FLT_MIN));
        }
    return C;
}

// KL(P||Q) over the sparse pattern with the exact normaliser Z = sum_{i!=j}
// 1/(1+d_ij^2): the objective whose gradient is 4*dC.  Terms with p = 0 are 0.
OR_API double or_kl_exact_sparse(const int* row_P, const int* col_P, const double* val_P, const double* Y, int N,
                                 int dims)
{
    double Z = 0.0;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < N; m++) {
            if (n == m) continue;
            double D = 0.0;
            for (int d = 0; d < dims; d++) {
                double t = Y[(size_t)n * dims + d] - Y[(size_t)m * dims + d];
                D += t * t;
            }
            Z += 1.0 / (1.0 + D);
        }
    double C = 0.0;
    for (int n = 0; n < N; n++)
        for (int i = row_P[n]; i < row_P[n + 1]; i++) {
            const double p = val_P[i];
            if (p <= 0) continue;
            double D = 0.0;
            for (int d = 0; d < dims; d++) {
                double t = Y[(size_t)n * dims + d] - Y[(size_t)col_P[i] * dims + d];
                D += t * t;
            }
            C += p * std::log(p * Z * (1.0 + D));
        }
    return C;
}

// ---------------------------------------------------------------------------
// Optimiser step: tsne.hpp:152-167 (gains, momentum, update) + zeroMean.
// sign(0) == 0 (tsne.hpp:53-56).
// ---------------------------------------------------------------------------
static inline double sgn(double x) { return x == .0 ? .0 : (x < .0 ? -1.0 : 1.0); }

OR_API void or_update(const double* dY, double* uY, double* gains, double* Y, int N, int dims, double momentum,
                      double eta)
{
    const size_t n = (size_t)N * dims;
    for (size_t i = 0; i < n; i++) gains[i] = (sgn(dY[i]) != sgn(uY[i])) ? (gains[i] + .2) : (gains[i] * .8);
    for (size_t i = 0; i < n; i++)
        if (gains[i] < .01) gains[i] = .01;
    for (size_t i = 0; i < n; i++) uY[i] = momentum * uY[i] - eta * gains[i] * dY[i];
    for (size_t i = 0; i < n; i++) Y[i] = Y[i] + uY[i];
    or_zero_mean(Y, N, dims);
}

// ---------------------------------------------------------------------------
// Full run: tsne.hpp:61-205.
// ---------------------------------------------------------------------------
typedef struct or_run_options {
    int max_iter;         // 1000
    int stop_lying_iter;  // 250
    int mom_switch_iter;  // 250
    int knn_mode;         // 0 = VP-tree as the reference, 1 = brute force
    int exact_p_mode;     // theta == 0: 0 = dense P (reference), 1 = kNN P + exact repulsion
    int dd_compat;        // theta == 0 dense path: 1 = reference's DD overwrite
    int y0_mode;          // 0 = std::rand as the reference (after the VP-tree's draws), 1 = explicit y0
    const double* y0;     // N x no_dims when y0_mode == 1
    int log_every;        // 50 (0 = only the reference's schedule disabled)
    // outputs
    double* kl_trace;     // optional, capacity kl_cap
    int* kl_iter;
    int kl_cap;
    int kl_count;
    double ms_p;          // similarities phase
    double ms_loop;       // main loop
    long long nnz;
} or_run_options;

#include <chrono>

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

OR_API int or_tsne_run(double* X, int N, int D, double* Y, int no_dims, double perplexity, double theta,
                       or_run_options* opt)
{
    const bool exact = (theta == .0);
    const int max_iter = opt->max_iter, stop_lying_iter = opt->stop_lying_iter, mom_switch_iter = opt->mom_switch_iter;
    double momentum = .5;
    const double final_momentum = .8, eta = 200.0;
    const size_t nd = (size_t)N * no_dims;
    std::vector<double> dY(nd), uY(nd, 0.0), gains(nd, 1.0);
    std::vector<double> P;
    std::vector<int> row_P, col_P;
    std::vector<double> val_P;
    const bool sparse = !exact || opt->exact_p_mode == 1;
    opt->kl_count = 0;

    double t0 = now_ms();
    or_normalize(X, N, D);
    if (!sparse) {
        std::vector<double> DD((size_t)N * N);
        or_sqdist(X, N, D, DD.data(), opt->dd_compat);
        P.resize((size_t)N * N);
        or_perplexity_dense(DD.data(), N, perplexity, P.data(), nullptr);
        for (int n = 0; n < N; n++)
            for (int m = n + 1; m < N; m++) {
                P[(size_t)n * N + m] += P[(size_t)m * N + n];
                P[(size_t)m * N + n] = P[(size_t)n * N + m];
            }
        double s = 0.0;
        for (size_t i = 0; i < P.size(); i++) s += P[i];
        for (size_t i = 0; i < P.size(); i++) P[i] /= s;
        for (size_t i = 0; i < P.size(); i++) P[i] *= 12.0;
        opt->nnz = (long long)N * N;
    } else {
        const int K = (int)(3 * perplexity);
        std::vector<int> idx((size_t)N * K);
        std::vector<double> dist((size_t)N * K), val((size_t)N * K);
        if (opt->knn_mode == 0)
            or_vptree_knn(X, N, D, K, idx.data(), dist.data());
        else
            or_knn_brute(X, N, D, K, idx.data(), dist.data());
        or_perplexity_sparse(dist.data(), N, K, perplexity, val.data(), nullptr, nullptr);
        std::vector<int> r(N + 1);
        for (int n = 0; n <= N; n++) r[n] = n * K;
        row_P.resize(N + 1);
        col_P.resize(2 * (size_t)N * K);
        val_P.resize(2 * (size_t)N * K);
        long long nnz = or_symmetrize(r.data(), idx.data(), val.data(), N, row_P.data(), col_P.data(), val_P.data());
        col_P.resize(nnz);
        val_P.resize(nnz);
        double s = 0.0;
        for (long long i = 0; i < nnz; i++) s += val_P[i];
        for (long long i = 0; i < nnz; i++) val_P[i] /= s;
        for (long long i = 0; i < nnz; i++) val_P[i] *= 12.0;
        opt->nnz = nnz;
    }
    if (opt->y0_mode == 1) {
        for (size_t i = 0; i < nd; i++) Y[i] = opt->y0[i];
    } else {
        for (size_t i = 0; i < nd; i++) Y[i] = or_gaussian_random() * .0001;
    }
    double t1 = now_ms();
    opt->ms_p = t1 - t0;

    for (int iter = 0; iter < max_iter; iter++) {
        if (!sparse)
            or_gradient_exact_dense(P.data(), Y, N, no_dims, dY.data(), opt->dd_compat);
        else if (exact)
            or_gradient_exact_sparse(row_P.data(), col_P.data(), val_P.data(), Y, N, no_dims, dY.data());
        else if (no_dims == 3)
            or_gradient_bh3(row_P.data(), col_P.data(), val_P.data(), Y, N, theta, dY.data());
        else
            or_gradient_bh(row_P.data(), col_P.data(), val_P.data(), Y, N, theta, dY.data());
        or_update(dY.data(), uY.data(), gains.data(), Y, N, no_dims, momentum, eta);
        if (iter == stop_lying_iter) {
            if (!sparse)
                for (size_t i = 0; i < P.size(); i++) P[i] /= 12.0;
            else
                for (size_t i = 0; i < val_P.size(); i++) val_P[i] /= 12.0;
        }
        if (iter == mom_switch_iter) momentum = final_momentum;
        if (opt->log_every > 0 && iter > 0 && ((iter % opt->log_every == 0) || (iter == max_iter - 1))) {
            double C;
            if (!sparse)
                C = or_kl_exact_dense(P.data(), Y, N, 2, opt->dd_compat);
            else if (exact)
                C = or_kl_exact_sparse(row_P.data(), col_P.data(), val_P.data(), Y, N, no_dims);
            else if (no_dims == 3)
                C = or_kl_bh3(row_P.data(), col_P.data(), val_P.data(), Y, N, theta);
            else
                C = or_kl_bh(row_P.data(), col_P.data(), val_P.data(), Y, N, theta);
            if (opt->kl_trace && opt->kl_count < opt->kl_cap) {
                opt->kl_trace[opt->kl_count] = C;
                opt->kl_iter[opt->kl_count] = iter;
            }
            opt->kl_count++;
        }
    }
    opt->ms_loop = now_ms() - t1;
    return 0;
}

// Sparse P pipeline only (normalise -> kNN -> perplexity -> symmetrise ->
// normalise), without the x12 exaggeration, for P-parity tests.  X is
// normalised in place.  Outputs must hold N+1 / 2*N*K entries.
OR_API long long or_build_p(double* X, int N, int D, double perplexity, int K, int knn_mode, int* row_P, int* col_P,
                            double* val_P)
{
    or_normalize(X, N, D);
    std::vector<int> idx((size_t)N * K);
    std::vector<double> dist((size_t)N * K), val((size_t)N * K);
    if (knn_mode == 0)
        or_vptree_knn(X, N, D, K, idx.data(), dist.data());
    else
        or_knn_brute(X, N, D, K, idx.data(), dist.data());
    or_perplexity_sparse(dist.data(), N, K, perplexity, val.data(), nullptr, nullptr);
    std::vector<int> r(N + 1);
    for (int n = 0; n <= N; n++) r[n] = n * K;
    long long nnz = or_symmetrize(r.data(), idx.data(), val.data(), N, row_P, col_P, val_P);
    double s = 0.0;
    for (long long i = 0; i < nnz; i++) s += val_P[i];
    for (long long i = 0; i < nnz; i++) val_P[i] /= s;
    return nnz;
}

OR_API int or_run_options_size(void) { return (int)sizeof(or_run_options); }
