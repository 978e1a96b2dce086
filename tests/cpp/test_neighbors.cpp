// The GPU neighbour backend (include/tapkee_amd/neighbors.hpp) against the
// reference's own neighbour tests (test/unit/neighbors.cpp:17-164): 100
// objects at 0, 1, ..., 99 on a line, k = 10; every object's neighbours are
// k distinct other objects and include its k/2 nearest on each side (fewer
// at the ends).  Those tests use a distance callback and a linear-kernel
// callback, sqrt(k(a,a) - 2k(a,b) + k(b,b)) = |a - b|: both are the
// Euclidean distance of the one-feature vectors here.  Then the connectivity
// retry (neighbors.hpp:196-205): two groups far apart are not connected at
// k = 10, 20 or 40; the search is repeated with 2k until k = 80.
// Without --gpu only the connectivity test of fixed graphs runs (no device).
#include <tapkee_amd/neighbors.hpp>

#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <vector>

using namespace tapkee;
using namespace tapkee::tapkee_internal;

static int failures = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "FAIL line %d: %s\n", __LINE__, #c);   \
            failures++;                                            \
        }                                                          \
    } while (0)

struct Lines : LoggerImplementation {
    std::vector<std::string> v;
    void message_info(const std::string& m) override { v.push_back("info:" + m); }
    void message_warning(const std::string& m) override { v.push_back("warning:" + m); }
    void message_debug(const std::string& m) override { v.push_back("debug:" + m); }
    void message_error(const std::string& m) override { v.push_back("error:" + m); }
    void message_benchmark(const std::string& m) override { v.push_back("benchmark:" + m); }
};

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    // fixed graphs: connected.hpp follows the directed lists from object 0
    CHECK(is_connected(Neighbors{{1}, {2}, {0}}));
    CHECK(!is_connected(Neighbors{{1}, {0}, {0}}));
    CHECK(is_connected(Neighbors{{1, 2}, {0, 2}, {0, 1}}));
    if (gpu) {
        const int N = 100, k = 10;
        DenseMatrix line(1, N);
        for (int i = 0; i < N; i++) line(0, i) = (double)i;
        for (const NeighborsMethod& m : {Brute, VpTree, CoverTree}) {
            Neighbors nb = find_neighbors(m, line, k, true);
            CHECK((int)nb.size() == N);
            for (int i = 0; i < N; i++) {
                CHECK((int)nb[i].size() == k);
                std::set<IndexType> s(nb[i].begin(), nb[i].end());
                CHECK((int)s.size() == k);   // no repeated neighbours
                CHECK(!s.count(i));          // not its own neighbour
                const int left = std::min(i, k / 2), right = std::min(N - i - 1, k / 2);
                for (int j = 0; j < left; j++) CHECK(s.count(i - j - 1));
                for (int j = 0; j < right; j++) CHECK(s.count(i + j + 1));
            }
        }
        Lines* log = new Lines;
        Logging::instance().set_logger_impl(log);
        Logging::instance().enable_info();
        DenseMatrix two(2, N);
        for (int i = 0; i < N; i++) {
            two(0, i) = (i < 50 ? 0.0 : 1000.0) + (double)(i % 50);
            two(1, i) = 0.0;
        }
        Neighbors nb = find_neighbors(Brute, two, 10, true);
        CHECK((int)nb[0].size() == 80);
        CHECK(is_connected(nb));
        int retries = 0;
        for (const std::string& l : log->v)
            if (l.find("is not connected. Recomputing with a larger number of neighbors") != std::string::npos) retries++;
        CHECK(retries == 3);  // 10 -> 20 -> 40 -> 80
        bool said = false;
        for (const std::string& l : log->v) said = said || l == "info:The neighborhood graph with 80 neighbors is connected.";
        CHECK(said);
        // k above N-1 is lowered, with the reference's warning
        log->v.clear();
        Neighbors all = find_neighbors(Brute, line, 500, false);
        CHECK((int)all[3].size() == N - 1);
        CHECK(!log->v.empty() && log->v[0].rfind("warning:Number of neighbors is greater", 0) == 0);
        Logging::instance().disable_info();
    }
    if (failures) {
        fprintf(stderr, "%d failure(s)\n", failures);
        return 1;
    }
    printf("all neighbour checks passed%s\n", gpu ? " (with GPU)" : "");
    return 0;
}
