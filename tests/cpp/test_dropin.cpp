// Drop-in boundary checks for include/tapkee_amd/tapkee.hpp, written like
// tapkee's own interface tests (test/unit/interface.cpp): each case expects
// a specific exception type.  Without arguments only the cases that fail
// before any device work run (CPU); with --gpu a real embedding and the
// out-of-memory mapping run too.
#include <tapkee_amd/tapkee.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

using namespace tapkee;

static int failures = 0;

#define EXPECT_THROW_T(stmt, T)                                                        \
    do {                                                                               \
        bool ok_ = false;                                                              \
        try {                                                                          \
            stmt;                                                                      \
        } catch (const T&) {                                                           \
            ok_ = true;                                                                \
        } catch (const std::exception& e) {                                            \
            fprintf(stderr, "line %d: wrong exception: %s\n", __LINE__, e.what());     \
        }                                                                              \
        if (!ok_) {                                                                    \
            fprintf(stderr, "FAIL line %d: expected %s from %s\n", __LINE__, #T, #stmt); \
            failures++;                                                                \
        }                                                                              \
    } while (0)

static DenseMatrix swissroll(int N)
{
    // features x samples, as test/unit/data.hpp builds it
    std::mt19937 g(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    DenseMatrix X(3, N);
    for (int i = 0; i < N; i++) {
        const double t = (3.0 * M_PI / 4.0) * (u(g) + 0.5), h = u(g) - 0.5;
        X(0, i) = t * std::cos(t);
        X(1, i) = 10 * h;
        X(2, i) = t * std::sin(t);
    }
    return X;
}

static bool cancel_now() { return true; }

// Collects the engine's messages (utils/logging.hpp:45-59 interface).
struct CollectingLogger : LoggerImplementation {
    std::vector<std::string> lines;
    void message_info(const std::string& m) override { lines.push_back("info:" + m); }
    void message_warning(const std::string& m) override { lines.push_back("warning:" + m); }
    void message_debug(const std::string& m) override { lines.push_back("debug:" + m); }
    void message_error(const std::string& m) override { lines.push_back("error:" + m); }
    void message_benchmark(const std::string& m) override { lines.push_back("benchmark:" + m); }
    bool has(const std::string& prefix) const
    {
        for (const auto& l : lines)
            if (l.compare(0, prefix.size(), prefix) == 0) return true;
        return false;
    }
};

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    DenseMatrix X = swissroll(100);

    // interface.cpp:198-210 missed_parameter_error when no method is given
    EXPECT_THROW_T(with((sne_perplexity = 10.0)).embedUsing(X), missed_parameter_error);
    // interface.cpp:160-174 multiple_parameter_error on a duplicate keyword
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                         sne_perplexity = 5.0)).embedUsing(X),
                   multiple_parameter_error);
    // interface.cpp:102-119 no_data_error on empty input
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding)).embedUsing(DenseMatrix(3, 0)),
                   no_data_error);
    // methods/tsne.hpp:20 perplexity in [0, (N-1)/3]
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 40.0)).embedUsing(X),
                   wrong_parameter_error);
    // methods/tsne.hpp:21 theta >= 0
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                         sne_theta = -0.1)).embedUsing(X),
                   wrong_parameter_error);
    // methods/base.hpp:43 target dimension in [1, N)
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                         target_dimension = 0)).embedUsing(X),
                   wrong_parameter_error);
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                         target_dimension = 100)).embedUsing(X),
                   wrong_parameter_error);
    // documented deviation: Barnes-Hut runs on a quadtree or an octree, so
    // it needs 2 or 3 output dimensions
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                         target_dimension = 4)).embedUsing(X),
                   wrong_parameter_error);
    // interface.cpp:181-196 cancelled_exception
    EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                         cancel_function = &cancel_now)).embedUsing(X),
                   cancelled_exception);
    // other methods are outside this engine
    EXPECT_THROW_T(with((method = Isomap, num_neighbors = 10)).embedUsing(X), unsupported_method_error);
    // non-finite input
    {
        DenseMatrix B = X;
        B(1, 7) = NAN;
        EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0)).embedUsing(B),
                       wrong_parameter_error);
    }

    // the reference's default levels (logging.hpp:96-98): warning and error on
    if (!Logging::instance().is_warning_enabled() || !Logging::instance().is_error_enabled() ||
        Logging::instance().is_info_enabled() || Logging::instance().is_benchmark_enabled()) {
        fprintf(stderr, "FAIL: default logging levels\n");
        failures++;
    }
    CollectingLogger* logger = new CollectingLogger;
    Logging::instance().set_logger_impl(logger);  // owned by Logging from here
    if (gpu) {
        Logging::instance().enable_info();
        Logging::instance().enable_benchmark();
        // methods.cpp:16-32 smoke test: shape N x 2, here also finite and centred
        TapkeeOutput out = with((method = tDistributedStochasticNeighborEmbedding, target_dimension = 2,
                                 num_neighbors = 20, sne_perplexity = 10.0))
                               .embedUsing(X);
        if (out.embedding.rows() != 100 || out.embedding.cols() != 2) {
            fprintf(stderr, "FAIL: shape %d x %d\n", out.embedding.rows(), out.embedding.cols());
            failures++;
        }
        double mx = 0.0, my = 0.0;
        for (int i = 0; i < 100; i++) {
            if (!std::isfinite(out.embedding(i, 0)) || !std::isfinite(out.embedding(i, 1))) failures++;
            mx += out.embedding(i, 0);
            my += out.embedding(i, 1);
        }
        if (std::fabs(mx) > 1e-6 || std::fabs(my) > 1e-6) {
            fprintf(stderr, "FAIL: embedding not centred (%g, %g)\n", mx, my);
            failures++;
        }
        printf("kl trace %d entries, last %.6f, final exact %.6f\n", out.stats.kl_count,
               out.stats.kl_value[out.stats.kl_count - 1], out.stats.kl_final_exact);
        // tsne.hpp:68, :87, :142, :191 through the installed logger
        for (const char* want : {"info:Using Barnes-Hut-SNE algorithm", "benchmark:Input similarities computation took ",
                                 "benchmark:Main t-SNE loop took ", "info:Iteration 50: error is "}) {
            if (!logger->has(want)) {
                fprintf(stderr, "FAIL: no log line '%s'\n", want);
                failures++;
            }
        }
        Logging::instance().disable_info();
        Logging::instance().disable_benchmark();
        logger->lines.clear();
        with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0)).embedUsing(X);
        if (logger->has("info:") || logger->has("benchmark:")) {
            fprintf(stderr, "FAIL: disabled levels still delivered\n");
            failures++;
        }
        // interface.cpp:245-261 not_enough_memory_error (forced by a cap)
        EngineOptions eo;
        eo.opt.device_mem_cap = 1024;
        EXPECT_THROW_T(with((method = tDistributedStochasticNeighborEmbedding, sne_perplexity = 10.0,
                             engine_options = eo)).embedUsing(X),
                       not_enough_memory_error);
    }
    if (failures) {
        fprintf(stderr, "%d failure(s)\n", failures);
        return 1;
    }
    printf("all drop-in checks passed%s\n", gpu ? " (with GPU)" : "");
    return 0;
}
