// Drop-in C++ front end of the MI355X t-SNE engine.
//
//   #include <tapkee_amd/tapkee.hpp>
//   using namespace tapkee;
//   TapkeeOutput out = with((method = tDistributedStochasticNeighborEmbedding,
//                            sne_perplexity = 30.0, sne_theta = 0.5)).embedUsing(M);
//
// mirrors tapkee's chain (include/tapkee/chain_interface.hpp:471-500), keyword
// objects (include/tapkee/defines/keywords.hpp, include/stichwort/), defaults
// merge (include/tapkee/embed.hpp:103-104), validation
// (include/tapkee/methods/tsne.hpp:18-22, methods/base.hpp:38-43) and
// exception types (include/tapkee/exceptions.hpp).  M is features x samples,
// column-major (tapkee::DenseMatrix); the result's embedding is
// samples x target_dimension.  Header-only C++17; links against
// libtapkee_amd.so through include/tapkee_amd.h.  Eigen is not required: with
// TAPKEE_AMD_USE_EIGEN defined, DenseMatrix is Eigen's dynamic double matrix.
#pragma once

#include <any>
#include <cstdint>
#include <functional>
#include <iostream>
#include <list>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../tapkee_amd.h"

#ifdef TAPKEE_AMD_USE_EIGEN
#include <Eigen/Dense>
#endif

namespace tapkee {

typedef double ScalarType;
typedef int IndexType;

#ifdef TAPKEE_AMD_USE_EIGEN
typedef Eigen::Matrix<ScalarType, Eigen::Dynamic, Eigen::Dynamic> DenseMatrix;
#else
// Column-major dense matrix with the members the t-SNE path uses.
class DenseMatrix {
  public:
    DenseMatrix() = default;
    DenseMatrix(IndexType r, IndexType c) : r_(r), c_(c), v_((size_t)r * c, 0.0) {}
    IndexType rows() const { return r_; }
    IndexType cols() const { return c_; }
    ScalarType* data() { return v_.data(); }
    const ScalarType* data() const { return v_.data(); }
    ScalarType& operator()(IndexType i, IndexType j) { return v_[(size_t)j * r_ + i]; }
    ScalarType operator()(IndexType i, IndexType j) const { return v_[(size_t)j * r_ + i]; }
    DenseMatrix transpose() const
    {
        DenseMatrix t(c_, r_);
        for (IndexType j = 0; j < c_; j++)
            for (IndexType i = 0; i < r_; i++) t(j, i) = (*this)(i, j);
        return t;
    }

  private:
    IndexType r_ = 0, c_ = 0;
    std::vector<ScalarType> v_;
};
#endif

// ---------------------------------------------------------------- exceptions
class no_data_error : public std::logic_error {
  public:
    explicit no_data_error() : std::logic_error("No data to embed") {}
};
class unsupported_method_error : public std::logic_error {
  public:
    explicit unsupported_method_error(const std::string& m) : std::logic_error(m) {}
};
class not_enough_memory_error : public std::runtime_error {
  public:
    explicit not_enough_memory_error(const std::string& m) : std::runtime_error(m) {}
};
class cancelled_exception : public std::exception {
  public:
    explicit cancelled_exception() : std::exception() {}
};
class eigendecomposition_error : public std::runtime_error {
  public:
    explicit eigendecomposition_error(const std::string& m) : std::runtime_error(m) {}
};
class missed_parameter_error : public std::logic_error {
  public:
    explicit missed_parameter_error(const std::string& m) : std::logic_error(m) {}
};
class wrong_parameter_error : public std::logic_error {
  public:
    explicit wrong_parameter_error(const std::string& m) : std::logic_error(m) {}
};
class wrong_parameter_type_error : public std::logic_error {
  public:
    explicit wrong_parameter_type_error(const std::string& m) : std::logic_error(m) {}
};
class multiple_parameter_error : public std::runtime_error {
  public:
    explicit multiple_parameter_error(const std::string& m) : std::runtime_error(m) {}
};
// A HIP runtime failure (no counterpart in the reference).
class device_error : public std::runtime_error {
  public:
    explicit device_error(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------- methods
struct DimensionReductionMethod {
    DimensionReductionMethod(const char* n, bool engine) : name_(n), runs_on_engine(engine) {}
    bool operator==(const DimensionReductionMethod& o) const { return name_ == o.name_; }
    std::string name() const { return name_; }
    std::string name_;
    bool runs_on_engine;
};

// Method objects of include/tapkee/defines/methods.hpp; only t-SNE runs here,
// the others exist so that code naming them compiles and fails at run time
// with unsupported_method_error.
static const DimensionReductionMethod tDistributedStochasticNeighborEmbedding(
    "t-distributed Stochastic Neighbor Embedding (t-SNE)", true);
static const DimensionReductionMethod KernelLocallyLinearEmbedding("Kernel Locally Linear Embedding (KLLE)", false);
static const DimensionReductionMethod KernelLocalTangentSpaceAlignment("Local Tangent Space Alignment (LTSA)", false);
static const DimensionReductionMethod DiffusionMap("Diffusion Map (DM)", false);
static const DimensionReductionMethod MultidimensionalScaling("Classic Multidimensional Scaling (MDS)", false);
static const DimensionReductionMethod LandmarkMultidimensionalScaling("Landmark Multidimensional Scaling (L-MDS)", false);
static const DimensionReductionMethod Isomap("Isomap", false);
static const DimensionReductionMethod LandmarkIsomap("Landmark Isomap (L-Isomap)", false);
static const DimensionReductionMethod NeighborhoodPreservingEmbedding("Neighborhood Preserving Embedding (NPE)", false);
static const DimensionReductionMethod LinearLocalTangentSpaceAlignment("Linear Local Tangent Space Alignment (LLTSA)", false);
static const DimensionReductionMethod HessianLocallyLinearEmbedding("Hessian Locally Linear Embedding (HLLE)", false);
static const DimensionReductionMethod LaplacianEigenmaps("Laplacian Eigenmaps (LE)", false);
static const DimensionReductionMethod LocalityPreservingProjections("Locality Preserving Projections (LPP)", false);
static const DimensionReductionMethod PrincipalComponentAnalysis("Principal Component Analysis (PCA)", false);
static const DimensionReductionMethod KernelPrincipalComponentAnalysis("Kernel Principal Component Analysis (KPCA)", false);
static const DimensionReductionMethod RandomProjection("Random Projection (RP)", false);
static const DimensionReductionMethod StochasticProximityEmbedding("Stochastic Proximity Embedding (SPE)", false);
static const DimensionReductionMethod PassThru("passthru", false);
static const DimensionReductionMethod FactorAnalysis("Factor Analysis (FA)", false);
static const DimensionReductionMethod ManifoldSculpting("Manifold Sculpting (MS)", false);

// Engine options for the t-SNE path (no counterpart in tapkee).
struct EngineOptions {
    tka_tsne_options opt;
    EngineOptions() { tka_default_options(&opt); }
};

// ---------------------------------------------------------------- keywords
namespace stichwort_like {
class ParametersSet;

// A named, type-erased value (include/stichwort/parameter.hpp).
class Parameter {
  public:
    Parameter() = default;
    template <class T>
    static Parameter create(const std::string& name, const T& value)
    {
        Parameter p;
        p.name_ = name;
        p.value_ = value;
        return p;
    }
    const std::string& name() const { return name_; }
    template <class T>
    T get() const
    {
        try {
            return std::any_cast<T>(value_);
        } catch (const std::bad_any_cast&) {
            throw wrong_parameter_type_error("Wrong value type for " + name_);
        }
    }
    template <class T>
    bool holds() const
    {
        return value_.type() == typeid(T);
    }
    inline ParametersSet operator,(const Parameter& p) const;

  private:
    std::string name_;
    std::any value_;
};

// Keyword set with duplicate detection and defaults merge
// (stichwort/parameter.hpp:232-311).
class ParametersSet {
  public:
    void add(const Parameter& p)
    {
        if (map_.count(p.name())) dups_.push_back(p.name());
        map_[p.name()] = p;
    }
    void check() const
    {
        if (!dups_.empty()) {
            std::stringstream ss;
            ss << "The following parameters are set more than once: ";
            for (const auto& d : dups_) ss << d << " ";
            throw multiple_parameter_error(ss.str());
        }
    }
    bool contains(const std::string& n) const { return map_.count(n) > 0; }
    void merge(const ParametersSet& o)
    {
        for (const auto& kv : o.map_)
            if (!map_.count(kv.first)) map_[kv.first] = kv.second;
    }
    const Parameter& operator[](const std::string& n) const
    {
        auto it = map_.find(n);
        if (it == map_.end()) throw missed_parameter_error(n + " is missed");
        return it->second;
    }
    ParametersSet& operator,(const Parameter& p)
    {
        add(p);
        return *this;
    }

  private:
    std::map<std::string, Parameter> map_;
    std::list<std::string> dups_;
};

inline ParametersSet Parameter::operator,(const Parameter& p) const
{
    ParametersSet s;
    s.add(*this);
    s.add(p);
    return s;
}

template <class T>
struct ParameterKeyword {
    ParameterKeyword(const std::string& n, const T& dv) : name(n), default_value(dv) {}
    Parameter operator=(const T& v) const { return Parameter::create(name, v); }
    Parameter by_default() const { return Parameter::create(name, default_value); }
    std::string name;
    T default_value;
};
}  // namespace stichwort_like

typedef stichwort_like::ParametersSet ParametersSet;
typedef stichwort_like::Parameter Parameter;

// Keyword names and defaults as include/tapkee/defines/keywords.hpp.
static const stichwort_like::ParameterKeyword<DimensionReductionMethod> method("dimension reduction method", PassThru);
static const stichwort_like::ParameterKeyword<IndexType> target_dimension("target dimension", 2);
static const stichwort_like::ParameterKeyword<ScalarType> sne_perplexity("SNE perplexity", 30.0);
static const stichwort_like::ParameterKeyword<ScalarType> sne_theta("SNE theta", 0.5);
static const stichwort_like::ParameterKeyword<IndexType> num_neighbors("number of neighbors", 5);
static const stichwort_like::ParameterKeyword<IndexType> max_iteration("maximal iteration", 100);
static const stichwort_like::ParameterKeyword<bool (*)()> cancel_function("cancel function", nullptr);
static const stichwort_like::ParameterKeyword<void (*)(double)> progress_function("progress function", nullptr);
// ---------------------------------------------------------------- logging
// The reference's logger interface (include/tapkee/utils/logging.hpp:45-144):
// a LoggerImplementation receives the messages of the enabled levels; the
// Logging singleton holds it and the levels (info, debug and benchmark off,
// warning and error on by default).  The engine's own messages (the t-SNE
// lines of tsne.hpp and the "... took <s> seconds." benchmark lines) reach
// the installed implementation through tka_set_logger.
class LoggerImplementation {
  public:
    LoggerImplementation() = default;
    virtual ~LoggerImplementation() = default;
    LoggerImplementation(const LoggerImplementation&) = delete;
    LoggerImplementation& operator=(const LoggerImplementation&) = delete;
    virtual void message_info(const std::string& msg) = 0;
    virtual void message_warning(const std::string& msg) = 0;
    virtual void message_debug(const std::string& msg) = 0;
    virtual void message_error(const std::string& msg) = 0;
    virtual void message_benchmark(const std::string& msg) = 0;
};

// Writes "[level] message" lines to std::cerr.
class DefaultLoggerImplementation : public LoggerImplementation {
  public:
    void message_info(const std::string& m) override { put("info", m); }
    void message_warning(const std::string& m) override { put("warning", m); }
    void message_debug(const std::string& m) override { put("debug", m); }
    void message_error(const std::string& m) override { put("error", m); }
    void message_benchmark(const std::string& m) override { put("benchmark", m); }

  private:
    static void put(const char* level, const std::string& m) { std::cerr << "[" << level << "] " << m << "\n"; }
};

class Logging {
  public:
    static Logging& instance()
    {
        static Logging s;
        return s;
    }
    LoggerImplementation* get_logger_impl() const { return impl_.get(); }
    // takes ownership, as the reference's setter (it deletes the previous one)
    void set_logger_impl(LoggerImplementation* i)
    {
        impl_.reset(i);
        sync();
    }
    void enable_info() { set(TKA_LOG_INFO, true); }
    void disable_info() { set(TKA_LOG_INFO, false); }
    bool is_info_enabled() const { return on(TKA_LOG_INFO); }
    void message_info(const std::string& m) { if (on(TKA_LOG_INFO) && impl_) impl_->message_info(m); }
    void enable_warning() { set(TKA_LOG_WARNING, true); }
    void disable_warning() { set(TKA_LOG_WARNING, false); }
    bool is_warning_enabled() const { return on(TKA_LOG_WARNING); }
    void message_warning(const std::string& m) { if (on(TKA_LOG_WARNING) && impl_) impl_->message_warning(m); }
    void enable_debug() { set(TKA_LOG_DEBUG, true); }
    void disable_debug() { set(TKA_LOG_DEBUG, false); }
    bool is_debug_enabled() const { return on(TKA_LOG_DEBUG); }
    void message_debug(const std::string& m) { if (on(TKA_LOG_DEBUG) && impl_) impl_->message_debug(m); }
    void enable_error() { set(TKA_LOG_ERROR, true); }
    void disable_error() { set(TKA_LOG_ERROR, false); }
    bool is_error_enabled() const { return on(TKA_LOG_ERROR); }
    void message_error(const std::string& m) { if (on(TKA_LOG_ERROR) && impl_) impl_->message_error(m); }
    void enable_benchmark() { set(TKA_LOG_BENCHMARK, true); }
    void disable_benchmark() { set(TKA_LOG_BENCHMARK, false); }
    bool is_benchmark_enabled() const { return on(TKA_LOG_BENCHMARK); }
    void message_benchmark(const std::string& m) { if (on(TKA_LOG_BENCHMARK) && impl_) impl_->message_benchmark(m); }

  private:
    Logging() : impl_(new DefaultLoggerImplementation) { sync(); }
    Logging(const Logging&) = delete;
    Logging& operator=(const Logging&) = delete;
    bool on(tka_log_level l) const { return (levels_ >> l) & 1u; }
    void set(tka_log_level l, bool v)
    {
        levels_ = v ? (levels_ | (1u << l)) : (levels_ & ~(1u << l));
        sync();
    }
    // the engine reports every enabled message here; the level was checked
    static void deliver(int32_t level, const char* msg, void* self)
    {
        LoggerImplementation* i = static_cast<Logging*>(self)->impl_.get();
        if (!i) return;
        const std::string m(msg);
        switch (level) {
            case TKA_LOG_INFO: i->message_info(m); break;
            case TKA_LOG_WARNING: i->message_warning(m); break;
            case TKA_LOG_DEBUG: i->message_debug(m); break;
            case TKA_LOG_ERROR: i->message_error(m); break;
            default: i->message_benchmark(m); break;
        }
    }
    void sync() { tka_set_logger(&Logging::deliver, this, levels_); }
    std::unique_ptr<LoggerImplementation> impl_;
    uint32_t levels_ = TKA_LOG_DEFAULT_LEVELS;
};

// Engine-only keyword: tka_tsne_options for the t-SNE path.
static const stichwort_like::ParameterKeyword<EngineOptions> engine_options("MI355X engine options", EngineOptions());

struct ProjectingFunction {
    bool valid() const { return false; }  // t-SNE has no projection (projection.hpp:66-69)
};

struct TapkeeOutput {
    TapkeeOutput() = default;
    TapkeeOutput(const DenseMatrix& e, const ProjectingFunction& p) : embedding(e), projection(p) {}
    DenseMatrix embedding;
    ProjectingFunction projection;
    tka_tsne_stats stats{};
};

namespace tapkee_internal {

inline void throw_status(tka_status s)
{
    const std::string m = tka_last_error();
    switch (s) {
        case TKA_OK: return;
        case TKA_E_WRONG_PARAMETER: throw wrong_parameter_error(m);
        case TKA_E_NO_DATA: throw no_data_error();
        case TKA_E_NOT_ENOUGH_MEMORY: throw not_enough_memory_error(m);
        case TKA_E_CANCELLED: throw cancelled_exception();
        case TKA_E_UNSUPPORTED: throw unsupported_method_error(m);
        case TKA_E_DEVICE: throw device_error(m);
        default: throw std::runtime_error(m);
    }
}

// Set to the caller's cancel function for the duration of one embedUsing.
inline thread_local bool (*g_cancel)() = nullptr;
inline int32_t cancel_trampoline() { return (g_cancel && g_cancel()) ? 1 : 0; }

class ParametersInitializedState {
  public:
    explicit ParametersInitializedState(const ParametersSet& p) : parameters(p) {}

    // matrix: features x samples, column-major, data() contiguous
    template <class Matrix>
    TapkeeOutput embedUsing(const Matrix& matrix) const
    {
        ParametersSet p = parameters;
        p.check();                                               // embed.hpp:103
        ParametersSet defaults;
        defaults.add(target_dimension.by_default());
        defaults.add(sne_perplexity.by_default());
        defaults.add(sne_theta.by_default());
        defaults.add(cancel_function.by_default());
        defaults.add(engine_options.by_default());
        const DimensionReductionMethod m = p[method.name].get<DimensionReductionMethod>();  // missed -> throws
        p.merge(defaults);                                       // embed.hpp:104
        const IndexType N = (IndexType)matrix.cols();
        const IndexType D = (IndexType)matrix.rows();
        if (N == 0) throw no_data_error();                      // methods/base.hpp:40-41
        bool (*cancel)() = p[cancel_function.name].get<bool (*)()>();
        if (cancel && cancel()) throw cancelled_exception();     // methods.hpp:30-31
        if (!m.runs_on_engine)
            throw unsupported_method_error(m.name() + " is not provided by this engine (only t-SNE runs on it)");
        const IndexType dims = p[target_dimension.name].get<IndexType>();
        const ScalarType perp = p[sne_perplexity.name].get<ScalarType>();
        const ScalarType theta = p[sne_theta.name].get<ScalarType>();
        tka_tsne_options opt = p[engine_options.name].get<EngineOptions>().opt;
        g_cancel = cancel;
        opt.cancel = cancel ? &cancel_trampoline : nullptr;
        std::vector<double> Y((size_t)N * (dims > 0 ? dims : 1));
        TapkeeOutput out;
        const tka_status s = tka_tsne_run(matrix.data(), N, D, Y.data(), dims, perp, theta, &opt, &out.stats);
        g_cancel = nullptr;
        throw_status(s);
        out.embedding = DenseMatrix(N, dims);
        for (IndexType i = 0; i < N; i++)
            for (IndexType d = 0; d < dims; d++) out.embedding(i, d) = Y[(size_t)i * dims + d];
        return out;
    }

  private:
    ParametersSet parameters;
};

}  // namespace tapkee_internal

// chain_interface.hpp:497-500
inline tapkee_internal::ParametersInitializedState with(const ParametersSet& parameters)
{
    return tapkee_internal::ParametersInitializedState(parameters);
}
inline tapkee_internal::ParametersInitializedState with(const Parameter& parameter)
{
    ParametersSet s;
    s.add(parameter);
    return tapkee_internal::ParametersInitializedState(s);
}

}  // namespace tapkee
